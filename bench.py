"""Benchmark: batched transforms/s and HBM roofline fraction of the MI355X 3D FFT.

One step = one forward execute of the headline workload (17x17x17 complex128, batch 1024,
inputs iid uniform(-1,1)) on device-resident buffers.  Each step uses the next of R
(in, out) buffer pairs, R chosen so the pairs total >= 2 GiB: every working set here fits
the MI355X's 256 MB Infinity Cache, so same-buffer repetition would time the cache, not
HBM.  The timed region is K steps between HIP events on the launch stream, bracketed by
device synchronisations.  The other workloads (8^3 x 4096, 36^3 x 128, 24x24x48 x 64 and
the 17^3 inverse) are timed the same way and reported under "secondary".

Prints ONE JSON line.  `python bench.py --gpus N` with N > 1 exits: this build runs on one
GPU (batches shard by volume with no communication).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
# HBM bytes per launch of the headline kernel from rocprofv3 FETCH_SIZE + WRITE_SIZE
# (profiles/), or None until measured.
TRAFFIC = None
ROTATE_BYTES = 2 << 30

HEADLINE = dict(name="L17_c128_fwd_b1024", shape=(17, 17, 17), batch=1024, dist="uniform",
                direction=-1)
SECONDARY = [
    dict(name="L8_c128_fwd_b4096", shape=(8, 8, 8), batch=4096, dist="uniform", direction=-1),
    dict(name="L17_c128_bwd_b1024", shape=(17, 17, 17), batch=1024, dist="uniform",
         direction=+1),
    dict(name="L36_c128_fwd_b128", shape=(36, 36, 36), batch=128, dist="normal_half",
         direction=-1),
    dict(name="24x24x48_c128_fwd_b64", shape=(24, 24, 48), batch=64, dist="normal_half",
         direction=-1),
]


def make_input(shape, batch, seed, dist):
    rng = np.random.default_rng(seed)
    full = (batch,) + tuple(shape)
    if dist == "uniform":
        re, im = rng.uniform(-1, 1, full), rng.uniform(-1, 1, full)
    else:
        s = math.sqrt(0.5)
        re, im = rng.normal(0, s, full), rng.normal(0, s, full)
    return (re + 1j * im).astype(np.complex128)


def measure(pkg, wl, steps, warmup, check=True):
    f, h = pkg.fft3d, pkg.hip
    shape, batch = wl["shape"], wl["batch"]
    V = shape[0] * shape[1] * shape[2]
    nbytes = V * batch * 16
    rot = max(1, math.ceil(ROTATE_BYTES / (2 * nbytes)))
    x = make_input(shape, batch, seed=2024 + batch, dist=wl["dist"])
    ins = [h.DeviceBuffer(nbytes) for _ in range(rot)]
    outs = [h.DeviceBuffer(nbytes) for _ in range(rot)]
    ins[0].upload(x)
    for b in ins[1:]:
        h.check(h.lib().hipMemcpy(b.ptr, ins[0].ptr, nbytes, h.D2D), "hipMemcpy D2D")
    plan = f.Plan(shape, batch, wl["direction"], f.C128)
    stream = h.Stream()
    e0, e1 = h.Event(), h.Event()
    for i in range(warmup):
        plan.execute(ins[i % rot].addr, outs[i % rot].addr, stream=stream)
    h.synchronize()
    e0.record(stream)
    for i in range(steps):
        plan.execute(ins[i % rot].addr, outs[i % rot].addr, stream=stream)
    e1.record(stream)
    e1.synchronize()
    h.synchronize()
    ms = e0.elapsed_ms(e1)
    # Per-launch durations (events bracketing each execute on its stream) for the roofline.
    pairs = [(h.Event(), h.Event()) for _ in range(min(steps, 50))]
    for i, (a, b) in enumerate(pairs):
        a.record(stream)
        plan.execute(ins[i % rot].addr, outs[i % rot].addr, stream=stream)
        b.record(stream)
    pairs[-1][1].synchronize()
    per = [a.elapsed_ms(b) for a, b in pairs]
    for a, b in pairs:
        a.close()
        b.close()
    launch_ms = sum(per) / len(per)
    res = dict(kernel=plan.kernel, rotate_pairs=rot, steps=steps, ms_total=ms,
               ms_per_execute=ms / steps, transforms_per_s=batch * steps / (ms * 1e-3),
               bracketed_us_mean=launch_ms * 1e3, bracketed_us_min=min(per) * 1e3)
    # Kernel duration for the roofline: the back-to-back timed region per execute (one
    # kernel per execute for the fused paths; the whole two-kernel sequence for pencils).
    # Events bracketing single launches add several microseconds of event overhead.
    algo = 2.0 * nbytes
    res["algorithmic_bytes_per_execute"] = algo
    res["achieved_GBs"] = algo / (ms * 1e-3 / steps) / 1e9
    res["roofline_frac"] = res["achieved_GBs"] / PEAK_HBM_GBS
    if check:
        import oracle
        # the last step's output pair: ((steps-1) % rot); its input is x
        y = outs[(steps - 1) % rot].download(np.complex128, x.shape)
        res["rel_l2_vs_numpy"] = oracle.rel_l2(y, oracle.fftn_numpy(x, sign=wl["direction"]))
    for b in ins + outs:
        b.close()
    stream.close()
    plan.close()
    return res


def _ref_exe(kind):
    isa = "v4" if "avx512f" in open("/proc/cpuinfo").read() else "v3"
    p = os.path.join(ROOT, "oracle", "_ref", isa, f"ref_bench_{kind}")
    return p if os.path.exists(p) else None


def cpu_baseline(seconds=10.0):
    """The reference's own CPU kernels on this host, on a bounded sample of the headline
    workload (17^3 complex128): all cores through its OpenMP kernel, plus one pinned core
    through its best-exemplar library.  Falls back to the oracle's numpy.fft ("port")."""
    import subprocess
    mt, st = _ref_exe("mt"), _ref_exe("st")
    if mt and st:
        ncores = len(os.sched_getaffinity(0))
        env = dict(os.environ, OMP_NUM_THREADS=str(ncores), OMP_PROC_BIND="close",
                   OMP_PLACES="cores")
        r = subprocess.run([mt, "--L", "17", "--batch", "1024", "--seconds", str(seconds)],
                           capture_output=True, text=True, env=env, timeout=300)
        r1 = subprocess.run(["taskset", "-c", "0", st, "--L", "17", "--batch", "256",
                             "--seconds", str(seconds / 2)], capture_output=True, text=True,
                            timeout=300)
        if r.returncode == 0 and r1.returncode == 0:
            m, s1 = json.loads(r.stdout.strip().splitlines()[-1]), json.loads(r1.stdout.strip().splitlines()[-1])
            model = ""
            try:
                model = [l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo")
                         if l.startswith("model name")][0]
            except Exception:
                pass
            return dict(value=m["transforms_per_s"], unit="transforms/s", cores=m["threads"],
                        kind="reference",
                        sample=f"reference {m['kernel']} (bench/mt/impl_4, OpenMP, "
                               f"{m['threads']} threads) on 17^3 complex128 x 1024, "
                               f"{m['reps']} executes in {m['seconds']:.1f} s on {model}",
                        single_core=dict(value=s1["transforms_per_s"], cores=1,
                                         kernel=s1["kernel"],
                                         sample=f"reference best-exemplar library "
                                                f"(bench/geom/best, picked {s1['kernel']}) "
                                                f"on 17^3 x 256, taskset -c 0, "
                                                f"{s1['seconds']:.1f} s"))
    import oracle
    x = make_input((17, 17, 17), 64, seed=7, dist="uniform")
    n, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        oracle.fftn_numpy(x)
        n += x.shape[0]
    dt = time.perf_counter() - t0
    return dict(value=n / dt, unit="transforms/s", cores=1, kind="port",
                sample=f"numpy.fft.fftn (pocketfft) on 64 x 17^3 complex128 volumes, "
                       f"repeated for {dt:.1f} s ({n} transforms), 1 thread")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-secondary", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU only; batches of independent volumes "
              "shard across GPUs with no communication (see DESIGN.md).", file=sys.stderr)
        sys.exit(2)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    # The CPU baseline spawns the reference binaries, so it runs before HIP is initialised.
    cpu = None if a.no_cpu else cpu_baseline()
    if cpu:
        print(f"# cpu_baseline: {cpu}", file=sys.stderr, flush=True)
    import __graft_entry__
    pkg = __graft_entry__.load_package()
    if pkg.hip.device_count() < 1:
        print("bench.py: no HIP device", file=sys.stderr)
        sys.exit(1)
    pkg.hip.check(pkg.hip.lib().hipSetDevice(0), "hipSetDevice")

    head = measure(pkg, HEADLINE, a.steps, a.warmup)
    print(f"# headline {HEADLINE['name']}: {head}", file=sys.stderr, flush=True)
    secondary = {}
    if not a.no_secondary:
        for wl in SECONDARY:
            secondary[wl["name"]] = measure(pkg, wl, max(10, a.steps // 2), a.warmup)
            print(f"# {wl['name']}: {secondary[wl['name']]}", file=sys.stderr, flush=True)
    V = 17 ** 3
    line = {
        "metric": "batched transforms/s + % HBM roofline at L=8/17/36 fp64; serving p99 latency",
        "value": head["transforms_per_s"],
        "unit": "transforms/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": head["ms_per_execute"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: iid uniform(-1,1) real and imaginary parts, seeded",
        "config": {"workload": "17x17x17 complex128 forward 3D FFT, batch 1024, "
                               "device-resident, rotated buffers",
                   "n0": 17, "n1": 17, "n2": 17, "batch": 1024, "precision": "complex128",
                   "direction": "forward", "kernel": head["kernel"],
                   "rotate_pairs": head["rotate_pairs"]},
        "gflops_nominal": 5.0 * V * math.log2(V) * head["transforms_per_s"] / 1e9,
        "rel_l2_vs_numpy": head["rel_l2_vs_numpy"],
        "roofline": {"bound": "hbm", "achieved": head["achieved_GBs"], "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": head["roofline_frac"], "traffic": TRAFFIC,
                     "kernel_us_mean": head["ms_per_execute"] * 1e3,
                     "algorithmic_bytes_per_launch": head["algorithmic_bytes_per_execute"]},
        "cpu_baseline": cpu,
        "secondary": secondary,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
