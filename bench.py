"""Benchmark: batched transforms/s and HBM roofline fraction of the MI355X 3D FFT.

One step = one forward execute of the headline workload (17x17x17 complex128, batch 1024,
inputs iid uniform(-1,1)) on device-resident buffers.  Each step uses the next of R
(in, out) buffer pairs, R chosen so the pairs total >= 2 GiB: every working set here fits
the MI355X's 256 MB Infinity Cache, so same-buffer repetition would time the cache, not
HBM.  The timed region is K steps between HIP events on the launch stream, bracketed by
device synchronisations.  The other workloads (8^3 x 4096, 36^3 x 128, 24x24x48 x 64 and
the 17^3 inverse) are timed the same way and reported under "secondary".

"precision_variants": complex64 on the uneven batch set 6^3 x 65536, 8^3 x 32768,
12^3 x 8192, 17^3 x 2048 -- as one persistent mixed launch (FFT, and the DFT-as-matrix
matrix-core form), as per-geometry FFT plans, and as per-geometry matrix-core plans (split
bf16 / fp16, three products, FP32 accumulate) -- with relative L2 and max error on 256
volumes per geometry against
numpy.fft in float64, and for the matrix-core kernels the executed MFMA TFLOP/s against the
2.5 PFLOP/s dense peak.

"serving": the request server under Poisson arrivals (20,000 requests, the mixed stream of
opt-fft_amd/serving.py, 64 GB budget, 16 launches in flight, 30 us batching window for the
continuously batched server): the highest arrival rate whose p99 latency stays
within 2 ms, for continuous batching and for the plain one-launch-per-request path, and
the throughput ratio of the two at that cap.

Prints ONE JSON line.  `python bench.py --gpus N` with N > 1 exits: the benchmark runs on one
GPU (batches of independent volumes shard by volume with no communication).
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
# HBM bytes per launch of the headline kernel (pipe_17x17x17_c128_fwd_v2t640p1_fold_st3, the
# race's pick) from rocprofv3 FETCH_SIZE + WRITE_SIZE, each scaled by the same counter on the
# runtime's copy of a known size: 39,995.5 KiB x 2 + 78,796.9 KiB
# (profiles/k_fpipe_17x17x17_c128_fwd_b1024.md).
TRAFFIC = int((39995.5 * 2 + 78796.859375) * 1024)
TRAFFIC_KERNEL = "pipe_17x17x17_c128_fwd_v2t640p1_fold_st3"
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


# Accuracy of a measured output, against numpy.fft in float64 (the reference's boundary check,
# bench/geom/check.py:28-33); the unnormalised backward transform is V * ifftn.
def _np_dft(x, sign=-1):
    x = np.asarray(x, dtype=np.complex128)
    if sign < 0:
        return np.fft.fftn(x, axes=(-3, -2, -1))
    return np.fft.ifftn(x, axes=(-3, -2, -1)) * (x.shape[-3] * x.shape[-2] * x.shape[-1])


def _rel_l2(y, ref):
    return float(np.linalg.norm((np.asarray(y, np.complex128) - ref).ravel()) /
                 np.linalg.norm(ref.ravel()))


def _rel_max(y, ref):
    return float(np.max(np.abs(np.asarray(y, np.complex128) - ref)) / np.max(np.abs(ref)))


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
        # the last step's output pair: ((steps-1) % rot); its input is x
        y = outs[(steps - 1) % rot].download(np.complex128, x.shape)
        res["rel_l2_vs_numpy"] = _rel_l2(y, _np_dft(x, wl["direction"]))
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
    through its best-exemplar library; the same two for 8^3 and 36^3 on 3 s samples
    ("other_geometries").  Falls back to the oracle's numpy.fft ("port")."""
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
        # The other configurations' geometries, shorter samples: all cores and one core.
        others = {}
        for L, b_mt, b_st in ((8, 4096, 2048), (36, 128, 32)):
            rm = subprocess.run([mt, "--L", str(L), "--batch", str(b_mt), "--seconds", "3"],
                                capture_output=True, text=True, env=env, timeout=300)
            rs = subprocess.run(["taskset", "-c", "0", st, "--L", str(L), "--batch", str(b_st),
                                 "--seconds", "3"], capture_output=True, text=True, timeout=300)
            if rm.returncode == 0 and rs.returncode == 0:
                a_, b_ = (json.loads(x.stdout.strip().splitlines()[-1]) for x in (rm, rs))
                others[f"L{L}"] = dict(
                    all_cores=dict(value=a_["transforms_per_s"], cores=a_["threads"],
                                   kernel=a_["kernel"], batch=b_mt, seconds=a_["seconds"]),
                    single_core=dict(value=b_["transforms_per_s"], cores=1,
                                     kernel=b_["kernel"], batch=b_st, seconds=b_["seconds"]))
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
                                                f"{s1['seconds']:.1f} s"),
                        other_geometries=others)
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


# Precision variants on uneven batches (complex64): one persistent mixed launch, the
# per-geometry complex64 FFT plans, and the DFT-as-matrix variant on the matrix cores.
C4_ITEMS = [((6, 6, 6), 65536), ((8, 8, 8), 32768), ((12, 12, 12), 8192), ((17, 17, 17), 2048)]
C4_CHECK = 256          # volumes per item compared with numpy.fft in float64
PEAK_MFMA_TFLOPS = 2500.0   # dense bf16/fp16 matrix peak of one MI355X


def _uniform_c64(rng, shape, batch):
    full = (batch,) + tuple(shape)
    return (rng.uniform(-1, 1, full) + 1j * rng.uniform(-1, 1, full)).astype(np.complex64)


# Volumes per block of the matrix-core kernels (standalone and mixed launch alike).
MFMA_VPW = {(6, 6, 6): 16, (8, 8, 8): 8, (12, 12, 12): 2, (17, 17, 17): 1, (8, 8, 16): 4}


def mfma_flops(shape, nvol):
    """MFMA flops the matrix-core kernel executes for nvol volumes: per block of VPW volumes
    and per axis pass of length N, ceil(lines/16) tiles of T x T x 3 16x16x16 products
    (T = ceil(2N/16), three products per tile for the compensated split)."""
    vpw = MFMA_VPW[tuple(shape)]
    V = shape[0] * shape[1] * shape[2]
    per_block = 0
    for n in shape:
        T = (2 * n + 15) // 16
        per_block += -(-(vpw * V // n) // 16) * T * T * 3
    full, part = divmod(nvol, vpw)
    insts = full * per_block
    if part:
        for n in shape:
            T = (2 * n + 15) // 16
            insts += -(-(part * V // n) // 16) * T * T * 3
    return insts * 2.0 * 16 * 16 * 16


def precision_mix(pkg, steps=20):
    f, h = pkg.fft3d, pkg.hip
    rng = np.random.default_rng(44)
    xs = [_uniform_c64(rng, sh, b) for sh, b in C4_ITEMS]
    nbytes = sum(x.nbytes for x in xs)
    sets = 2
    ins = [[h.DeviceBuffer(x.nbytes) for x in xs] for _ in range(sets)]
    outs = [[h.DeviceBuffer(x.nbytes) for x in xs] for _ in range(sets)]
    for s_ in range(sets):
        for b, x in zip(ins[s_], xs):
            b.upload(x)
    stream = h.Stream()
    e0, e1 = h.Event(), h.Event()

    def timed(launch):
        for i in range(3):
            launch(i % sets)
        h.synchronize()
        e0.record(stream)
        for i in range(steps):
            launch(i % sets)
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_ms(e1) / steps * 1e-3

    def errors(bufs):
        out = {}
        for (sh, b), x, ob in zip(C4_ITEMS, xs, bufs):
            y = ob.download(np.complex64, x.shape)[:C4_CHECK]
            ref = _np_dft(x[:C4_CHECK])
            out[f"L{sh[0]}"] = dict(rel_l2=_rel_l2(y, ref), rel_max=_rel_max(y, ref))
        return out

    res = {"items": [dict(shape=list(sh), batch=b) for sh, b in C4_ITEMS],
           "bytes_per_launch": 2 * nbytes, "error_sample_volumes": C4_CHECK}
    mp = f.MultiPlan([(sh, b, -1) for sh, b in C4_ITEMS], f.C64)
    t = timed(lambda s_: mp.execute([b.addr for b in ins[s_]], [b.addr for b in outs[s_]],
                                    stream=stream))
    h.synchronize()
    res["one_launch_fft_c64"] = dict(kernel=mp.info()[0], us=t * 1e6, GBs=2 * nbytes / t / 1e9,
                                     errors=errors(outs[0]))
    mp.close()
    for tag, flag in (("one_launch_mfma_bf16x3", f.MFMA_BF16), ("one_launch_mfma_fp16x3", f.MFMA_FP16)):
        mp = f.MultiPlan([(sh, b, -1) for sh, b in C4_ITEMS], f.C64, flags=flag)
        t = timed(lambda s_: mp.execute([b.addr for b in ins[s_]], [b.addr for b in outs[s_]],
                                        stream=stream))
        h.synchronize()
        fl = sum(mfma_flops(sh, b) for sh, b in C4_ITEMS)
        res[tag] = dict(kernel=mp.info()[0], us=t * 1e6, GBs=2 * nbytes / t / 1e9,
                        mfma_tflops=fl / t / 1e12, mfma_util=fl / t / 1e12 / PEAK_MFMA_TFLOPS,
                        errors=errors(outs[0]))
        mp.close()
    for tag, flag in (("fft_c64", 0), ("mfma_bf16x3", f.MFMA_BF16), ("mfma_fp16x3", f.MFMA_FP16)):
        per = {}
        for k, ((sh, b), x) in enumerate(zip(C4_ITEMS, xs)):
            p = f.Plan(sh, b, f.FORWARD, f.C64, flags=flag)
            tk = timed(lambda s_: p.execute(ins[s_][k].addr, outs[s_][k].addr, stream=stream))
            h.synchronize()
            d = dict(kernel=p.kernel, us=tk * 1e6, GBs=2 * x.nbytes / tk / 1e9)
            y = outs[(steps - 1) % sets][k].download(np.complex64, x.shape)[:C4_CHECK]
            ref = _np_dft(x[:C4_CHECK])
            d["rel_l2"] = _rel_l2(y, ref)
            d["rel_max"] = _rel_max(y, ref)
            if flag:
                fl = mfma_flops(sh, b)
                d["mfma_tflops"] = fl / tk / 1e12
                d["mfma_util"] = d["mfma_tflops"] / PEAK_MFMA_TFLOPS
            per[f"L{sh[0]}"] = d
            p.close()
        res[tag] = per
    for lst in ins + outs:
        for b in lst:
            b.close()
    return res


def serving(pkg, n=20000, p99_cap_ms=2.0, streams=16, window_us=30.0):
    """Load test of the request server: the highest Poisson arrival rate whose p99 latency
    stays within p99_cap_ms, for continuous batching and for the plain path."""
    f, sv = pkg.fft3d, pkg.serving
    mix = sv.draw_mix(n, seed=0)
    pools = sv.InputPools(seed=0)
    reqs = sv.make_requests(mix, pools, seed=1)
    req_bytes = sum(2 * r.batch * r.n0 * r.n1 * r.n2 * (16 if r.precision == 0 else 8)
                    for r in reqs) / n
    out = {"requests": n, "mean_io_bytes_per_request": req_bytes, "p99_cap_ms": p99_cap_ms,
           "budget_bytes": 64 << 30, "streams": streams, "window_us": window_us}
    for mode in ("plain", "batched"):
        srv = f.Server(budget_bytes=64 << 30, streams=streams, batched=(mode == "batched"),
                       window_us=window_us if mode == "batched" else 0.0)
        ok, bad, pts = None, None, []
        rate = 4000.0
        while rate < 2e6:
            pt = sv.load_point(srv, reqs, rate)
            pts.append(pt)
            if pt["p99_ms"] <= p99_cap_ms:
                ok = pt
                rate *= 1.5
            else:
                bad = pt
                break
        if ok and bad:
            lo, hi = ok["rate"], bad["rate"]
            for _ in range(4):
                mid = 0.5 * (lo + hi)
                pt = sv.load_point(srv, reqs, mid)
                pts.append(pt)
                if pt["p99_ms"] <= p99_cap_ms:
                    ok, lo = pt, mid
                else:
                    hi = mid
        st = srv.stats()
        srv.close()
        out[mode] = dict(max_rate_req_s=ok["rate"] if ok else 0.0,
                         at_max=ok, stats=st,
                         points=[{k: round(v, 4) for k, v in p.items()} for p in pts])
        print(f"# serving {mode}: max rate {out[mode]['max_rate_req_s']:.0f}/s", file=sys.stderr,
              flush=True)
    pools.close()
    pl, ba = out["plain"]["at_max"], out["batched"]["at_max"]
    if pl and ba:
        out["throughput_ratio_at_p99_cap"] = ba["transforms_per_s"] / pl["transforms_per_s"]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-secondary", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-precision", action="store_true")
    ap.add_argument("--workload", default=None,
                    help="time only this secondary workload (e.g. L8_c128_fwd_b4096) and print "
                         "its measurement; for profiling runs, no JSON line")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: measures one GPU only; batches of independent volumes "
              "shard across GPUs with no communication.", file=sys.stderr)
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

    if a.workload:
        wl = [w for w in SECONDARY if w["name"] == a.workload]
        if not wl:
            print(f"bench.py: unknown workload {a.workload}", file=sys.stderr)
            sys.exit(2)
        print(f"# {a.workload}: {measure(pkg, wl[0], a.steps, a.warmup)}", flush=True)
        return
    head = measure(pkg, HEADLINE, a.steps, a.warmup)
    print(f"# headline {HEADLINE['name']}: {head}", file=sys.stderr, flush=True)
    secondary = {}
    if not a.no_secondary:
        for wl in SECONDARY:
            secondary[wl["name"]] = measure(pkg, wl, max(10, a.steps // 2), a.warmup)
            print(f"# {wl['name']}: {secondary[wl['name']]}", file=sys.stderr, flush=True)
    precision = None if a.no_precision else precision_mix(pkg)
    if precision:
        print(f"# precision variants: {precision}", file=sys.stderr, flush=True)
    serve = None if a.no_serving else serving(pkg)
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
                     "unit": "GB/s", "frac": head["roofline_frac"],
                     "traffic": TRAFFIC if head["kernel"] == TRAFFIC_KERNEL else None,
                     "kernel_us_mean": head["ms_per_execute"] * 1e3,
                     "algorithmic_bytes_per_launch": head["algorithmic_bytes_per_execute"]},
        "cpu_baseline": cpu,
        "secondary": secondary,
        "precision_variants": precision,
        "serving": serve,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
