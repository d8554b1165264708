"""GPU parity at the configurations' own sizes, beyond the headline shapes:

* every persistent kernel variant (forced by name) at a batch of at least four times its
  largest resident grid, so each workgroup walks several claimed blocks with the next
  block's lines prefetched (the matrix-core forms: several blocks at a fixed stride with
  their matrices built once), against numpy.fft on the whole batch;
* the uneven complex64 mix in ONE persistent launch (6^3 x 65536, 8^3 x 32768,
  12^3 x 8192, 17^3 x 2048) as the FFT and as both matrix-core splits: 256 sampled volumes
  per item against numpy.fft in float64, Parseval and linearity over the full batch;
* a stream of 2,000 requests of the serving mix (batches log-uniform in [1, 512], both
  directions, complex128 and complex64) through the continuously batched server under a
  64 GB budget, outputs in caller buffers: every request checked on sampled volumes, and a
  sample of requests bit-identical when served alone.

Tolerances (relative L2): complex128 1e-13; complex64 2e-6; matrix-core bf16x3 2e-5,
fp16x3 6e-7."""
import ctypes
import math
import re

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

CUS = 256
TOL = {0: 1e-13, 1: 2e-6}
MFMA_TOL = {"bf16x3": 2e-5, "fp16x3": 6e-7}   # the matrix-core variants' stated accuracy


def _tol(name, prec):
    for k, v in MFMA_TOL.items():
        if name.startswith("mfma_") and name.endswith(k):
            return v
    return TOL[prec]


def _persistent_groups():
    """(shape, prec, direction) -> [(name, batch needed)] for every persistent variant."""
    import __graft_entry__
    f = __graft_entry__.load_package().fft3d
    groups = {}
    for name, shape, prec, d in f.kernels():
        m = re.search(r"_v(\d+)t\d+(?:r\d+)?p(\d+)", name)
        if m and int(m.group(2)) > 0:
            vpw, P = int(m.group(1)), int(m.group(2))
            need = 4 * P * CUS * vpw
        else:
            m = re.search(r"_s(\d+)t\d+b\d+c\d+p(\d+)", name)
            if not (m and int(m.group(2)) > 0):
                continue
            slabs, P = int(m.group(1)), int(m.group(2))
            need = math.ceil(4 * P * CUS / math.ceil(shape[0] / slabs))
        groups.setdefault((shape, prec, d), []).append((name, need))
    return groups


GROUPS = _persistent_groups()


@pytest.mark.parametrize("key", sorted(GROUPS), ids=lambda k: "%dx%dx%d_%s_%s" % (
    *k[0], "c128" if k[1] == 0 else "c64", "fwd" if k[2] < 0 else "bwd"))
def test_persistent_variants_walk_many_blocks(gpu, key):
    f = gpu.fft3d
    shape, prec, d = key
    batch = max(n for _, n in GROUPS[key])
    dt = np.complex128 if prec == 0 else np.complex64
    x = oracle.make_input(shape, batch, seed=31 + batch, dtype=dt)
    want = oracle.fftn_numpy(x, sign=d)
    for name, _ in GROUPS[key]:
        print(f"{name} B={batch} ...", end=" ", flush=True)
        plan = f.Plan(shape, batch, d, prec, kernel=name)
        y = plan.run(x)
        plan.close()
        err = oracle.rel_l2(y, want)
        print(f"rel_l2={err:.2e}", flush=True)
        assert np.all(np.isfinite(y)), name
        assert err <= _tol(name, prec), (name, err)


# --------------------------------------------------------------------------- uneven mix
UNEVEN_MIX = [((6, 6, 6), 65536), ((8, 8, 8), 32768), ((12, 12, 12), 8192), ((17, 17, 17), 2048)]
MIX_TOL = {0: 2e-6, 4: 2e-5, 8: 6e-7}   # flags -> rel L2 bar


def _c64(seed, shape, batch):
    return oracle.make_input(shape, batch, seed=seed, dtype=np.complex64)


def _energy(a):
    a = a.reshape(a.shape[0], -1)
    return np.sum(a.real.astype(np.float64) ** 2 + a.imag.astype(np.float64) ** 2)


@pytest.mark.parametrize("flags", [0, 4, 8], ids=["fft", "mfma_bf16x3", "mfma_fp16x3"])
def test_uneven_mix_one_launch_full_size(gpu, flags):
    f = gpu.fft3d
    tol = MIX_TOL[flags]
    x1 = [_c64(100 + k, s, b) for k, (s, b) in enumerate(UNEVEN_MIX)]
    x2 = [_c64(200 + k, s, b) for k, (s, b) in enumerate(UNEVEN_MIX)]
    a, c = 0.75 - 0.5j, -1.25 + 0.25j
    x3 = [(a * p + c * q).astype(np.complex64) for p, q in zip(x1, x2)]
    m = f.MultiPlan([(s, b, -1) for s, b in UNEVEN_MIX], precision=f.C64, flags=flags)
    name, ntasks, grid = m.info()
    print(f"{name}: {ntasks} tasks on {grid} workgroups", flush=True)
    y1, y2, y3 = m.run(x1), m.run(x2), m.run(x3)
    m.close()
    rng = np.random.default_rng(5)
    for k, (s, b) in enumerate(UNEVEN_MIX):
        V = s[0] * s[1] * s[2]
        idx = np.sort(rng.choice(b, 256, replace=False))
        ref = oracle.fftn_numpy(x1[k][idx].astype(np.complex128))
        err = oracle.rel_l2(y1[k][idx], ref)
        pars = abs(_energy(y1[k]) - V * _energy(x1[k])) / (V * _energy(x1[k]))
        lin = oracle.rel_l2(y3[k], a * y1[k].astype(np.complex128) + c * y2[k].astype(np.complex128))
        print(f"  {s} x {b}: rel_l2 {err:.2e}  parseval {pars:.2e}  linearity {lin:.2e}",
              flush=True)
        assert np.all(np.isfinite(y1[k])) and np.all(np.isfinite(y3[k]))
        assert err <= tol, (s, err)
        assert pars <= 2 * tol, (s, pars)
        assert lin <= 3 * tol, (s, lin)


# ------------------------------------------------------------------------ serving stream
def _download_volumes(buf_addr, dtype, V, first, count):
    import __graft_entry__
    h = __graft_entry__.load_package().hip
    out = np.empty((count, V), dtype=dtype)
    h.check(h.lib().hipMemcpy(out.ctypes.data, ctypes.c_void_p(buf_addr + first * V * out.itemsize),
                              out.nbytes, h.D2H), "hipMemcpy D2H")
    return out


def test_serving_stream_2000_requests_full_batches(gpu):
    f, h, sv = gpu.fft3d, gpu.hip, gpu.serving
    n = 2000
    mix = sv.draw_mix(n, seed=11)
    assert max(b for _, b, _, _ in mix) > 256
    pools = sv.InputPools(seed=4, keep_host=True)
    rng = np.random.default_rng(12)
    firsts = [int(rng.integers(0, sv.MAX_BATCH + 1)) for _ in mix]
    sizes = [b * s[0] * s[1] * s[2] * (16 if p == f.C128 else 8) for s, b, _, p in mix]
    offs = np.concatenate([[0], np.cumsum([(z + 255) // 256 * 256 for z in sizes])])
    arena = h.DeviceBuffer(int(offs[-1]))
    reqs = [f.SrvRequest(*s, b, d, p, pools.addr(s, p, first), arena.addr + int(o))
            for (s, b, d, p), first, o in zip(mix, firsts, offs[:-1])]
    srv = f.Server(budget_bytes=64 << 30, streams=16, batched=True, window_us=30.0)
    arrivals = sv.poisson_arrivals(n, 60000.0, seed=13)
    lat = srv.run_trace(reqs, arrivals)
    st = srv.stats()
    print(st, f"p50 {np.percentile(lat, 50) * 1e3:.3f} ms p99 {np.percentile(lat, 99) * 1e3:.3f} ms",
          flush=True)
    assert st["completed"] == n and np.all(lat > 0)
    assert st["mixed_launches"] > 0 and st["peak_inflight_bytes"] <= 64 << 30
    worst = {0: 0.0, 1: 0.0}
    samples = {}
    for k, ((s, b, d, p), first, o) in enumerate(zip(mix, firsts, offs[:-1])):
        V = s[0] * s[1] * s[2]
        dt = np.complex128 if p == f.C128 else np.complex64
        pick = sorted({0, b // 3, (2 * b) // 3, b - 1})
        host = pools.host(s, p)
        for i in pick:
            y = _download_volumes(arena.addr + int(o), dt, V, i, 1).reshape((1,) + s)
            x = host[first + i].reshape((1,) + s).astype(np.complex128)
            err = oracle.rel_l2(y, oracle.fftn_numpy(x, sign=d))
            worst[p] = max(worst[p], err)
            assert err <= TOL[p], (k, s, b, d, p, i, err)
        if k % 20 == 0:
            samples[k] = _download_volumes(arena.addr + int(o), dt, V, 0, b)
    print(f"worst rel_l2: c128 {worst[0]:.2e}  c64 {worst[1]:.2e}", flush=True)
    # the same requests alone, one at a time, into a scratch output
    for k, y_mix in samples.items():
        s, b, d, p = mix[k]
        scratch = h.DeviceBuffer(sizes[k])
        rid = srv.submit(s, b, d, p, pools.addr(s, p, firsts[k]), scratch.addr)
        srv.wait(rid)
        alone = _download_volumes(scratch.addr, y_mix.dtype, y_mix.shape[1], 0, b)
        scratch.close()
        assert np.array_equal(alone.view(np.uint8), y_mix.view(np.uint8)), (k, s, b)
    srv.close()
    arena.close()
    pools.close()


# --------------------------------------------------------------------------- dense floor
@pytest.mark.parametrize("shape,batch", [((17, 17, 17), 64), ((36, 36, 36), 8), ((24, 24, 48), 6),
                                         ((8, 8, 8), 512)])
def test_dense_floor_agrees_with_the_fast_kernels(gpu, shape, batch):
    """The dense row-column floor (no factorization) and the plan's fast kernel, both on the
    device, against each other and the oracle."""
    f = gpu.fft3d
    x = oracle.make_input(shape, batch, seed=3 + batch, dist="normal_half")
    name = "dense_%dx%dx%d_c128_fwd_l1" % shape
    dense = f.Plan(shape, batch, f.FORWARD, f.C128, kernel=name)
    fast = f.Plan(shape, batch, f.FORWARD, f.C128)
    yd, yf = dense.run(x), fast.run(x)
    print(f"{name} vs {fast.kernel}", flush=True)
    dense.close()
    fast.close()
    want = oracle.fftn_numpy(x)
    assert oracle.rel_l2(yd, want) <= 1e-13
    assert oracle.rel_l2(yd, yf) <= 1e-13


# ------------------------------------------------------------------------ batch classes
@pytest.mark.parametrize("L,prec", [(24, 0), (24, 1), (36, 1)], ids=["24_c128", "24_c64", "36_c64"])
def test_plan_batch_classes(gpu, L, prec):
    """A plan raced at batch 512 and at 64 (the small-batch class), at geometries where the
    split-resident single pass and the pencil trade places: executes on both sides of the
    class boundary, each against numpy.fft."""
    f = gpu.fft3d
    dt = np.complex128 if prec == 0 else np.complex64
    plan = f.Plan((L, L, L), 512, f.FORWARD, prec)
    small = plan.kernel_small
    print(f"{plan.kernel} small class {small}", flush=True)
    if small is not None:
        assert small[1] == int(math.sqrt(512 * 64))
    for b in (1, 5, 64, 181, 182, 300, 512):
        x = oracle.make_input((L, L, L), b, seed=70 + b, dtype=dt)
        err = oracle.rel_l2(plan.run(x), oracle.fftn_numpy(x))
        print(f"  batch {b}: rel_l2 {err:.2e}", flush=True)
        assert err <= TOL[prec], (b, err)
    plan.close()
