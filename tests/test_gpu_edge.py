"""GPU edge cases: batches far past the workload sizes (checked through the round trip and
spot volumes against the oracle), one plan executed concurrently on several streams (each
launch of a persistent kernel takes its own claim counter), partial batches that must leave
the rest of the output untouched, and plan creation that does not leak device memory."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _uniform(shape, batch, seed, dtype=np.complex128):
    rng = np.random.default_rng(seed)
    full = (batch,) + tuple(shape)
    return (rng.uniform(-1, 1, full) + 1j * rng.uniform(-1, 1, full)).astype(dtype)


@pytest.mark.parametrize("shape,batch", [((8, 8, 8), 262144), ((17, 17, 17), 16384)])
def test_large_batch_round_trip_and_spot_volumes(gpu, shape, batch):
    """2.1-1.3 GB per buffer: forward then normalised backward returns the input (1e-13),
    and 16 volumes spread over the batch match the oracle's forward transform."""
    f, h = gpu.fft3d, gpu.hip
    x = _uniform(shape, batch, seed=batch)
    bx, by, bz = (h.DeviceBuffer(x.nbytes) for _ in range(3))
    bx.upload(x)
    fwd = f.Plan(shape, batch, f.FORWARD, f.C128)
    bwd = f.Plan(shape, batch, f.BACKWARD, f.C128, flags=f.NORMALIZE)
    print(f"{shape} B={batch}: {fwd.kernel} / {bwd.kernel}", flush=True)
    fwd.execute(bx.addr, by.addr)
    bwd.execute(by.addr, bz.addr)
    h.synchronize()
    y = by.download(np.complex128, x.shape)
    pick = np.linspace(0, batch - 1, 16).astype(int)
    assert oracle.rel_l2(y[pick], oracle.fftn_numpy(x[pick])) <= 1e-13
    del y
    z = bz.download(np.complex128, x.shape)
    assert oracle.rel_l2(z, x) <= 1e-13
    for b in (bx, by, bz):
        b.close()
    fwd.close()
    bwd.close()


def test_one_plan_on_several_streams_at_once(gpu):
    """Four streams run the same 17^3 x 1024 plan (a persistent kernel that claims blocks)
    three times each without synchronising in between; every result matches the oracle."""
    f, h = gpu.fft3d, gpu.hip
    shape, batch, nstream = (17, 17, 17), 1024, 4
    plan = f.Plan(shape, batch)
    xs = [_uniform(shape, batch, seed=40 + s) for s in range(nstream)]
    ins = [h.DeviceBuffer(x.nbytes) for x in xs]
    outs = [h.DeviceBuffer(x.nbytes) for x in xs]
    streams = [h.Stream() for _ in range(nstream)]
    for b, x in zip(ins, xs):
        b.upload(x)
    for rep in range(3):
        for s in range(nstream):
            plan.execute(ins[s].addr, outs[s].addr, stream=streams[s])
    h.synchronize()
    for s in range(nstream):
        y = outs[s].download(np.complex128, xs[s].shape)
        err = oracle.rel_l2(y, oracle.fftn_numpy(xs[s]))
        assert err <= 1e-13, (s, err)
    for b in ins + outs:
        b.close()
    for st in streams:
        st.close()
    plan.close()


@pytest.mark.parametrize("shape", [(12, 12, 12), (36, 36, 36)])
def test_partial_batches_leave_the_rest_untouched(gpu, shape):
    f, h = gpu.fft3d, gpu.hip
    batch = 300 if shape[0] == 12 else 40
    x = _uniform(shape, batch, seed=7)
    bi, bo = h.DeviceBuffer(x.nbytes), h.DeviceBuffer(x.nbytes)
    bi.upload(x)
    plan = f.Plan(shape, batch)
    poison = np.full(x.shape, np.nan + 1j * np.nan)
    ref = oracle.fftn_numpy(x)
    for n in (0, 1, 7, batch - 1, batch):
        bo.upload(poison)
        plan.execute(bi.addr, bo.addr, batch=n)
        h.synchronize()
        y = bo.download(np.complex128, x.shape)
        if n:
            assert oracle.rel_l2(y[:n], ref[:n]) <= 1e-13, n
        assert np.all(np.isnan(y[n:])), f"n={n}: volumes past the batch were written"
    with pytest.raises(RuntimeError):
        plan.execute(bi.addr, bo.addr, batch=batch + 1)
    for b in (bi, bo):
        b.close()
    plan.close()


def test_plans_do_not_leak_device_memory(gpu):
    f, h = gpu.fft3d, gpu.hip
    for shape in ((17, 17, 17), (36, 36, 36), (8, 8, 8)):  # warm the runtime's pools
        f.Plan(shape, 64, flags=f.NO_TUNE).close()
    h.synchronize()
    free0, _ = h.mem_info()
    for i in range(60):
        for shape in ((17, 17, 17), (36, 36, 36), (8, 8, 8)):
            f.Plan(shape, 64, flags=f.NO_TUNE).close()
    mp = [f.MultiPlan([((8, 8, 8), 16, -1), ((17, 17, 17), 3, 1)], f.C64) for _ in range(20)]
    for m in mp:
        m.close()
    h.synchronize()
    free1, _ = h.mem_info()
    assert free0 - free1 < (64 << 20), (free0, free1)
