"""GPU: the request server (include/fft3d_serve.h).

Each request's result, computed inside a continuously batched mix, matches the oracle
(complex128 <= 1e-13, complex64 <= 2e-6 relative L2) and is bit-identical to the same
request served alone; admission never exceeds the byte budget; the plain path agrees."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _mix(gpu, n, cap, seed):
    s = gpu.serving
    return [(shape, min(b, cap), d, p) for shape, b, d, p in s.draw_mix(n, seed)]


def _setup(gpu, mix, seed):
    f, h = gpu.fft3d, gpu.hip
    xs, ins, outs = [], [], []
    for k, (shape, b, d, p) in enumerate(mix):
        dt = np.complex128 if p == f.C128 else np.complex64
        x = oracle.make_input(shape, b, seed=seed + k, dtype=dt)
        bi, bo = h.DeviceBuffer(x.nbytes), h.DeviceBuffer(x.nbytes)
        bi.upload(x)
        xs.append(x)
        ins.append(bi)
        outs.append(bo)
    return xs, ins, outs


def _check(gpu, mix, xs, outs):
    res = []
    for (shape, b, d, p), x, bo in zip(mix, xs, outs):
        y = bo.download(x.dtype, x.shape)
        err = oracle.rel_l2(y, oracle.fftn_numpy(x, sign=d))
        assert err <= (1e-13 if p == 0 else 2e-6), (shape, b, d, p, err)
        res.append(y)
    return res


@pytest.mark.parametrize("streams,window_us", [(4, 0.0), (16, 30.0)])
def test_mix_matches_oracle_and_alone(gpu, streams, window_us):
    f = gpu.fft3d
    mix = _mix(gpu, 120, 24, seed=3)
    xs, ins, outs = _setup(gpu, mix, 100)
    srv = f.Server(budget_bytes=1 << 30, streams=streams, batched=True, window_us=window_us)
    reqs = [f.SrvRequest(*shape, b, d, p, bi.addr, bo.addr)
            for (shape, b, d, p), bi, bo in zip(mix, ins, outs)]
    # a burst: everything arrives within 200 us, so launches pack several requests
    lat = srv.run_trace(reqs, np.linspace(0, 2e-4, len(reqs)))
    assert np.all(lat > 0)
    st = srv.stats()
    print(st, flush=True)
    assert st["completed"] == len(reqs)
    assert st["mixed_launches"] > 0 and st["mean_items_per_mixed_launch"] > 1.0
    mixed = _check(gpu, mix, xs, outs)
    # the same requests one at a time through the same server
    for k, ((shape, b, d, p), bi, bo) in enumerate(zip(mix, ins, outs)):
        bo.fill_zero()
        rid = srv.submit(shape, b, d, p, bi.addr, bo.addr)
        srv.wait(rid)
        alone = bo.download(xs[k].dtype, xs[k].shape)
        assert np.array_equal(alone.view(np.uint8), mixed[k].view(np.uint8)), (k, shape)
    srv.close()
    for b in ins + outs:
        b.close()


@pytest.mark.parametrize("mode,streams", [(0, 1), (2, 4), (2, 16)])
def test_plain_path_matches_oracle(gpu, mode, streams):
    """The plain one-launch-per-request modes: synchronous, and asynchronous with several
    launches in flight."""
    f = gpu.fft3d
    mix = _mix(gpu, 60, 16, seed=8)
    xs, ins, outs = _setup(gpu, mix, 300)
    srv = f.Server(budget_bytes=1 << 30, streams=streams, batched=mode)
    reqs = [f.SrvRequest(*shape, b, d, p, bi.addr, bo.addr)
            for (shape, b, d, p), bi, bo in zip(mix, ins, outs)]
    srv.run_trace(reqs, np.linspace(0, 1e-3, len(reqs)))
    assert srv.stats()["mixed_launches"] == 0
    _check(gpu, mix, xs, outs)
    srv.close()
    for b in ins + outs:
        b.close()


def test_bad_configuration_is_rejected(gpu):
    f = gpu.fft3d
    for kw in (dict(streams=0), dict(streams=17), dict(max_items=0), dict(max_items=17),
               dict(window_us=-1.0), dict(window_us=2e6), dict(budget_bytes=1 << 10),
               dict(batched=3), dict(batched=-1)):
        with pytest.raises(RuntimeError):
            f.Server(**kw)


def test_admission_respects_budget(gpu):
    f = gpu.fft3d
    budget = 48 << 20
    mix = [((36, 36, 36), 8, -1, 0)] * 6 + [((17, 17, 17), 64, 1, 0)] * 6
    xs, ins, outs = _setup(gpu, mix, 500)
    srv = f.Server(budget_bytes=budget, streams=4, batched=True)
    reqs = [f.SrvRequest(*shape, b, d, p, bi.addr, bo.addr)
            for (shape, b, d, p), bi, bo in zip(mix, ins, outs)]
    srv.run_trace(reqs, np.zeros(len(reqs)))
    st = srv.stats()
    assert st["completed"] == len(reqs)
    assert 0 < st["peak_inflight_bytes"] <= budget
    _check(gpu, mix, xs, outs)
    with pytest.raises(ValueError):
        srv.submit((36, 36, 36), 512, -1, 0, ins[0].addr)   # larger than the budget
    srv.close()
    for b in ins + outs:
        b.close()
