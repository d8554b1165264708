"""fft3d_mx_execute_noalloc against the plan path it mirrors (fft3d_mx_create with
FFT3D_MX_NO_TUNE picks the same kernel): bit-identical outputs on two calls in a row, the
second on a smaller batch, through the C ABI; no device memory taken after set-up; sizes
and arguments beyond what it accepts refused with a message."""
import ctypes

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

CASES = [((17, 17, 17), 1024, 100, -1, 0, 0),
         ((8, 8, 8), 4096, 33, 1, 0, 0),
         ((36, 36, 36), 16, 3, -1, 0, 0),
         ((24, 24, 48), 8, 1, 1, 0, 0),
         ((12, 12, 12), 512, 7, -1, 1, 1),
         ((6, 6, 6), 3000, 5, 1, 1, 0)]


def test_noalloc_matches_plan_twice(gpu):
    f, h = gpu.fft3d, gpu.hip
    lib = f.load()
    # set-up call (batch 0): the only call that may move device memory
    assert lib.fft3d_mx_execute_noalloc(8, 8, 8, 0, -1, 0, 0, None, None, None) == 0
    for shape, b1, b2, d, prec, flags in CASES:
        print("noalloc", shape, b1, b2, d, prec, flags, flush=True)
        dt = np.complex128 if prec == 0 else np.complex64
        x = oracle.make_input(shape, b1, seed=sum(shape) + b1, dtype=dt)
        din = h.DeviceBuffer(x.nbytes)
        din.upload(x)
        dref, dout = h.DeviceBuffer(x.nbytes), h.DeviceBuffer(x.nbytes)
        plan = f.Plan(shape, b1, d, prec, flags=flags | f.NO_TUNE)
        stream = h.Stream()
        for batch in (b1, b2):
            dref.fill_zero()
            dout.fill_zero()
            plan.execute(din.addr, dref.addr, batch=batch, stream=stream)
            free0 = h.mem_info()[0]
            f.execute_noalloc(shape, batch, d, prec, din.addr, dout.addr, stream=stream,
                              flags=flags)
            stream.synchronize()
            assert h.mem_info()[0] >= free0  # nothing allocated on the device
            got = dout.download(dt, x.shape)
            ref = dref.download(dt, x.shape)
            assert np.array_equal(got[:batch].view(np.uint8), ref[:batch].view(np.uint8))
            assert not np.any(got[batch:])  # nothing written past the batch
            scale = 1.0 / (shape[0] * shape[1] * shape[2]) if flags else 1.0
            want = oracle.fftn_numpy(x[:min(batch, 8)], sign=d) * scale
            tol = 1e-13 if prec == 0 else 2e-6
            assert oracle.rel_l2(got[:min(batch, 8)], want) <= tol
        plan.close()
        stream.close()
        for b in (din, dref, dout):
            b.close()


def test_noalloc_refusals(gpu):
    f = gpu.fft3d
    lib = f.load()
    buf = gpu.hip.DeviceBuffer(1 << 20)
    for args, msg in [((8, 8, 8, 1 << 31, -1, 0, 0), "outside"),
                      ((8, 8, 8, -1, -1, 0, 0), "outside"),
                      ((7, 7, 7, 1, -1, 0, 0), "no kernel"),
                      ((8, 8, 8, 1, 0, 0, 0), "direction"),
                      ((8, 8, 8, 1, -1, 2, 0), "precision"),
                      ((8, 8, 8, 1, -1, 0, f.NO_TUNE), "flags")]:
        lib.fft3d_mx_clear_error()
        assert lib.fft3d_mx_execute_noalloc(*args, buf.addr, buf.addr + 8192, None) < 0
        assert msg in lib.fft3d_mx_last_error().decode()
    lib.fft3d_mx_clear_error()
    assert lib.fft3d_mx_execute_noalloc(8, 8, 8, 1, -1, 0, 0, buf.addr, buf.addr, None) < 0
    assert "distinct" in lib.fft3d_mx_last_error().decode()
    buf.close()
