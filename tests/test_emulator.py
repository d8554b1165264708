"""Host emulation of every kernel variant (CPU only).

tests/emu builds the kernel translation units host-only with an emulating launcher: each
workgroup's phases run in barrier order, threads one after another (ascending, then again
descending -- a result that depends on the order means an intra-phase race), LDS starts
NaN-filled and is fenced by guard elements.  This checks the kernels' index arithmetic,
tail handling and codelets against the oracle before any GPU run; the ASan sweep checks
every global access against exactly-sized buffers.
"""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EMU = os.path.join(ROOT, "tests", "emu")


@pytest.fixture(scope="module")
def emu():
    subprocess.run(["make", "-s", "-C", EMU, "libfft3d_emu.so"], check=True)
    lib = ctypes.CDLL(os.path.join(EMU, "libfft3d_emu.so"))
    lib.emu_kernel_name.restype = ctypes.c_char_p
    lib.emu_transform.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_void_p,
                                  ctypes.c_long, ctypes.c_double, ctypes.c_int]
    return lib


def parse(name):
    m = re.search(r"_(\d+)x(\d+)x(\d+)_(c128|c64)_(fwd|bwd)_[sv](\d+)", name)
    shape = tuple(int(m.group(k)) for k in (1, 2, 3))
    return shape, (0 if m.group(4) == "c128" else 1), (-1 if m.group(5) == "fwd" else 1), int(m.group(6))


def test_emulator_matches_product_registry(emu, pkg):
    names = [emu.emu_kernel_name(i).decode() for i in range(emu.emu_kernel_count())]
    assert names == [k[0] for k in pkg.fft3d.kernels()]


def test_every_variant_on_the_emulator(emu):
    n = emu.emu_kernel_count()
    assert n >= 100
    for i in range(n):
        name = emu.emu_kernel_name(i).decode()
        shape, prec, d, per_wg = parse(name)
        dt = np.complex128 if prec == 0 else np.complex64
        tol = 1e-13 if prec == 0 else 2e-6
        vol = shape[0] * shape[1] * shape[2]
        # pencil groups count slabs, not volumes: a batch past one group suffices there
        batch = 2 if vol > 10000 else per_wg + 1
        x = oracle.make_input(shape, batch, seed=i, dist="uniform", dtype=dt)
        for reverse in (0, 1):
            out = np.full_like(x, np.nan)
            rc = emu.emu_transform(name.encode(), x.ctypes.data, out.ctypes.data, batch, 0.25,
                                   reverse)
            assert rc == 0, f"{name}: LDS guard broken in block {-1000 - rc}"
            if reverse == 0:
                first = out.copy()
                err = oracle.rel_l2(out, 0.25 * oracle.fftn_numpy(x, sign=d))
                assert err <= tol, f"{name} B={batch}: {err:.3e}"
            else:
                assert np.array_equal(first.view(np.uint8), out.view(np.uint8)), \
                    f"{name}: result depends on thread order"


def test_asan_sweep():
    """Every variant at 4 batch sizes with exactly-sized heap buffers under AddressSanitizer."""
    subprocess.run(["make", "-s", "-C", EMU, "emu_asan"], check=True)
    r = subprocess.run([os.path.join(EMU, "emu_asan")], capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "ok:" in r.stdout
