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
    m = re.search(r"_(\d+)x(\d+)x(\d+)_(c128|c64)_(fwd|bwd)_[svwl](\d+)", name)
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
        if name.startswith("mfma_"):
            continue  # wave-level matrix-core kernels: checked on the GPU only
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


def test_pipelined_variants_claim_blocks_dynamically(emu):
    """Persistent pipelined kernels on a batch of 8 blocks: the emulated grid of 3 workgroups
    draws blocks from the launch's claim counter in interleaved order (both thread orders),
    and the counter must come back at zero (emu_transform returns -3 otherwise)."""
    n = emu.emu_kernel_count()
    seen = 0
    for i in range(n):
        name = emu.emu_kernel_name(i).decode()
        if not name.startswith(("pipe_", "bal_", "split_")) or name.endswith("p0") or "p0_" in name:
            continue
        shape, prec, d, per_wg = parse(name)
        dt = np.complex128 if prec == 0 else np.complex64
        tol = 1e-13 if prec == 0 else 2e-6
        batch = 7 * per_wg + 1
        x = oracle.make_input(shape, batch, seed=i, dist="uniform", dtype=dt)
        ref = oracle.fftn_numpy(x, sign=d)
        for reverse in (0, 1):
            out = np.full_like(x, np.nan)
            rc = emu.emu_transform(name.encode(), x.ctypes.data, out.ctypes.data, batch, 1.0,
                                   reverse)
            assert rc == 0, f"{name}: rc {rc}"
            assert oracle.rel_l2(out, ref) <= tol, name
        seen += 1
    assert seen >= 20


def test_residue_split_variants_over_several_volume_groups(emu):
    """Residue-split kernels at batch 19 (two full groups of eight volumes, one per XCD, and
    a ragged third whose surplus workgroups return at once), both thread orders, into a
    NaN-filled output: every element written, results against the oracle."""
    seen = 0
    for i in range(emu.emu_kernel_count()):
        name = emu.emu_kernel_name(i).decode()
        if not name.startswith("xres_"):
            continue
        shape, prec, d, _ = parse(name)
        dt = np.complex128 if prec == 0 else np.complex64
        tol = 1e-13 if prec == 0 else 2e-6
        batch = 19
        x = oracle.make_input(shape, batch, seed=i, dist="uniform", dtype=dt)
        ref = oracle.fftn_numpy(x, sign=d)
        for reverse in (0, 1):
            out = np.full_like(x, np.nan)
            rc = emu.emu_transform(name.encode(), x.ctypes.data, out.ctypes.data, batch, 1.0,
                                   reverse)
            assert rc == 0, f"{name}: rc {rc}"
            assert np.all(np.isfinite(out)), f"{name}: element left unwritten"
            assert oracle.rel_l2(out, ref) <= tol, name
        seen += 1
    assert seen >= 10


def test_asan_sweep():
    """Every variant at 4 batch sizes with exactly-sized heap buffers under AddressSanitizer."""
    subprocess.run(["make", "-s", "-C", EMU, "emu_asan"], check=True)
    r = subprocess.run([os.path.join(EMU, "emu_asan")], capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "ok:" in r.stdout


MIX = [((6, 6, 6), 20, -1), ((8, 8, 8), 11, 1), ((12, 12, 12), 3, -1), ((17, 17, 17), 2, 1),
       ((8, 8, 16), 5, -1), ((8, 8, 8), 1, -1), ((6, 6, 6), 0, 1)]


def test_mixed_batch_kernels_on_the_emulator(emu):
    emu.emu_multi_name.restype = ctypes.c_char_p
    n = len(MIX)
    assert emu.emu_multi_count() >= 4
    for i in range(emu.emu_multi_count()):
        name = emu.emu_multi_name(i)
        if b"mfma" in name:
            continue  # wave-level matrix-core kernel: checked on the GPU only
        prec = 0 if b"c128" in name else 1
        dt = np.complex128 if prec == 0 else np.complex64
        xs = [oracle.make_input(s, max(b, 1), seed=k, dtype=dt)[:b] for k, (s, b, d) in enumerate(MIX)]
        ys = [np.full_like(x, np.nan) for x in xs]
        scale = [1.0 if d < 0 else 0.5 for _, _, d in MIX]
        rc = emu.emu_multi(name, n, (ctypes.c_int * (3 * n))(*[v for s, _, _ in MIX for v in s]),
                           (ctypes.c_int * n)(*[b for _, b, _ in MIX]),
                           (ctypes.c_int * n)(*[d for _, _, d in MIX]),
                           (ctypes.c_double * n)(*scale),
                           (ctypes.c_void_p * n)(*[x.ctypes.data for x in xs]),
                           (ctypes.c_void_p * n)(*[y.ctypes.data for y in ys]))
        assert rc == 0, (name, rc)
        for x, y, (s, b, d), sc in zip(xs, ys, MIX, scale):
            if b == 0:
                continue
            err = oracle.rel_l2(y, sc * oracle.fftn_numpy(x, sign=d))
            assert err <= (1e-13 if prec == 0 else 2e-6), (name, s, d, err)


@pytest.mark.parametrize("grid", [1, 3, 31, 32, 33, 256, 2048])
def test_claim_queues_hand_out_every_block_once(emu, grid):
    """The claim queues (device arithmetic, host counters) at real grid sizes: every block
    runs exactly once and every counter returns to zero, whatever the interleaving."""
    emu.emu_claim_sim.argtypes = [ctypes.c_long, ctypes.c_long, ctypes.c_uint]
    for nblocks in sorted({grid, grid + 1, 2 * grid, 2 * grid + 31, 10 * grid + 7}):
        for seed in range(3):
            assert emu.emu_claim_sim(grid, nblocks, seed) == 0, (grid, nblocks, seed)
