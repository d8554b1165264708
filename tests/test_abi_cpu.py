"""CPU checks of the C-ABI library: it loads, exports every symbol the public headers
declare, enumerates its kernels consistently, builds the plan-time constant bank correctly
(checked against the oracle's independent computation and closed forms), and rejects bad
arguments through its error channel.  No transform runs here."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in
           ("fft3d_api.h", "fft3d_gpu_api.h", "fft3d_mi355x.h", "fft3d_serve.h")]


def declared_functions():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(fft3d_\w+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_every_declared_symbol_is_exported(pkg):
    lib = pkg.fft3d.load()
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    assert set(pkg.fft3d.EXPORTED) == names


def test_names_and_supports(pkg):
    lib = pkg.fft3d.load()
    assert lib.fft3d_name() == b"mi355x_hip"
    assert b" " not in lib.fft3d_name() and b" " not in lib.fft3d_gpu_name()
    assert lib.fft3d_description() and lib.fft3d_gpu_description()
    for L in (6, 8, 12, 17, 24, 36):
        assert lib.fft3d_supports(L) and lib.fft3d_gpu_supports(L)
    for L in (0, 1, 5, 7, 13, 64, -8):
        assert not lib.fft3d_supports(L) and not lib.fft3d_gpu_supports(L)


def test_kernel_registry_is_complete(pkg):
    ks = pkg.fft3d.kernels()
    names = [k[0] for k in ks]
    assert len(names) == len(set(names)), "duplicate kernel names"
    geoms = {(6, 6, 6), (8, 8, 8), (12, 12, 12), (17, 17, 17), (24, 24, 24), (36, 36, 36),
             (24, 24, 48), (8, 8, 16)}
    for g in geoms:
        for prec in (0, 1):
            assert pkg.fft3d.supports(g, prec), (g, prec)
            for d in (-1, 1):
                assert any(k[1] == g and k[2] == prec and k[3] == d for k in ks), (g, prec, d)
    for name, shape, prec, d in ks:
        assert f"{shape[0]}x{shape[1]}x{shape[2]}" in name
        assert ("c128" if prec == 0 else "c64") in name
        assert ("fwd" if d == -1 else "bwd") in name
    lib = pkg.fft3d.load()
    assert lib.fft3d_mx_kernel_name(len(ks), None, None, None, None, None) is None
    assert lib.fft3d_mx_kernel_name(-1, None, None, None, None, None) is None


def test_volume_bytes(pkg):
    lib = pkg.fft3d.load()
    assert lib.fft3d_mx_volume_bytes(17, 17, 17, 0) == 78608
    assert lib.fft3d_mx_volume_bytes(17, 17, 17, 1) == 39304
    assert lib.fft3d_mx_volume_bytes(0, 17, 17, 0) == 0
    assert lib.fft3d_mx_volume_bytes(8, 8, 8, 7) == 0


LENS = [2, 3, 4, 6, 8, 9, 12, 16, 17, 24, 36, 48]


def _split_bank(b, prec):
    nw = sum(LENS)
    w = b[: 2 * nw].reshape(-1, 2)
    w = w[:, 0] + 1j * w[:, 1]
    rad = b[2 * nw: 2 * nw + 64].reshape(2, 16, 2)
    rad = rad[..., 0] + 1j * rad[..., 1]
    cos17 = b[2 * nw + 64: 2 * nw + 81]
    sin17 = b[2 * nw + 81: 2 * nw + 98]
    assert b.size == 2 * nw + 98
    return w, rad, cos17, sin17


@pytest.mark.parametrize("prec", [0, 1])
def test_bank_twiddles(pkg, prec):
    b = pkg.fft3d.bank_image(prec)
    w, rad, c17, s17 = _split_bank(b, prec)
    # reference in extended precision (the bank is long double rounded once)
    tol = 1.2e-16 if prec == 0 else 6e-8
    pi_ld = np.longdouble("3.14159265358979323846264338327950288")
    off = 0
    for n in LENS:
        a = 2 * pi_ld * np.arange(n, dtype=np.longdouble) / np.longdouble(n)
        ref = (np.cos(a) - 1j * np.sin(a)).astype(np.complex128)
        assert np.max(np.abs(w[off:off + n] - ref)) <= tol, n
        # quarter turns are exact
        for q in range(4):
            if (q * n) % 4 == 0:
                e = w[off + q * n // 4]
                assert e == np.round(ref[q * n // 4].real) + 1j * np.round(ref[q * n // 4].imag)
        off += n
    a = 2 * pi_ld * np.arange(17, dtype=np.longdouble) / np.longdouble(17)
    assert np.max(np.abs(c17 - np.cos(a).astype(np.float64))) <= tol
    assert np.max(np.abs(s17 - np.sin(a).astype(np.float64))) <= tol


@pytest.mark.parametrize("prec", [0, 1])
def test_bank_rader_spectra(pkg, prec):
    """The plan step's Rader-17 spectra against the oracle's independent C computation and
    the closed forms Bhat[0] = -1/16, Bhat[8] = sqrt(17)/16 (bench/geom/strategies/L17_rader.md)."""
    _, rad, _, _ = _split_bank(pkg.fft3d.bank_image(prec), prec)
    tol = 1e-17 if prec == 0 else 1e-8
    for d, sign in ((0, -1), (1, 1)):
        _, ref = oracle.rader17(sign)
        if prec == 1:
            ref = ref.astype(np.complex64)
        assert np.max(np.abs(rad[d] - ref)) <= tol
        assert rad[d][0] == np.float64(-1 / 16) if prec == 0 else abs(rad[d][0] + 1 / 16) < 1e-8
        assert rad[d][0].imag == 0 and rad[d][8].imag == 0
        assert abs(rad[d][8].real - math.sqrt(17) / 16) <= (1e-17 if prec == 0 else 1e-8)
    # backward spectrum is the index-negated conjugate of the forward one
    neg = (-np.arange(16)) % 16
    assert np.max(np.abs(rad[1] - np.conj(rad[0][neg]))) <= (1e-17 if prec == 0 else 1e-7)


def test_create_rejects_bad_arguments(pkg):
    lib = pkg.fft3d.load()
    cases = [((7, 7, 7, 4, -1, 0, 0), "no kernel"),
             ((8, 8, 8, 4, 0, 0, 0), "direction"),
             ((8, 8, 8, 4, -1, 5, 0), "precision"),
             ((0, 8, 8, 4, -1, 0, 0), "extents"),
             ((8, 8, 8, -1, -1, 0, 0), "batch")]
    for args, word in cases:
        lib.fft3d_mx_clear_error()
        assert not lib.fft3d_mx_create(*args)
        assert word in lib.fft3d_mx_last_error().decode()
    assert not lib.fft3d_mx_create_named(8, 8, 8, 4, -1, 0, 0, b"")
    # a registered kernel of another geometry, and a name no kernel has
    other = b"pipe_17x17x17_c128_fwd_v1t320p1_fold"
    assert other.decode() in [k[0] for k in pkg.fft3d.kernels()]
    for name in (other, b"no_such_kernel"):
        lib.fft3d_mx_clear_error()
        assert not lib.fft3d_mx_create_named(8, 8, 8, 4, -1, 0, 0, name)
        assert "does not serve" in lib.fft3d_mx_last_error().decode()
    assert lib.fft3d_mx_execute(None, None, None, None) == -1
    assert lib.fft3d_mx_plan_get_info(None, None) == -1
    assert not lib.fft3d_create(7, 1) and not lib.fft3d_gpu_create(7, 1)
    assert not lib.fft3d_create(8, 0)


def test_create_without_device_fails_cleanly(pkg):
    if pkg.hip.device_count() > 0:
        pytest.skip("a HIP device is present")
    lib = pkg.fft3d.load()
    lib.fft3d_mx_clear_error()
    assert not lib.fft3d_mx_create(8, 8, 8, 4, -1, 0, 0)
    assert lib.fft3d_mx_last_error().decode() != ""
