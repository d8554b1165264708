"""ORACLE -- test infrastructure only.

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import this module, and only as the checker (or as the timed CPU port).  The product in
``opt-fft_amd/`` never imports it and has no CPU fallback.

Restatement of the reference's FP64 oracle path for the batched small-volume 3D FFT:

* ``fftn_numpy``   -- ``np.fft.fftn(x, axes=(-3,-2,-1))`` exactly as bench/geom/check.py:28
                      does, extended with the backward sign of python/slow_dft.py:27-28.
* ``dft3d``        -- the long-double dense separable DFT in ``dft_ref.c`` (the method of
                      python/slow_dft.py:93-106 and python/test_fft3d.py:546-555).
* ``dft3d_loops``  -- the definitional six-fold sum, python/slow_dft.py:54-73 (tiny shapes).
* ``rel_l2``, ``rel_max`` -- the accuracy metrics of bench/geom/check.py:29-33.
* ``chain_expected`` -- the unitary chain closed form of bench/geom/check.py:43-61.
* ``make_input``   -- the seeded generator of bench/geom/gen_input.py:12-15, with the
                      per-workload distributions (uniform(-1,1) or N(0,1/2) per component).
* ``rader17``      -- Rader-17 plan constants of bench/geom/strategies/L17_rader.md:21-55.
* ``pfa_maps``     -- Good-Thomas index maps of bench/geom/strategies/L36_pfa.md:10-24.

Parity of every function here is pinned in tests/test_oracle.py (numpy vs long double vs
brute force, plus closed forms); none is "parity unpinned".
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

FORWARD = -1
BACKWARD = +1

_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
        lib = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        fp = ctypes.POINTER(ctypes.c_float)
        ip = ctypes.POINTER(ctypes.c_int)
        lib.oracle_dft3d_c128.argtypes = [dp, dp, ctypes.c_long, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_int, ctypes.c_int]
        lib.oracle_dft3d_c128.restype = ctypes.c_int
        lib.oracle_dft3d_c64.argtypes = [fp, dp, ctypes.c_long, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_int]
        lib.oracle_dft3d_c64.restype = ctypes.c_int
        lib.oracle_rader17.argtypes = [ctypes.c_int, ip, dp]
        lib.oracle_rader17.restype = None
        _lib = lib
    return _lib


def _as_batched(x):
    x = np.asarray(x)
    if x.ndim == 3:
        return x[None], True
    if x.ndim != 4:
        raise ValueError(f"expected (B,n0,n1,n2) or (n0,n1,n2), got shape {x.shape}")
    return x, False


def dft3d(x, sign=FORWARD, normalize=False):
    """Long-double dense separable 3D DFT over the last three axes; returns complex128."""
    xb, squeeze = _as_batched(x)
    B, n0, n1, n2 = xb.shape
    out = np.empty((B, n0, n1, n2), dtype=np.complex128)
    lib = _load()
    dp = ctypes.POINTER(ctypes.c_double)
    if xb.dtype == np.complex64:
        src = np.ascontiguousarray(xb)
        rc = lib.oracle_dft3d_c64(src.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                  out.ctypes.data_as(dp), B, n0, n1, n2, sign, int(normalize))
    else:
        src = np.ascontiguousarray(xb, dtype=np.complex128)
        rc = lib.oracle_dft3d_c128(src.ctypes.data_as(dp), out.ctypes.data_as(dp),
                                   B, n0, n1, n2, sign, int(normalize))
    if rc != 0:
        raise ValueError("oracle_dft3d rejected its arguments")
    return out[0] if squeeze else out


def fftn_numpy(x, sign=FORWARD, normalize=False):
    """numpy.fft over the last three axes in complex128 (bench/geom/check.py:28).

    Forward is unnormalised.  Backward is the FFTW-style unnormalised sum unless
    ``normalize`` is set, which gives numpy's ``ifftn`` (1/V)."""
    x = np.asarray(x).astype(np.complex128)
    axes = (-3, -2, -1)
    if sign == FORWARD:
        y = np.fft.fftn(x, axes=axes)
        if normalize:
            y = y / float(np.prod(x.shape[-3:]))
        return y
    if sign == BACKWARD:
        y = np.fft.ifftn(x, axes=axes)
        if not normalize:
            y = y * float(np.prod(x.shape[-3:]))
        return y
    raise ValueError("sign must be -1 or +1")


def dft3d_loops(x, sign=FORWARD):
    """The six-fold definitional sum (python/slow_dft.py:54-73); keep the volume tiny."""
    x = np.asarray(x, dtype=np.complex128)
    n0, n1, n2 = x.shape
    out = np.zeros_like(x)
    for k0 in range(n0):
        for k1 in range(n1):
            for k2 in range(n2):
                acc = 0j
                for j0 in range(n0):
                    for j1 in range(n1):
                        for j2 in range(n2):
                            ph = (k0 * j0 % n0) / n0 + (k1 * j1 % n1) / n1 + (k2 * j2 % n2) / n2
                            acc += x[j0, j1, j2] * np.exp(sign * 2j * np.pi * ph)
                out[k0, k1, k2] = acc
    return out


def rel_l2(got, ref):
    """Relative L2 error over the flattened batch (bench/geom/check.py:29-30)."""
    got = np.asarray(got).astype(np.complex128).reshape(-1)
    ref = np.asarray(ref).astype(np.complex128).reshape(-1)
    denom = np.linalg.norm(ref)
    return float(np.linalg.norm(got - ref) / (denom if denom else 1.0))


def rel_max(got, ref):
    """max|got-ref| / max|ref| (bench/geom/check.py:31-33)."""
    got = np.asarray(got).astype(np.complex128)
    ref = np.asarray(ref).astype(np.complex128)
    m = float(np.abs(ref).max()) if ref.size else 0.0
    return float(np.abs(got - ref).max() / (m if m else 1.0)) if ref.size else 0.0


def chain_expected(x, m):
    """End state of m unitary forward steps (each FFT then 1/sqrt(V)), the closed form of
    bench/geom/check.py:43-61: FFT^2 = V * index reversal."""
    x = np.asarray(x, dtype=np.complex128)
    n0, n1, n2 = x.shape[-3:]
    rev = x[..., (-np.arange(n0)) % n0, :, :][..., :, (-np.arange(n1)) % n1, :][..., (-np.arange(n2)) % n2]
    r = m % 4
    v = float(n0 * n1 * n2)
    if r == 0:
        return x.copy()
    if r == 1:
        return fftn_numpy(x) / np.sqrt(v)
    if r == 2:
        return rev
    return fftn_numpy(rev) / np.sqrt(v)


def make_input(shape, batch, seed, dist="uniform", dtype=np.complex128):
    """Seeded complex input (B, n0, n1, n2), real part drawn first, then imaginary, each a
    full-shape draw (bench/geom/gen_input.py:12-14).

    dist: "uniform" -> U(-1,1) per component; "normal_half" -> N(0, 1/2) per component
    (standard complex normal); "normal" -> N(0,1) per component (the reference default)."""
    rng = np.random.default_rng(seed)
    full = (int(batch),) + tuple(int(n) for n in shape)
    if dist == "uniform":
        re = rng.uniform(-1.0, 1.0, full)
        im = rng.uniform(-1.0, 1.0, full)
    elif dist == "normal_half":
        s = np.sqrt(0.5)
        re = rng.normal(0.0, s, full)
        im = rng.normal(0.0, s, full)
    elif dist == "normal":
        re = rng.standard_normal(full)
        im = rng.standard_normal(full)
    else:
        raise ValueError(f"unknown dist {dist!r}")
    return (re + 1j * im).astype(dtype)


def rader17(sign=FORWARD):
    """(perm[16], Bhat[16] complex) for Rader-17 with generator 3."""
    lib = _load()
    perm = (ctypes.c_int * 16)()
    bh = (ctypes.c_double * 32)()
    lib.oracle_rader17(sign, perm, bh)
    b = np.frombuffer(bh, dtype=np.float64).copy()
    return np.array(list(perm), dtype=np.int64), b[0::2] + 1j * b[1::2]


def pfa_maps(n1, n2):
    """Good-Thomas maps for N = n1*n2, gcd(n1,n2)=1 (bench/geom/strategies/L36_pfa.md:10-24):
    input  n = (n2*a + n1*b) mod N            for (a,b) in [0,n1) x [0,n2)
    output k = (n2*(n2^-1 mod n1)*c + n1*(n1^-1 mod n2)*d) mod N."""
    N = n1 * n2
    e1 = n2 * pow(n2, -1, n1) % N
    e2 = n1 * pow(n1, -1, n2) % N
    inp = np.array([[(n2 * a + n1 * b) % N for b in range(n2)] for a in range(n1)])
    outp = np.array([[(e1 * c + e2 * d) % N for d in range(n2)] for c in range(n1)])
    return inp, outp


def nominal_flops(shape, batch=1):
    """5 V log2 V per transform (bench/geom/driver.c:256-261)."""
    v = float(np.prod(shape))
    return 5.0 * v * np.log2(v) * batch
