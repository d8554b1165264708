"""Pins for the oracle (CPU only).

Every oracle function is checked against something independent of its own formula:
numpy.fft (pocketfft) vs the long-double dense DFT vs the six-fold definitional sum;
closed forms (delta, ones, a single mode -- which fixes the sign --, DFT^4 = V^2 I, FFT^2 =
V * index reversal); invariants (Parseval, linearity, shift theorem, Hermitian symmetry,
convolution theorem against brute force); the Rader-17 and PFA-36 constants of the
reference's strategy records; and golden fixtures.  The last group plants mistakes in the
oracle (wrong sign, dropped term, transposed axes, wrong normalisation) and shows that a
pin catches each one.
"""
import numpy as np
import pytest

import oracle

SHAPES = [(2, 3, 4), (6, 6, 6), (8, 8, 8), (5, 7, 3), (17, 2, 3), (12, 12, 12), (8, 8, 16),
          (1, 8, 8), (9, 4, 6)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("sign", [-1, 1])
def test_long_double_matches_numpy(shape, sign):
    x = oracle.make_input(shape, 2, seed=sum(shape), dist="normal")
    a = oracle.dft3d(x, sign=sign)
    b = oracle.fftn_numpy(x, sign=sign)
    assert oracle.rel_l2(a, b) < 1e-14


@pytest.mark.parametrize("shape", [(2, 3, 4), (3, 3, 3), (1, 5, 2), (4, 1, 3)])
def test_definitional_sum_matches(shape):
    x = oracle.make_input(shape, 1, seed=3)[0]
    for sign in (-1, 1):
        assert oracle.rel_l2(oracle.dft3d(x, sign=sign), oracle.dft3d_loops(x, sign=sign)) < 1e-14
        assert oracle.rel_l2(oracle.fftn_numpy(x, sign=sign), oracle.dft3d_loops(x, sign=sign)) < 1e-14


@pytest.mark.parametrize("fn", [oracle.dft3d, oracle.fftn_numpy])
def test_closed_forms(fn):
    shape = (4, 6, 5)
    V = 4 * 6 * 5
    delta = np.zeros(shape, complex)
    delta[0, 0, 0] = 1
    assert oracle.rel_l2(fn(delta), np.ones(shape)) < 1e-15
    spike = np.zeros(shape, complex)
    spike[0, 0, 0] = V
    assert oracle.rel_l2(fn(np.ones(shape, complex)), spike) < 1e-15
    # exp(+2 pi i k0.j/n) lands on k0 under the forward (-1) kernel: this pins the sign
    k0 = (1, 2, 3)
    j = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    mode = np.exp(2j * np.pi * sum(k * jj / n for k, jj, n in zip(k0, j, shape)))
    want = np.zeros(shape, complex)
    want[k0] = V
    assert oracle.rel_l2(fn(mode), want) < 1e-14
    # and the backward kernel sends the conjugate mode there
    assert oracle.rel_l2(fn(np.conj(mode), sign=1), want) < 1e-14


def test_dft_fourth_power_and_chain_closed_form():
    x = oracle.make_input((5, 3, 4), 1, seed=11)[0]
    V = 60
    y = x
    for _ in range(4):
        y = oracle.dft3d(y)
    assert oracle.rel_l2(y, V * V * x) < 1e-14
    # unitary chain: each step FFT then 1/sqrt(V) (bench/geom/check.py:43-61)
    y = x
    for m in range(1, 6):
        y = oracle.dft3d(y) / np.sqrt(V)
        assert oracle.rel_l2(y, oracle.chain_expected(x, m)) < 1e-14, m


def test_invariants():
    shape = (6, 4, 10)
    V = 240
    x = oracle.make_input(shape, 1, seed=1, dist="normal")[0]
    y = oracle.make_input(shape, 1, seed=2, dist="normal")[0]
    a, b = 0.3 - 1.7j, 2.1 + 0.4j
    X, Y = oracle.dft3d(x), oracle.dft3d(y)
    assert oracle.rel_l2(oracle.dft3d(a * x + b * y), a * X + b * Y) < 1e-14
    assert abs(np.sum(np.abs(X) ** 2) - V * np.sum(np.abs(x) ** 2)) / (V * np.sum(np.abs(x) ** 2)) < 1e-14
    shift = (2, 3, 1)
    k = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    phase = np.exp(-2j * np.pi * sum(kk * d / n for kk, d, n in zip(k, shift, shape)))
    assert oracle.rel_l2(oracle.dft3d(np.roll(x, shift, axis=(0, 1, 2))), phase * X) < 1e-14
    r = x.real.astype(complex)
    R = oracle.dft3d(r)
    mirrored = np.conj(R[np.ix_(*[(-np.arange(n)) % n for n in shape])])
    assert oracle.rel_l2(R, mirrored) < 1e-14
    # backward with 1/V inverts forward
    assert oracle.rel_l2(oracle.dft3d(X, sign=1, normalize=True), x) < 1e-14


def test_convolution_theorem_brute_force():
    shape = (3, 4, 5)
    x = oracle.make_input(shape, 1, seed=3)[0]
    y = oracle.make_input(shape, 1, seed=4)[0]
    direct = np.zeros(shape, complex)
    for a in range(3):
        for b in range(4):
            for c in range(5):
                direct += x[a, b, c] * np.roll(y, (a, b, c), axis=(0, 1, 2))
    got = oracle.dft3d(oracle.dft3d(x) * oracle.dft3d(y), sign=1, normalize=True)
    assert oracle.rel_l2(got, direct) < 1e-14


def test_rader17_constants():
    """Constants of bench/geom/strategies/L17_rader.md:21-55, and the construction itself."""
    perm, bh = oracle.rader17(-1)
    assert list(perm) == [1, 3, 9, 10, 13, 5, 15, 11, 16, 14, 8, 7, 4, 12, 2, 6]
    assert abs(bh[0] - (-1 / 16)) < 1e-17
    assert abs(bh[8] - np.sqrt(17) / 16) < 1e-16
    # the construction reproduces the 17-point DFT in both directions
    x = oracle.make_input((17,), 1, seed=17, dist="normal")[0]
    for sign in (-1, 1):
        perm, bh = oracle.rader17(sign)
        a = x[perm]
        A = np.fft.fft(a)
        P = A * bh
        P[0] += x[0]
        X = np.empty(17, complex)
        X[0] = x[0] + A[0]
        X[perm] = np.fft.fft(P)
        ref = np.exp(sign * 2j * np.pi * np.outer(np.arange(17), np.arange(17)) / 17) @ x
        assert oracle.rel_l2(X, ref) < 1e-14
        assert abs(bh[8].imag) < 1e-17 and abs(bh[0].imag) < 1e-17


def test_pfa36_maps():
    """bench/geom/strategies/L36_pfa.md:15-16: n=(9n1+4n2) mod 36, k=(9k1+28k2) mod 36."""
    inp, outp = oracle.pfa_maps(4, 9)
    assert sorted(inp.reshape(-1)) == list(range(36))
    assert sorted(outp.reshape(-1)) == list(range(36))
    assert inp[1, 0] == 9 and inp[0, 1] == 4 and outp[1, 0] == 9 and outp[0, 1] == 28
    for n1 in range(4):
        for n2 in range(9):
            for k1 in range(4):
                for k2 in range(9):
                    # W36^{nk} = W4^{n1 k1} W9^{n2 k2}, i.e. nk = 9 n1 k1 + 4 n2 k2 (mod 36)
                    assert (inp[n1, n2] * outp[k1, k2]) % 36 == (9 * n1 * k1 + 4 * n2 * k2) % 36


def test_metrics_and_generator():
    ref = np.array([3 + 4j, 0])
    assert oracle.rel_l2(ref, ref) == 0.0
    assert abs(oracle.rel_l2(np.array([3 + 4j, 1]), ref) - 0.2) < 1e-15
    assert abs(oracle.rel_max(np.array([3 + 4j, 1]), ref) - 0.2) < 1e-15
    u = oracle.make_input((8, 8, 8), 64, seed=5)
    assert u.dtype == np.complex128 and u.shape == (64, 8, 8, 8)
    assert np.all(np.abs(u.real) < 1) and np.all(np.abs(u.imag) < 1)
    assert abs(np.var(u.real) - 1 / 3) < 0.01
    g = oracle.make_input((8, 8, 8), 64, seed=5, dist="normal_half")
    assert abs(np.var(g.real) - 0.5) < 0.02 and abs(np.var(g.imag) - 0.5) < 0.02
    assert np.array_equal(u, oracle.make_input((8, 8, 8), 64, seed=5))
    assert oracle.make_input((6, 6, 6), 2, 1, dtype=np.complex64).dtype == np.complex64
    assert oracle.nominal_flops((8, 8, 8)) == 5 * 512 * 9


# ------------------------------------------------------------------ planted mistakes
def _planted_sign(x, sign=-1, normalize=False):
    return oracle.dft3d(x, sign=-sign, normalize=normalize)


def _planted_transpose(x, sign=-1, normalize=False):
    # transforms the right number of axes but with (n0,n1,n2) read as (n2,n1,n0)
    y = oracle.dft3d(np.ascontiguousarray(np.swapaxes(x, -1, -3)).reshape(x.shape), sign=sign)
    return y


def _planted_dropped_term(x, sign=-1, normalize=False):
    y = np.array(x, dtype=complex)
    y[..., -1] = 0  # drop the last z term of every line
    return oracle.dft3d(y, sign=sign, normalize=normalize)


def test_planted_mistakes_are_caught():
    shape = (4, 6, 5)
    j = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    k0 = (1, 2, 3)
    mode = np.exp(2j * np.pi * sum(k * jj / n for k, jj, n in zip(k0, j, shape)))
    want = np.zeros(shape, complex)
    want[k0] = 120
    # wrong sign: the single-mode pin fails
    assert oracle.rel_l2(_planted_sign(mode), want) > 0.5
    # transposed operand on an anisotropic shape: numpy disagrees
    x = oracle.make_input(shape, 1, seed=9)[0]
    assert oracle.rel_l2(_planted_transpose(x), oracle.fftn_numpy(x)) > 0.5
    # dropped term: Parseval fails
    X = _planted_dropped_term(x)
    assert abs(np.sum(np.abs(X) ** 2) - 120 * np.sum(np.abs(x) ** 2)) / (120 * np.sum(np.abs(x) ** 2)) > 1e-3
    # missing 1/V: the round trip fails
    assert oracle.rel_l2(oracle.dft3d(oracle.dft3d(x), sign=1, normalize=False), x) > 1.0
