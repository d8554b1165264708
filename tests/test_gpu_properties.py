"""GPU: properties the DFT must satisfy, checked at the benchmark workload sizes without
any reference implementation (the closed forms and identities of python/test_fft3d.py:
201-279, applied to the product path):

  delta -> all ones; ones -> V at k = 0; one Fourier mode -> V at its frequency
  (the sign convention: exp(+2 pi i k0.j/n) lands on +k0 under the forward transform);
  linearity; Parseval (sum |X|^2 = V sum |x|^2); the shift theorem.

complex128 bar 1e-13 relative (the path's tolerance); complex64 2e-6.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

WORKLOADS = [((8, 8, 8), 4096, -1), ((17, 17, 17), 1024, -1), ((17, 17, 17), 1024, 1),
             ((36, 36, 36), 128, -1), ((24, 24, 48), 64, -1)]


def _plan(f, shape, batch, d, prec=0):
    return f.Plan(shape, batch, d, prec)


def _rand(shape, batch, seed, dtype=np.complex128):
    rng = np.random.default_rng(seed)
    full = (batch,) + tuple(shape)
    return (rng.uniform(-1, 1, full) + 1j * rng.uniform(-1, 1, full)).astype(dtype)


@pytest.mark.parametrize("shape,batch,d", WORKLOADS)
def test_closed_forms(gpu, shape, batch, d):
    f = gpu.fft3d
    V = shape[0] * shape[1] * shape[2]
    p = _plan(f, shape, batch, d)
    print(f"{shape} B={batch} dir={d}: {p.kernel}", flush=True)
    x = np.zeros((batch,) + shape, np.complex128)
    x[:, 0, 0, 0] = 1.0
    y = p.run(x)
    assert np.max(np.abs(y - 1.0)) <= 1e-15, "delta -> ones"
    y = p.run(np.ones_like(x))
    ref = np.zeros_like(x)
    ref[:, 0, 0, 0] = V
    assert np.max(np.abs(y - ref)) <= 1e-13 * V, "ones -> V delta"
    # volume b carries mode k0(b) = (b, 3b, 5b) mod shape
    j = np.indices(shape)
    for b in range(min(batch, 8)):
        k0 = (b % shape[0], (3 * b) % shape[1], (5 * b) % shape[2])
        ph = sum(k0[a] * j[a] / shape[a] for a in range(3))
        x[b] = np.exp(2j * np.pi * ph)
    y = p.run(x)
    for b in range(min(batch, 8)):
        k0 = (b % shape[0], (3 * b) % shape[1], (5 * b) % shape[2])
        # forward (-1) puts exp(+2 pi i k0.j/n) at +k0, backward (+1) at -k0
        k = tuple((c if d < 0 else -c) % n for c, n in zip(k0, shape))
        ref = np.zeros(shape, np.complex128)
        ref[k] = V
        assert np.max(np.abs(y[b] - ref)) <= 1e-13 * V, (b, k0)
    p.close()


@pytest.mark.parametrize("shape,batch,d", WORKLOADS)
def test_linearity_parseval_shift(gpu, shape, batch, d):
    f = gpu.fft3d
    V = shape[0] * shape[1] * shape[2]
    p = _plan(f, shape, batch, d)
    x = _rand(shape, batch, seed=batch + 1)
    z = _rand(shape, batch, seed=batch + 2)
    a, c = 0.75 - 0.5j, -1.25 + 2.0j
    X, Z, W = p.run(x), p.run(z), p.run(a * x + c * z)
    lin = np.linalg.norm((W - (a * X + c * Z)).ravel()) / np.linalg.norm(W.ravel())
    assert lin <= 1e-13, f"linearity {lin:.2e}"
    ex = np.sum(np.abs(X) ** 2, axis=(1, 2, 3))
    ey = V * np.sum(np.abs(x) ** 2, axis=(1, 2, 3))
    assert np.max(np.abs(ex - ey) / ey) <= 1e-13, "Parseval"
    # shift by s in every axis multiplies X[k] by exp(d 2 pi i k.s/n) ... in the forward
    # convention exp(-2 pi i k.s/n) for x[j - s]
    s = (1, 2, 3)
    Y = p.run(np.roll(x, s, axis=(1, 2, 3)))
    k = np.indices(shape)
    ph = np.exp(d * 2j * np.pi * sum(k[a_] * s[a_] / shape[a_] for a_ in range(3)))
    sh = np.linalg.norm((Y - X * ph).ravel()) / np.linalg.norm(Y.ravel())
    assert sh <= 1e-13, f"shift theorem {sh:.2e}"
    print(f"{shape} B={batch} dir={d} {p.kernel}: linearity {lin:.1e} shift {sh:.1e}", flush=True)
    p.close()


def test_properties_complex64(gpu):
    f = gpu.fft3d
    shape, batch = (17, 17, 17), 2048
    V = 17 ** 3
    p = f.Plan(shape, batch, -1, f.C64)
    x = _rand(shape, batch, seed=9, dtype=np.complex64)
    z = _rand(shape, batch, seed=10, dtype=np.complex64)
    X, Z, W = (p.run(v).astype(np.complex128) for v in (x, z, (x + z).astype(np.complex64)))
    lin = np.linalg.norm((W - X - Z).ravel()) / np.linalg.norm(W.ravel())
    assert lin <= 2e-6, lin
    ex = np.sum(np.abs(X) ** 2, axis=(1, 2, 3))
    ey = V * np.sum(np.abs(x.astype(np.complex128)) ** 2, axis=(1, 2, 3))
    assert np.max(np.abs(ex - ey) / ey) <= 4e-6
    d = np.zeros((batch,) + shape, np.complex64)
    d[:, 0, 0, 0] = 1.0
    assert np.max(np.abs(p.run(d) - 1.0)) <= 1e-7
    p.close()
