"""Golden fixtures (tests/golden/golden.npz, written by tests/golden/make_golden.py).

CPU: the oracle reproduces every stored output, and the seeded generator reproduces every
stored input.  GPU: the product's result for each fixture matches the stored output."""
import os

import numpy as np
import pytest

import oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden.npz")


def fixtures():
    z = np.load(GOLD, allow_pickle=False)
    names = sorted({k.split("/")[0] for k in z.files})
    out = []
    for n in names:
        m = z[n + "/meta"]
        d = dict(name=n, shape=tuple(int(v) for v in m[:3]), batch=int(m[3]), seed=int(m[4]),
                 sign=int(m[5]), normalize=bool(m[6]), prec=int(m[7]), dist=str(z[n + "/dist"]))
        for k in ("in", "out", "idx", "vals", "sum", "energy"):
            if f"{n}/{k}" in z.files:
                d[k] = z[f"{n}/{k}"]
        out.append(d)
    return out


FIX = fixtures()


def _input(f):
    if "in" in f:
        return f["in"]
    dt = np.complex128 if f["prec"] == 0 else np.complex64
    return oracle.make_input(f["shape"], f["batch"], f["seed"], f["dist"], dt)


def _check(f, y, tol):
    if "out" in f:
        assert oracle.rel_l2(y, f["out"]) <= tol, f["name"]
    else:
        flat = y.reshape(-1)
        assert oracle.rel_l2(flat[f["idx"]], f["vals"]) <= tol, f["name"]
        assert abs(y.sum() - f["sum"]) <= tol * np.sqrt(f["energy"]) * 10, f["name"]
        assert abs(np.sum(np.abs(y) ** 2) - f["energy"]) <= 10 * tol * f["energy"], f["name"]


@pytest.mark.parametrize("f", FIX, ids=[f["name"] for f in FIX])
def test_oracle_reproduces_golden(f):
    x = _input(f)
    if "in" in f:
        dt = np.complex128 if f["prec"] == 0 else np.complex64
        regen = oracle.make_input(f["shape"], f["batch"], f["seed"], f["dist"], dt)
        assert np.array_equal(regen, f["in"]), "seeded generator drifted"
    _check(f, oracle.dft3d(x, sign=f["sign"], normalize=f["normalize"]), 1e-15)
    _check(f, oracle.fftn_numpy(x, sign=f["sign"], normalize=f["normalize"]), 1e-14)


@pytest.mark.gpu
@pytest.mark.parametrize("f", FIX, ids=[f["name"] for f in FIX])
def test_gpu_matches_golden(gpu, f):
    fx = gpu.fft3d
    x = _input(f)
    plan = fx.Plan(f["shape"], f["batch"], f["sign"], f["prec"],
                   flags=fx.NORMALIZE if f["normalize"] else 0)
    y = plan.run(x)
    plan.close()
    _check(f, y, 1e-13 if f["prec"] == 0 else 2e-6)
