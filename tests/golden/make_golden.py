"""Write tests/golden/golden.npz: seeded inputs and their 3D DFTs from the long-double oracle.

The reference holds no golden vector files (inputs are regenerated from seeds there too),
so these fixtures are this repository's own.  Each expected output comes from the
long-double dense DFT and is accepted only if numpy.fft agrees with it to 1e-14.  Small
geometries store input and output in full; L = 36, 24 and 24x24x48 store the input seed,
the whole-batch sums sum(X) and sum|X|^2, and 64 sampled output entries.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle  # noqa: E402

FULL = [  # (name, shape, batch, seed, dist, sign, normalize, dtype)
    ("L6_fwd", (6, 6, 6), 2, 61, "uniform", -1, 0, "c128"),
    ("L8_fwd", (8, 8, 8), 2, 81, "uniform", -1, 0, "c128"),
    ("L8_bwd_norm", (8, 8, 8), 1, 82, "uniform", 1, 1, "c128"),
    ("L12_fwd", (12, 12, 12), 1, 121, "uniform", -1, 0, "c128"),
    ("L17_fwd", (17, 17, 17), 1, 171, "uniform", -1, 0, "c128"),
    ("L17_bwd", (17, 17, 17), 1, 172, "uniform", 1, 0, "c128"),
    ("8x8x16_fwd", (8, 8, 16), 1, 816, "uniform", -1, 0, "c128"),
    ("L17_fwd_c64", (17, 17, 17), 1, 173, "uniform", -1, 0, "c64"),
    ("L8_bwd_c64", (8, 8, 8), 2, 83, "uniform", 1, 0, "c64"),
]
SAMPLED = [
    ("L36_fwd", (36, 36, 36), 1, 361, "normal_half", -1, 0, "c128"),
    ("L24_fwd", (24, 24, 24), 1, 241, "normal_half", -1, 0, "c128"),
    ("24x24x48_fwd", (24, 24, 48), 1, 2448, "normal_half", -1, 0, "c128"),
]


def main():
    out = {}
    for name, shape, batch, seed, dist, sign, norm, dt in FULL + SAMPLED:
        dtype = np.complex128 if dt == "c128" else np.complex64
        x = oracle.make_input(shape, batch, seed, dist, dtype)
        y = oracle.dft3d(x, sign=sign, normalize=bool(norm))
        chk = oracle.rel_l2(oracle.fftn_numpy(x, sign=sign, normalize=bool(norm)), y)
        assert chk < 1e-14, (name, chk)
        meta = np.array([shape[0], shape[1], shape[2], batch, seed, sign, norm,
                         0 if dt == "c128" else 1], dtype=np.int64)
        out[name + "/meta"] = meta
        out[name + "/dist"] = np.array(dist)
        if (name, shape, batch, seed, dist, sign, norm, dt) in FULL:
            out[name + "/in"] = x
            out[name + "/out"] = y
        else:
            rng = np.random.default_rng(seed + 7)
            idx = rng.choice(y.size, 64, replace=False)
            out[name + "/idx"] = idx
            out[name + "/vals"] = y.reshape(-1)[idx]
            out[name + "/sum"] = np.array(y.sum())
            out[name + "/energy"] = np.array(float(np.sum(np.abs(y) ** 2)))
        print(f"{name:14s} numpy vs long double {chk:.2e}")
    np.savez_compressed(os.path.join(HERE, "golden.npz"), **out)


if __name__ == "__main__":
    main()
