"""The reference's own CPU kernels (oracle/_ref, built from the reference sources by
oracle/build_ref.sh) against the oracle on the same seeded input: pins the oracle to the
real reference implementation wherever it builds."""
import os
import subprocess

import numpy as np
import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def ref_bin(kind):
    isa = "v4" if "avx512f" in open("/proc/cpuinfo").read() else "v3"
    p = os.path.join(ROOT, "oracle", "_ref", isa, f"ref_bench_{kind}")
    if not os.path.exists(p):
        pytest.skip("reference CPU kernels not built (oracle/build_ref.sh needs the reference)")
    return p


@pytest.mark.parametrize("L,batch", [(6, 3), (8, 3), (17, 2), (36, 1)])
@pytest.mark.parametrize("kind", ["st", "mt"])
def test_reference_kernel_agrees_with_oracle(tmp_path, L, batch, kind):
    exe = ref_bin(kind)
    x = oracle.make_input((L, L, L), batch, seed=L * 1000 + batch, dist="normal")
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    x.tofile(fin)
    env = dict(os.environ, OMP_NUM_THREADS="2")
    r = subprocess.run([exe, "--L", str(L), "--batch", str(batch), "--seconds", "0.01",
                        "--in", str(fin), "--out", str(fout)], capture_output=True, text=True,
                       env=env, timeout=300)
    assert r.returncode == 0, r.stderr
    y = np.fromfile(fout, dtype=np.complex128).reshape(x.shape)
    assert oracle.rel_l2(y, oracle.dft3d(x)) < 1e-13
    assert oracle.rel_l2(y, oracle.fftn_numpy(x)) < 1e-13
