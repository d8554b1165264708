"""The reference's own CPU kernels, each linked with the reference's own timing driver
(oracle/_ref, built from the reference sources by oracle/build_ref.sh), against the oracle on
the same seeded input: pins the oracle to the real reference implementation wherever it
builds, through the reference's file format (raw complex128, bench/geom/driver.c:116-129 in,
:209-217 out)."""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ISA = "v4" if "avx512f" in open("/proc/cpuinfo").read() else "v3"
ST = ["L6_pfa", "L6_unrolled", "L8_batchsimd", "L8_fusedaxes", "L17_matrixsimd", "L17_rader",
      "L36_mixedradix", "L36_pfa", "L36_pencilfused"]
MT = ["L6_pfa", "L8_fusedaxes", "L8_batchsimd", "L17_winograd", "L17_matrixsimd",
      "L36_mixedradix", "L36_pfa"]
BATCH = {6: 3, 8: 3, 17: 2, 36: 1}


def ref_bin(*parts):
    p = os.path.join(ROOT, "oracle", "_ref", *parts)
    if not os.path.exists(p):
        pytest.skip("reference builds absent (oracle/build_ref.sh needs the reference sources)")
    return p


def run_driver(exe, L, batch, x, tmp_path, env=None, extra=()):
    fin, fout, js = tmp_path / "in.bin", tmp_path / "out.bin", tmp_path / "r.json"
    x.tofile(fin)
    r = subprocess.run([exe, "--L", str(L), "--batch", str(batch), "--in", str(fin),
                        "--out", str(fout), "--json", str(js), "--samples", "2", "--warmup", "1",
                        "--min-sample-ms", "1", *extra],
                       capture_output=True, text=True, env=env, timeout=300)
    return r, fout, js


@pytest.mark.parametrize("kind,kernel", [("st", k) for k in ST] + [("mt", k) for k in MT])
def test_reference_kernel_agrees_with_oracle(tmp_path, kind, kernel):
    exe = ref_bin(ISA, kind, kernel)
    L = int(kernel[1:].split("_")[0])
    batch = BATCH[L]
    x = oracle.make_input((L, L, L), batch, seed=L * 1000 + batch, dist="normal")
    env = dict(os.environ, OMP_NUM_THREADS="2")
    r, fout, js = run_driver(exe, L, batch, x, tmp_path, env=env)
    assert r.returncode == 0, r.stderr
    y = np.fromfile(fout, dtype=np.complex128).reshape(x.shape)
    assert oracle.rel_l2(y, oracle.dft3d(x)) < 1e-13
    assert oracle.rel_l2(y, oracle.fftn_numpy(x)) < 1e-13
    rec = json.loads(js.read_text())
    assert rec["supported"] and rec["L"] == L and rec["batch"] == batch


def test_reference_driver_on_this_library_unsupported_and_usage(tmp_path):
    """The reference's driver linked against libfft3d_mi355x.so: an unsupported L is exit 3
    with supported:false (bench/geom/driver.c:99-111), bad usage exit 2 -- neither touches a
    device."""
    exe = ref_bin("geom_driver_mi355x")
    x = oracle.make_input((7, 7, 7), 1, seed=1)
    r, _, js = run_driver(exe, 7, 1, x, tmp_path)
    assert r.returncode == 3
    rec = json.loads(js.read_text())
    assert rec == {"name": "mi355x_hip", "L": 7, "batch": 1, "supported": False}
    r = subprocess.run([exe, "--batch", "1"], capture_output=True)
    assert r.returncode == 2
