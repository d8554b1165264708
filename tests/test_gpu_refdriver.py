"""The drop-in boundary as the reference's own harness sees it: the reference's CPU timing
driver (bench/geom/driver.c, unchanged) linked against libfft3d_mi355x.so in place of a CPU
kernel (oracle/_ref/geom_driver_mi355x, built by oracle/build_ref.sh the way
bench/geom/Makefile:47-48 links a backend).  The driver reads the raw complex128 input file,
times calibrated samples, writes one execute's output (--out) and the unitary chain's end
state (<out>.chain), and exits 5 if the output does not follow the input; each result is
checked here against the oracle (bench/geom/check.py:28-34, :43-61)."""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "_ref", "geom_driver_mi355x")

pytestmark = pytest.mark.gpu


def driver(args, timeout=240):
    if not os.path.exists(EXE):
        pytest.skip("reference driver not built (oracle/build_ref.sh needs the reference "
                    "sources at build time)")
    print("geom_driver_mi355x", " ".join(args), flush=True)
    return subprocess.run([EXE, *args], capture_output=True, text=True, timeout=timeout)


@pytest.mark.parametrize("L,batch", [(8, 64), (17, 16), (36, 4), (6, 33), (12, 9), (24, 2)])
def test_reference_driver_output_matches_oracle(gpu, tmp_path, L, batch):
    x = oracle.make_input((L, L, L), batch, seed=L * 1000 + batch, dist="normal")
    fin, fout, js = tmp_path / "in.bin", tmp_path / "out.bin", tmp_path / "r.json"
    x.tofile(fin)
    r = driver(["--L", str(L), "--batch", str(batch), "--in", str(fin), "--out", str(fout),
                "--json", str(js), "--samples", "3", "--warmup", "1", "--min-sample-ms", "2"])
    assert r.returncode == 0, r.stdout + r.stderr
    y = np.fromfile(fout, dtype=np.complex128).reshape(x.shape)
    assert oracle.rel_l2(y, oracle.fftn_numpy(x)) <= 1e-13
    rec = json.loads(js.read_text())
    assert rec["name"] == "mi355x_hip" and rec["supported"] and rec["batch"] == batch
    assert rec["per_execute_seconds"]["min"] > 0


@pytest.mark.parametrize("L,batch,m", [(8, 8, 4), (17, 3, 2), (36, 1, 3)])
def test_reference_driver_unitary_chain_closed_form(gpu, tmp_path, L, batch, m):
    x = oracle.make_input((L, L, L), batch, seed=7 + L, dist="uniform")
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    x.tofile(fin)
    r = driver(["--L", str(L), "--batch", str(batch), "--in", str(fin), "--out", str(fout),
                "--samples", "2", "--warmup", "1", "--min-sample-ms", "1", "--chain", str(m),
                "--unitary"])
    assert r.returncode == 0, r.stdout + r.stderr
    end = np.fromfile(str(fout) + ".chain", dtype=np.complex128).reshape(x.shape)
    assert oracle.rel_l2(end, oracle.chain_expected(x, m)) <= 1e-13


def test_reference_driver_unsupported_size_exits_3(gpu, tmp_path):
    x = oracle.make_input((7, 7, 7), 1, seed=1)
    fin, js = tmp_path / "in.bin", tmp_path / "r.json"
    x.tofile(fin)
    r = driver(["--L", "7", "--batch", "1", "--in", str(fin), "--json", str(js)])
    assert r.returncode == 3
    assert json.loads(js.read_text())["supported"] is False
