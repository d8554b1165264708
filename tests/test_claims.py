"""Claim-counter assignment of plans (opt-fft_amd/csrc/claims.hpp) on the CPU, over a fake
runtime: one stream never pays for events; from the second stream on every slot reuse waits
for the launch that last held it (the 65th concurrent launch of a plan waits for the 1st);
captured launches each keep a counter and the 65th capture is refused."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_claim_slots(tmp_path):
    exe = tmp_path / "claims_check"
    subprocess.run(["g++", "-O1", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "opt-fft_amd", "csrc"),
                    "-o", str(exe), os.path.join(ROOT, "tests", "claims", "claims_check.cpp"),
                    "-lpthread"], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout.strip() == "OK", r.stdout + r.stderr
