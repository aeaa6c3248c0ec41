"""The stream-collide's wave-to-node division (lbm::FastDiv, a multiply-high
and a shift with the multiplier precomputed on the host) against the
processor's own integer division, by brute force over divisors and
dividends (tests/fastdiv_check.cpp, built with g++ from the header alone)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "lbm_multigpu_amd", "csrc")


def test_fastdiv_matches_integer_division(tmp_path):
    exe = tmp_path / "fastdiv_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", CSRC, "-o", str(exe),
                    os.path.join(HERE, "fastdiv_check.cpp")], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok ")
