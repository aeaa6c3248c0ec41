"""lbm_post, the reference's post-processing programs over this build's output
files (bgk_channel1/StaticInlet/ForceCalcs.cpp, VelProf2.cpp): force mean and
RMS over a window of a .force record, and the cross-stream and wake profiles
from a statistics slice file.  Host only.  Checked against closed forms: the
running mean/variance recurrences equal the window's mean and population
variance; the profiles sample known fields at the documented stations."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
POST = os.path.join(ROOT, "lbm_multigpu_amd", "lbm_post")


def _post(*args, cwd=None):
    if not os.path.exists(POST):
        pytest.fail("lbm_post not built: run __graft_entry__.build()")
    return subprocess.run([POST, *args], capture_output=True, text=True, cwd=cwd, timeout=60)


def _parse_force(out):
    vals = {}
    for line in out.strip().splitlines():
        a, b = line.split(",")
        k1, v1 = a.split("=")
        k2, v2 = b.split("=")
        vals[k1.strip()] = float(v1)
        vals[k2.strip()] = float(v2)
    return vals


def test_force_mean_and_rms_over_a_window(tmp_path):
    rng = np.random.default_rng(0)
    n = 500
    F = np.stack([1.5 + 0.2 * rng.standard_normal(n), 0.3 * np.sin(np.arange(n) / 7.0), np.full(n, -0.25)], 1)
    F = F.astype(np.float32)
    path = tmp_path / "c.force"
    with open(path, "w") as fh:
        for t in range(n):
            fh.write(f"{t + 1000}, {float(F[t, 0]):.9g}, {float(F[t, 1]):.9g}, {float(F[t, 2]):.9g}\n")
    for lo, hi in ((0, n), (100, 400)):
        r = _post("force", str(path), "--from", str(lo), "--to", str(hi))
        assert r.returncode == 0, r.stderr
        got = _parse_force(r.stdout)
        w = F[lo:hi].astype(np.float64)
        for a, name in enumerate(("Fx", "Fy", "Fz")):
            assert abs(got[name + "Av"] - w[:, a].mean()) <= 1e-5 * max(abs(w[:, a]).max(), 1e-30)
            assert abs(got[name + "RMS"] - w[:, a].std()) <= 1e-4 * max(w[:, a].std(), 1e-3)
    assert _post("force", str(path), "--from", "600").returncode == 1


def test_profiles_sample_the_slice_at_the_stations(tmp_path):
    """A slice (X = 24, Y = 120) whose statistics columns are known functions
    of (x, y): <u> = 0.05 (1 + x/100 + y/1000), <v> = 0.01 x y/1000,
    ufluc = (1e-3 (x + 1))^2, vfluc = (2e-3 (y + 1))^2."""
    X, Y, U, oby, d, H = 24, 120, 0.05, 20.5, 8.0, 40
    path = tmp_path / "case_slice.dat"
    with open(path, "w") as fh:
        fh.write('VARIABLES = "X","Y","Z","u","v","w","rho","velAv[0]","velAv[1]","velAv[2]","ufluc","vfluc","wfluc","Smag"\n')
        fh.write(f"ZONE F=POINT, I={X}, J={Y}, K=1\n")
        for y in range(Y):
            for x in range(X):
                ua = 0.05 * (1 + x / 100 + y / 1000)
                va = 0.01 * x * y / 1000
                uf = (1e-3 * (x + 1)) ** 2
                vf = (2e-3 * (y + 1)) ** 2
                fh.write(f"{x}, {y}, 1, 0,0,0,1,{ua!r},{va!r},0,{uf!r},{vf!r},0,0\n")
    r = _post("profiles", str(path), "--nx", str(X), "--ny", str(Y), "--obst-y", str(oby), "--height", str(H),
              "--diam", str(d), "--umax", str(U))
    assert r.returncode == 0, r.stderr
    text = (tmp_path / "case.prof2").read_text().strip().splitlines()
    blocks, cur = {}, None
    for line in text:
        if "," not in line:
            cur = line
            blocks[cur] = []
        else:
            blocks[cur].append([float(v) for v in line.split(",")])
    assert list(blocks) == ["x=1d", "x=3.5d", "x=6d", "x=8.5d", "wake profile"]
    for label, k in (("x=1d", 1.0), ("x=3.5d", 3.5), ("x=6d", 6.0), ("x=8.5d", 8.5)):
        j = int(np.float32(oby) + np.float32(d) * np.float32(0.5) + 1 + np.float32(k) * np.float32(d))
        rows = np.array(blocks[label])
        assert rows.shape == (H // 2 + 1, 5)
        x = np.arange(H // 2 + 1)
        assert np.allclose(rows[:, 0], x / d)
        assert np.allclose(rows[:, 1], 0.05 * (1 + x / 100 + j / 1000) / U, rtol=1e-5)
        assert np.allclose(rows[:, 2], 0.01 * x * j / 1000 / U, rtol=1e-5, atol=1e-7)
        assert np.allclose(rows[:, 3], 1e-3 * (x + 1) / U, rtol=1e-5)
        assert np.allclose(rows[:, 4], 2e-3 * (j + 1) / U, rtol=1e-5)
    wake = np.array(blocks["wake profile"])
    j0 = int(oby + d * 0.5 + 1)
    js = np.arange(j0, int(np.ceil(oby + d * 0.5 + 1 + 8 * d)))
    assert wake.shape == (len(js), 5)
    assert np.allclose(wake[:, 0], (js - oby - d * 0.5) / d, rtol=1e-6)
    assert np.allclose(wake[:, 1], 0.05 * (1 + (H // 2) / 100 + js / 1000) / U, rtol=1e-5)


def test_profiles_reject_a_slice_without_statistics(tmp_path):
    path = tmp_path / "s_slice.dat"
    path.write_text('VARIABLES = "X","Y","Z","u","v","w","rho"\nZONE F=POINT, I=30, J=2, K=1\n'
                    + "".join(f"{x}, {y}, 1, 0,0,0,1\n" for y in range(2) for x in range(30)))
    r = _post("profiles", str(path), "--nx", "30", "--ny", "2", "--obst-y", "0", "--height", "20", "--diam", "1",
              "--umax", "0.1")
    assert r.returncode == 1 and "statistics" in r.stderr
