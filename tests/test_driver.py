"""The C++ host driver lbm_run (the reference programs' main(),
MRTtest/DynInlet3.cu:2834-3335): it reads the same `.inputs` records as the
Python layer (same keys, separators and defaults) and writes the reference's
CASENAME.dat / .force / .inputs files."""
import os
import subprocess

import numpy as np
import pytest

import lbm_multigpu_amd as L
from lbm_multigpu_amd import config as C

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "lbm_multigpu_amd", "lbm_run")

# a run record as the reference's WriteInputs prints it (keys, separators and
# the "omega \t: v" form), no extension keys
REFERENCE_RECORD = """Base domain size \t402x66x66
Base blocksize: \t64x1x1
Obst1 location: \t(100.5,32.5,32.5)
Obst1 radius: \t8
Obst2 location: \t(0,0,0)
Obst2 radius: \t0
RE: \t100
UMAX: \t0.06
omega \t: 1.85185
DPDY \t: 0
TMAX: \t40000
STARTF: \t20000
START_VELAV: \t200000
START_VELFLUC: \t1600000
REFINEMENT: \t0
MODEL: \tMRT
Smagorinsky LES: \t1
CS: \t0.01
LR domain size \t128x128x64
LR factor \t0.5
LR location \t100.25x16.25x16.25
LR blocksize: \t64x1x1
omega in LR \t: 1.72414
GPUs per node \t: 1
"""


def _driver(*args, cwd=None):
    if not os.path.exists(DRIVER):
        pytest.fail("lbm_run not built: run __graft_entry__.build()")
    return subprocess.run([DRIVER, *args], capture_output=True, text=True, cwd=cwd, timeout=300)


@pytest.mark.parametrize("case", ["channel", "refined", "cpu3d", "reference"])
def test_driver_parses_inputs_like_the_python_layer(case, tmp_path):
    path = tmp_path / "case.inputs"
    if case == "reference":
        path.write_text(REFERENCE_RECORD)
    else:
        cfg = {"channel": C.channel(force_steps=100), "refined": C.refined_channel(),
               "cpu3d": C.cpu3d_case()}[case]
        C.write_inputs(cfg, str(path))
    r = _driver(str(path), "--parse-only")
    assert r.returncode == 0, r.stderr
    want = C.write_inputs(C.read_inputs(str(path)))
    assert r.stdout == want


def test_driver_usage_errors(tmp_path):
    assert _driver().returncode == 2
    r = _driver(str(tmp_path / "missing.inputs"), "--parse-only")
    assert r.returncode == 1 and "cannot read" in r.stderr


@pytest.mark.gpu
def test_driver_run_matches_solver(tmp_path):
    """A small channel run through lbm_run and through the Python Solver:
    the same fields, statistics and forces."""
    cfg = C.channel(nx=96, ny=32, nz=24, side=8, x_center=32, tau=0.52, force_steps=40)
    cfg.tmax, cfg.startf = 60, 20
    cfg.velav, cfg.start_velav, cfg.start_velfluc = True, 10, 40
    path = tmp_path / "drv.inputs"
    C.write_inputs(cfg, str(path))
    print("driver run", flush=True)
    r = _driver(str(path), cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr
    assert (tmp_path / "drv.inputs").read_text() == C.write_inputs(cfg)
    dat = np.loadtxt(tmp_path / "drv.dat", delimiter=",", skiprows=2)
    assert dat.shape == (96 * 32 * 24, 14)
    # the slice file is the z = 1 plane of the same columns (WriteResults' outputslice)
    sl = np.loadtxt(tmp_path / "drv_slice.dat", delimiter=",", skiprows=2)
    assert np.array_equal(sl, dat[96 * 32:2 * 96 * 32])
    # the control window (t > 10000) is not reached: an empty CASENAME.vel
    assert (tmp_path / "drv.vel").read_text() == ""
    force = np.loadtxt(tmp_path / "drv.force", delimiter=",")
    with L.Solver(C.read_inputs(str(path))) as s:
        s.set_geometry()
        s.init_equilibrium(L.INIT_MRT)
        s.step(60)
        rho, u, v, w = s.moments()
        mean, dev = s.stats()
        F = s.forces()
    scale = np.abs(u).max()
    for col, ref in ((3, u), (4, v), (5, w)):
        assert np.abs(dat[:, col] - ref.reshape(-1)).max() < 1e-5 * scale
    assert np.abs(dat[:, 6] - rho.reshape(-1)).max() < 1e-5
    for a in range(3):
        assert np.abs(dat[:, 7 + a] - mean[a].reshape(-1)).max() < 1e-5 * scale
        assert np.abs(dat[:, 10 + a] - dev[a].reshape(-1)).max() <= 1e-5 * max(np.abs(dev).max(), 1e-30)
    assert np.all(np.isfinite(dat[:, 13])) and dat[:, 13].max() > 0
    ref = np.float32(cfg.umax) ** 2 * cfg.nz * np.float32(cfg.obst1_r)
    assert force.shape == (40, 4)
    assert np.array_equal(force[:, 0], np.arange(20, 60))
    assert np.abs(force[:, 1:] - F / ref).max() < 1e-5 * np.abs(F / ref).max()
