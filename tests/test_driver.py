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
    assert np.all(force[0, 1:] == 0)  # step STARTF itself is not recorded (t > STARTF)


POST = os.path.join(ROOT, "lbm_multigpu_amd", "lbm_post")


@pytest.mark.gpu
def test_precursor_runs_end_to_end_through_lbm_run_and_lbm_post(tmp_path):
    """A small bgk_channel1/Re3000_Cyl2-style case: the periodic precursor
    rows 0..DYNY1 under the P-controlled body force, the im 80 recycling
    plane, a trip obstacle removed at step 100, the force-summed cylinder and
    the north outlet; lbm_run writes .dat/.force/.vel/_slice.dat equal to the
    Solver's run, and lbm_post profiles (VelProf2) reads the slice."""
    cfg = C.precursor(nx=42, ny=96, nz=4, dyn_y1=24, side=4, obst_y=44, trip=(14, 8, 3), trip_until=100,
                      ctrl=(50, 100))
    cfg.tmax, cfg.startf, cfg.force_steps = 400, 200, 200
    cfg.velav, cfg.start_velav, cfg.start_velfluc = True, 100, 250
    path = tmp_path / "pre.inputs"
    C.write_inputs(cfg, str(path))
    print("precursor driver run", flush=True)
    r = _driver(str(path), cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr
    dat = np.loadtxt(tmp_path / "pre.dat", delimiter=",", skiprows=2)
    force = np.loadtxt(tmp_path / "pre.force", delimiter=",")
    with L.Solver(C.read_inputs(str(path))) as s:
        s.set_geometry()
        s.init_equilibrium(L.INIT_MRT)
        s.step(400)
        rho, u, v, w = s.moments()
        F = s.forces()
        ctrl = s.inlet_control()
        m = s.get_mask()
    assert (m == L.NODE_RECYCLE).any() and not (m[:, 5:12, 10:18] == L.NODE_SOLID).any()  # trip gone
    scale = np.abs(v).max()
    assert np.abs(dat[:, 4] - v.reshape(-1)).max() < 1e-5 * scale
    assert np.abs(dat[:, 6] - rho.reshape(-1)).max() < 1e-5
    ref = np.float32(cfg.umax) ** 2 * cfg.nz * np.float32(cfg.obst1_r)
    assert force.shape == (200, 4) and np.array_equal(force[:, 0], np.arange(200, 400))
    assert np.abs(force[:, 1:] - F / ref).max() < 1e-5 * np.abs(F / ref).max()
    # control events at t = 150, 200, ... (t % 50 == 0, t > 100), the body force moved
    assert len(ctrl) == 5 and ctrl[0, 0] == 150 and ctrl[-1, 2] != cfg.dpdy
    # the flow runs along y through the recycling plane into the channel
    assert v[:, cfg.dyn_y1 + 2:-1, 1:-1].mean() > 0.5 * cfg.umax
    prof = tmp_path / "pre.prof2"
    p = subprocess.run([POST, "profiles", str(tmp_path / "pre_slice.dat"), "--nx", "42", "--ny", "96",
                        "--obst-y", str(cfg.obst1[1]), "--height", "40", "--diam", "4", "--umax", str(cfg.umax),
                        "--out", str(prof)], capture_output=True, text=True, timeout=60)
    assert p.returncode == 0, p.stderr
    lines = prof.read_text().splitlines()
    assert lines[0] == "x=1d" and "wake profile" in lines
    vals = np.array([[float(a) for a in ln.split(",")] for ln in lines if "," in ln])
    assert vals.shape[1] == 5 and np.all(np.isfinite(vals))


@pytest.mark.gpu
def test_refined_driver_normalises_the_fine_force_by_lrlevel_squared(tmp_path):
    """With refinement the .force file holds the fine level's force over
    UMAX^2 ZDIM OBSTR1 LRLEVEL^2 (WriteForces(..., REFINEMENT*LRLEVEL),
    MRTtest/DynInlet3.cu:2788-2796, :3300)."""
    cfg = C.channel(nx=64, ny=32, nz=32, side=6, x_center=22, tau=0.55, rates="lrmgpu", force_steps=20)
    cfg.refinement, cfg.lr_size, cfg.lr_loc = True, (32, 24, 24), (12.25, 4.25, 4.25)
    cfg.tmax, cfg.startf = 20, 0
    path = tmp_path / "lr.inputs"
    C.write_inputs(cfg, str(path))
    r = _driver(str(path), cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr
    force = np.loadtxt(tmp_path / "lr.force", delimiter=",")
    with L.Solver(C.read_inputs(str(path))) as s:
        s.set_geometry()
        s.init_equilibrium(L.INIT_MRT)
        s.step(20)
        F = s.forces()
    ref = np.float32(cfg.umax) ** 2 * cfg.nz * np.float32(cfg.obst1_r) * 4
    assert np.abs(F).max() > 0
    assert np.abs(force[:, 1:] - F / ref).max() < 1e-5 * np.abs(F / ref).max()
