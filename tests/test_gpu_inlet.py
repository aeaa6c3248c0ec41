"""Dynamic-inlet precursor pieces of MRTtest/DynInlet3.cu against the
oracle: the body force limited to rows y <= DYNY1 (:1744), AverageV over the
plane y = DYNY1 (:2561-2600) and the host P-controller on dpdy every
ctrl_every steps (:3147-3170).  Tolerances: single step 2e-6 of max|f| (the
FP32 single-step bar); control history: plane means within 1e-6 absolute
and dpdy within 1e-5 relative over 8 control events."""
import sys

import numpy as np
import pytest

import lbm_multigpu_amd as L
import oracle_lib as O
from test_gpu_parity import orc_params, random_state

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def box(dyn_y1, dpdy=1e-4, kp=0.0, every=100, start=10000):
    c = L.default_config()
    c.nx, c.ny, c.nz = 40, 24, 8
    c.model, c.omega, c.cs = L.MODEL_MRT_LES, 1.6, 0.01
    c.dpdy, c.umax = dpdy, 0.05
    c.dyn_y1, c.kp, c.ctrl_every, c.ctrl_start = dyn_y1, kp, every, start
    return c


def orc(c):
    p = orc_params(c)
    p.dyn_y1 = c.dyn_y1
    return p


def test_row_limited_force_matches_oracle():
    say("row-limited force")
    c = box(dyn_y1=11)
    f = random_state(np.random.default_rng(21), (c.nz, c.ny, c.nx))
    with L.Solver(c) as s:
        s.set_populations(f)
        s.step(1)
        g = s.populations()
    ref, _ = O.step(orc(c), f)
    assert np.abs(g - ref).max() < 2e-6 * np.abs(ref).max()
    c2 = box(dyn_y1=-1)
    full, _ = O.step(orc(c2), f)
    # the limit acts: rows past 11 differ from the everywhere-forced step
    assert np.abs(g[:, :, 12:] - full[:, :, 12:]).max() > 1e-6


def test_inlet_control_matches_oracle():
    """Periodic box driven along y, DYNY1 = 16, KP = 0.5, control every 5
    steps from t > 0: 40 steps, the control history and the final state
    against the oracle stepping with the same controller."""
    say("inlet P-control")
    c = box(dyn_y1=16, dpdy=2e-5, kp=0.5, every=5, start=0)
    f0 = random_state(np.random.default_rng(22), (c.nz, c.ny, c.nx), umax=0.02)
    T = 40
    with L.Solver(c) as s:
        s.set_populations(f0)
        s.step(T)
        hist = s.inlet_control()
        g = s.populations()
        d_gpu = s.dpdy()
        sm, mean = s.plane_mean_v(16)
    p = orc(c)
    f = f0.copy()
    dpdy = np.float32(c.dpdy)
    ref = []
    for t in range(T):
        p.dpdy = float(dpdy)
        f, _ = O.step(p, f)
        if t % 5 == 0 and t > 0:
            av, dpdy = O.inlet_control_update(dpdy, O.plane_v_sum(f, 16), c.nx, c.nz, c.umax, c.kp, c.dpdy)
            ref.append((t, av, dpdy))
    ref = np.array(ref)
    say(f"  {len(hist)} control events; dpdy {c.dpdy:.3e} -> {d_gpu:.4e} (oracle {ref[-1, 2]:.4e})")
    assert hist.shape == ref.shape == (7, 3)
    assert np.array_equal(hist[:, 0], ref[:, 0])
    assert np.abs(hist[:, 1] - ref[:, 1]).max() < 1e-6
    assert np.abs(hist[:, 2] - ref[:, 2]).max() < 1e-5 * np.abs(ref[:, 2]).max()
    assert abs(ref[-1, 2] - c.dpdy) > 1e-3 * c.dpdy  # the controller moved dpdy
    assert np.abs(g - f).max() < 1e-5 * np.abs(f).max()
    assert abs(sm - O.plane_v_sum(f, 16)) < 1e-5 * max(abs(sm), 1e-30) + 1e-9


def test_control_is_off_by_default_window():
    """The reference's window (every 100 steps after step 10000): a short run
    records nothing and keeps DPDY."""
    c = box(dyn_y1=-1, dpdy=1e-5, kp=0.5)
    with L.Solver(c) as s:
        s.init_equilibrium(L.INIT_BGK)
        s.step(30)
        assert s.inlet_control().shape == (0, 3)
        assert s.dpdy() == np.float32(1e-5)
