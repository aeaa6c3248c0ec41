"""Running velocity statistics (the reference's VELAV vel_av / vel_fluc,
MRTtest/DynInlet3.cu:1139-1184, :1773-1799) fused into the HIP
stream-collide, against the oracle's restatement applied after each oracle
step on the same inputs.  Tolerance: the populations of the two FP32 paths
agree to ~1e-7 per step, so the means agree within 1e-5 of their scale and
the deviations within 1e-3 of theirs."""
import os
import sys

import numpy as np
import pytest

import lbm_multigpu_amd as L
import oracle_lib as O
from lbm_multigpu_amd import config as cfgmod
from test_gpu_parity import NODE_TO_ORC, orc_params

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_channel_statistics_match_oracle():
    say("statistics channel")
    cc = cfgmod.channel(nx=64, ny=24, nz=16, side=6, x_center=20, tau=0.52)
    c = cc.to_c()
    T, A, FL = 40, 10, 28
    with L.Solver(c) as s:
        s.set_geometry(None)
        mask = np.vectorize(NODE_TO_ORC.get)(s.get_mask()).astype(np.uint8)
        s.init_equilibrium(L.INIT_BGK)
        f0 = s.populations()
        s.stats_enable(A, FL)
        s.step(T)
        mean, dev = s.stats()
    p = orc_params(c)
    u_in = O.inlet_profile(c.ny, c.nz, c.umax, c.inlet_amp, c.seed)
    stat = np.zeros((6,) + mask.shape, np.float32)
    f = f0
    for t in range(T):
        f, _ = O.step(p, f, mask=mask, u_in=u_in)
        if t >= FL:
            O.stats_update(p, f, stat, 2, float(t - FL), 1.0, mask=mask)
        elif t >= A:
            O.stats_update(p, f, stat, 1, float(t - A), 1.0, mask=mask)
    em = np.abs(mean - stat[:3]).max() / np.abs(stat[:3]).max()
    ed = np.abs(dev - stat[3:]).max() / np.abs(stat[3:]).max()
    say("  mean err", em, "dev err", ed, "dev scale", np.abs(stat[3:]).max())
    assert em < 1e-5 and ed < 1e-3
    solid = np.isin(mask, [O.SOLID, O.SOLID_FORCE])
    assert np.all(mean[:, solid] == 0) and np.all(dev[:, solid] == 0)
    assert np.abs(stat[3:]).max() > 0


def test_statistics_phases_and_fine_level():
    """Uniform flow: the mean is the flow's momentum on both levels (the
    fine level's LRFACTOR weighting keeps a constant sequence constant) and
    the deviations stay at the FP32 level; before start_av nothing is
    recorded.  The refined run keeps the stream uniform to ~1e-5 (FP32
    rounding of the interface interpolation), hence the 3e-5 bar."""
    say("statistics uniform + refinement")
    cc = cfgmod.Config(nx=32, ny=24, nz=24, bc=("periodic",) * 3, model="MRT", mrt_rates="dhumieres",
                       omega=1 / 0.55, umax=0.04, refinement=True, lr_size=(24, 16, 16),
                       lr_loc=(8.25, 8.25, 8.25))
    with L.Solver(cc) as s:
        s.init_equilibrium(L.INIT_BGK)   # uniform u = UMAX on both levels
        s.stats_enable(3, 8)
        s.step(2)
        m0, _ = s.stats(0)
        assert np.all(m0 == 0)
        s.step(10)
        for lev in (0, 1):
            mean, dev = s.stats(lev)
            say("  level", lev, np.abs(mean[0] - 0.04).max(), np.abs(mean[1:]).max(), np.abs(dev).max())
            assert np.abs(mean[0] - 0.04).max() < 3e-5
            assert np.abs(mean[1:]).max() < 3e-5
            assert np.abs(dev).max() < 1e-9
        # Tecplot: the fine level in WriteResultsLR's 12 columns (X,Y,Z in
        # coarse units, u,v,w,rho, velAv[0..1], ufluc, vfluc, Smag)
        import tempfile
        with tempfile.TemporaryDirectory() as d:
            s.write_tecplot(os.path.join(d, "st"))
            lr = np.loadtxt(os.path.join(d, "st_LR.dat"), delimiter=",", skiprows=2)
            mean, dev = s.stats(1)
            assert lr.shape == (24 * 16 * 16, 12)
            assert np.allclose(lr[:24, 0], 8.25 + 0.5 * np.arange(24))
            # the file holds 6 significant digits (ostream's default)
            for col, ref in ((7, mean[0]), (8, mean[1]), (9, dev[0]), (10, dev[1])):
                assert np.abs(lr[:, col] - ref.reshape(-1)).max() <= 1e-5 * np.abs(ref).max() + 1e-30
            sl = open(os.path.join(d, "st_slice.dat")).read().split("ZONE")
            assert len(sl) == 3  # coarse z = 1 and fine k = 3 zones


def test_statistics_errors():
    c = L.default_config()
    c.nx, c.ny, c.nz = 16, 8, 8
    with L.Solver(c) as s:
        with pytest.raises(L.LBMError):
            s.stats(0)                   # not enabled
        with pytest.raises(L.LBMError):
            s.stats_enable(10, 5)        # fluctuation phase before the mean phase
