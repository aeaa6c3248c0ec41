"""The y-streamwise inflow family (North/South Extrap, the recycling plane,
the im 70 plane), the ghost copies, the precursor's periodic y-wrap and the
time-windowed trip obstacle on the GPU, against the oracle on the same
inputs; and the fine level's momentum-exchange force under refinement.

Tolerances as the other parity tests: single steps within 2e-6 of max|f|,
runs within 1e-5 of the field's scale, forces within 1e-4 of max|F|.
"""
import numpy as np
import pytest

import lbm_multigpu_amd as L
import oracle_lib as O
from lbm_multigpu_amd import config as cfgmod
from test_gpu_parity import NODE_TO_ORC, orc_params, random_state, rel, rel_vel, say

pytestmark = pytest.mark.gpu


def to_orc(m):
    lut = np.zeros(256, np.uint8)
    for k, v in NODE_TO_ORC.items():
        lut[k] = v
    return lut[m]


def inflow_box(model=L.MODEL_MRT_LES, halfway=1):
    """Every new node code in one box: south inlet row 0, ghost of row
    DYNY1-1 in row 1, a force-summed block, the recycling plane, the im 70
    plane, ghosts of rows 1 and DYNY2, the north outlet, and x-ghosts; the
    precursor wrap on rows 0..DYNY1."""
    c = L.default_config()
    c.nx, c.ny, c.nz = 70, 16, 6
    c.bc[0], c.bc[1], c.bc[2] = L.BC_CLAMP, L.BC_PERIODIC, L.BC_WALL
    c.model, c.halfway = model, halfway
    c.omega, c.cs, c.dpdy, c.umax, c.inlet_amp, c.seed = 1.6, 0.01, 1e-5, 0.05, 2e-3, 3
    c.dyn_mode, c.dyn_y1, c.dyn_y2 = 1, 6, 9
    c.force_steps = 4
    m = np.zeros((c.nz, c.ny, c.nx), np.uint8)
    m[:, 0, 1:-1] = L.NODE_INLET_SOUTH
    m[:, 1, 1:-1] = L.NODE_GHOST_S
    m[:, 3:5, 30:34] = L.NODE_SOLID_FORCE
    m[:, 2:6, 0] = L.NODE_GHOST_W
    m[:, 2:6, -1] = L.NODE_GHOST_E
    m[:, 7, 1:-1] = L.NODE_RECYCLE
    m[:, 10, 1:-1] = L.NODE_VEL_FROM_BELOW
    m[:, 12, 1:-1] = L.NODE_GHOST_DYN
    m[:, 13, 1:-1] = L.NODE_GHOST_N
    m[:, c.ny - 1, 1:-1] = L.NODE_OUTLET_NORTH
    return c, m


@pytest.mark.parametrize("model,halfway", [(L.MODEL_MRT_LES, 1), (L.MODEL_BGK, 0), (L.MODEL_MRT_DHUMIERES, 1)])
def test_inflow_family_steps_match_oracle(model, halfway):
    """Three single steps from random populations, each against the oracle
    step of the GPU's own previous state, and the per-step forces of the
    force slots (rows t > STARTF = 0)."""
    say("inflow family", model, "halfway", halfway)
    c, m = inflow_box(model, halfway)
    rng = np.random.default_rng(7)
    f = random_state(rng, m.shape, amp=1e-2)
    om = to_orc(m)
    u_in = O.inlet_profile(c.ny, c.nz, c.umax, c.inlet_amp, c.seed)
    keep = ~np.isin(m, (L.NODE_SOLID, L.NODE_SOLID_FORCE)) if halfway else np.ones(m.shape, bool)
    Fo = []
    with L.Solver(c) as s:
        s.set_geometry(m)
        s.set_populations(f)
        for t in range(3):
            g0 = s.populations()
            s.step(1)
            g = s.populations()
            fo, F1 = O.run(orc_params(c), g0, 1, mask=om, u_in=u_in, forces=True)
            Fo.append(F1[0])
            err = np.abs(g - fo)[:, keep].max() / np.abs(fo).max()
            say(f"  step {t}: max diff {err:.2e}")
            assert err < 2e-6
        F = s.forces()
    Fo = np.array(Fo)
    assert np.all(F[0] == 0) and np.abs(Fo).max() > 0
    assert np.abs(F[1:3] - Fo[1:3]).max() < 1e-4 * np.abs(Fo).max()


def precursor_mask(c, t):
    """ImageFcn of bgk_channel1/Re3000_Cyl2/DynInlet3.cu:75-94 evaluated in
    numpy (float32 comparisons), at step t."""
    z, y, x = np.meshgrid(np.arange(c.nz), np.arange(c.ny), np.arange(c.nx), indexing="ij")
    f = np.float32
    xf, yf = x.astype(f), y.astype(f)
    m = np.zeros((c.nz, c.ny, c.nx), np.uint8)
    if t < c.trip_until:
        trip = (np.abs(xf - f(c.obst2_x)) < f(c.obst2_r)) & (np.abs(yf - f(c.obst2_y)) < f(c.obst2_r))
        trip |= ((np.abs(xf - f(c.obst2_x) - f(3)) < f(c.obst2_r)) & (np.abs(yf - f(c.obst2_y) - f(3)) < f(c.obst2_r))
                 & (z == c.trip_z))
        m[trip] = L.NODE_SOLID
    m[(np.abs(xf - f(c.obst_x)) < f(c.obst_r)) & (np.abs(yf - f(c.obst_y)) < f(c.obst_r))] = L.NODE_SOLID_FORCE
    side = (x == 0) | (x == c.nx - 1)
    m[side] = L.NODE_SOLID
    m[~side & (y == c.dyn_y1 + 1)] = L.NODE_RECYCLE
    m[~side & (y != c.dyn_y1 + 1) & (y == c.ny - 1)] = L.NODE_OUTLET_NORTH
    return m


@pytest.mark.parametrize("storage", ["f32", "f64"])
def test_precursor_run_matches_oracle(storage):
    """The bgk_channel1-style precursor (cfgmod.precursor) for 60 steps with
    the trip obstacle removed after step 30: the library re-classifies the
    mask itself, both masks are bit-exact against the ImageFcn restatement,
    and the fields after 60 steps agree with the oracle run over the same
    two masks (1e-5 of the field's scale; 1e-12 in FP64)."""
    say("precursor run", storage)
    cc = cfgmod.precursor(nx=42, ny=64, nz=6, dyn_y1=16, side=6, obst_y=36, trip=(14, 8, 3), trip_until=30,
                          storage=storage)
    cc.force_steps, cc.startf = 60, 0
    c = cc.to_c()
    with L.Solver(c) as s:
        s.set_geometry(None)
        m0 = s.get_mask()
        s.init_equilibrium(L.INIT_MRT)
        f0 = s.populations(f64=True)
        s.step(30)
        s.step(1)  # the first step without the trip
        m1 = s.get_mask()
        s.step(29)
        rho, u, v, w = s.moments()
        F = s.forces()
    assert np.array_equal(m0, precursor_mask(c, 0))
    assert np.array_equal(m1, precursor_mask(c, 30))
    assert (m0 != m1).any() and (m0 == L.NODE_SOLID).sum() > (m1 == L.NODE_SOLID).sum()
    dt = np.float64 if storage == "f64" else np.float32
    p = orc_params(c)
    fa, Fa = O.run(p, f0.astype(dt), 30, mask=to_orc(m0), forces=True)
    fb, Fb = O.run(p, fa, 30, mask=to_orc(m1), forces=True)
    r2, u2, v2, w2 = O.macro(p, fb)
    fl = np.isin(m1, (L.NODE_FLUID, L.NODE_RECYCLE, L.NODE_OUTLET_NORTH))
    tol = 1e-12 if storage == "f64" else 1e-5
    assert rel(rho[fl], r2[fl]) < tol
    assert rel_vel((u[fl], v[fl], w[fl]), (u2[fl], v2[fl], w2[fl])) < tol
    Fo = np.concatenate([Fa, Fb])
    assert np.all(F[0] == 0) and np.abs(Fo).max() > 0
    assert np.abs(F[1:] - Fo[1:60]).max() < (1e-9 if storage == "f64" else 1e-4) * np.abs(Fo).max()
    # the precursor carries the flow into the channel: positive mean v above the recycling plane
    assert v[:, cc.dyn_y1 + 2:, 1:-1].mean() > 0


def test_refined_channel_records_the_fine_level_force():
    """Under refinement the force history holds the fine level's momentum
    exchange of the first sub-step of each coarse step t > STARTF
    (MRTtest/DynInlet3.cu:3130, :3183), against the oracle's refined step."""
    say("refined forces")
    cc = cfgmod.channel(nx=64, ny=32, nz=32, side=6, x_center=22, tau=0.55, rates="lrmgpu")
    cc.refinement = True
    cc.lr_size = (32, 24, 24)
    cc.lr_loc = (12.25, 4.25, 4.25)
    cc.force_steps = 12
    c = cc.to_c()
    T = 10
    with L.Solver(c) as s:
        s.set_geometry(None)
        cm, fm = to_orc(s.get_mask(0)), to_orc(s.get_mask(1))
        assert (fm == O.SOLID_FORCE).any()
        s.init_equilibrium(L.INIT_MRT)
        c0, f0 = s.populations(0), s.populations(1)
        s.step(T)
        F = s.forces()
    pc = orc_params(c)
    r = O.refine([int(round(x - 0.25)) for x in cc.lr_loc], cc.lr_size, float(np.float32(c.omega)),
                 cs=float(np.float32(c.cs)), temporal=1, eq_kind=O.EQ_BGK_MEQ)
    pf = O.params(*cc.lr_size, bc=(O.CLAMP,) * 3, collision=pc.collision, halfway=1, omega=r.omega_f,
                  cs=pc.cs, dpdy=0.5 * pc.dpdy)
    u_in = O.inlet_profile(c.ny, c.nz, c.umax, c.inlet_amp, c.seed)
    _, _, Fo = O.refine_run(pc, pf, r, c0, f0, T, cmask=cm, fmask=fm, u_in=u_in, forces=True)
    assert F.shape == (T, 3) and np.all(F[0] == 0)
    assert np.abs(Fo).max() > 0
    assert np.abs(F[1:] - Fo[1:]).max() < 1e-4 * np.abs(Fo).max()
