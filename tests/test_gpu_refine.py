"""2:1 local refinement on the GPU against the oracle's restatement
(oracle/refine_impl.h, float and double instances) on the same inputs: both levels after several coarse
steps, with and without temporal interpolation and with both interface
equilibria; the fine mask bit-exact against the geometry evaluated at the
fine nodes; the full-size patch stays finite and tracks the coarse flow."""
import sys

import numpy as np
import pytest

import lbm_multigpu_amd as L
import oracle_lib as O
from lbm_multigpu_amd import config as cfgmod
from test_gpu_parity import NODE_TO_ORC, orc_params, rel, rel_vel

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def to_orc(m):
    lut = np.zeros(256, np.uint8)
    for k, v in NODE_TO_ORC.items():
        lut[k] = v
    return lut[m]


def small_case(temporal=True, interface_eq=0, rates="lrmgpu", dpdy=0.0):
    cc = cfgmod.channel(nx=64, ny=32, nz=32, side=6, x_center=22, tau=0.55, rates=rates)
    cc.dpdy = dpdy
    cc.refinement = True
    cc.lr_size = (32, 24, 24)
    cc.lr_loc = (12.25, 4.25, 4.25)
    cc.temporal_interp = temporal
    cc.interface_eq = interface_eq
    return cc


@pytest.mark.parametrize("temporal,interface_eq,dpdy", [(True, 0, 0.0), (False, 0, 0.0), (True, 1, 0.0),
                                                         (True, 0, 2e-5)])
def test_refined_channel_matches_oracle(temporal, interface_eq, dpdy):
    """With a body force the fine level takes LRFACTOR x DPDY (the force
    density of the 2:1 acoustic scaling)."""
    say("refine parity temporal", temporal, "eq", interface_eq, "dpdy", dpdy)
    cc = small_case(temporal, interface_eq, dpdy=dpdy)
    c = cc.to_c()
    T = 10
    with L.Solver(c) as s:
        s.set_geometry(None)
        cm, fm = to_orc(s.get_mask(0)), to_orc(s.get_mask(1))
        s.init_equilibrium(L.INIT_MRT)
        c0, f0 = s.populations(0), s.populations(1)
        s.step(T)
        cg, fg = s.moments(0), s.moments(1)
    pc = orc_params(c)
    r = O.refine([int(round(x - 0.25)) for x in cc.lr_loc], cc.lr_size, float(np.float32(c.omega)),
                 cs=float(np.float32(c.cs)), temporal=int(temporal),
                 eq_kind=(O.EQ_MRT_MEQ if interface_eq else O.EQ_BGK_MEQ))
    pf = O.params(*cc.lr_size, bc=(O.CLAMP,) * 3, collision=pc.collision, halfway=1, omega=r.omega_f,
                  cs=pc.cs, dpdy=0.5 * pc.dpdy)
    u_in = O.inlet_profile(c.ny, c.nz, c.umax, c.inlet_amp, c.seed)
    ca, fa = O.refine_run(pc, pf, r, c0, f0, T, cmask=cm, fmask=fm, u_in=u_in)
    co = O.macro(pc, ca)
    fo = O.macro(O.params(*cc.lr_size), fa)
    cfl = cm == O.FLUID
    ffl = fm == O.FLUID
    # the fine frame's outer layer is rewritten each coarse step (its own update is discarded)
    inner = np.zeros_like(ffl)
    inner[2:-2, 2:-2, 2:-2] = True
    for got, ref, sel in ((cg, co, cfl), (fg, fo, ffl & inner)):
        assert rel(got[0][sel], ref[0][sel]) < 1e-5
        assert rel_vel([g[sel] for g in got[1:]], [x[sel] for x in ref[1:]]) < 1e-5


@pytest.mark.parametrize("storage,tol", [("f64", 1e-5), ("f16", 2e-3), ("bf16", 2e-2)])
def test_refined_channel_other_storage_formats(storage, tol):
    """Both levels, the interface kernels and the 16-bit offsets with FP64,
    FP16 and BF16 populations: 10 coarse steps against the FP32 oracle from
    the same (stored) initial state.  FP64 differs from it by FP32 rounding
    only; the 16-bit bars are those of the uniform-grid 16-bit tests."""
    say("refine storage", storage)
    cc = small_case()
    cc.storage = storage
    c = cc.to_c()
    T = 10
    with L.Solver(c) as s:
        s.set_geometry(None)
        cm, fm = to_orc(s.get_mask(0)), to_orc(s.get_mask(1))
        s.init_equilibrium(L.INIT_MRT)
        c0, f0 = s.populations(0), s.populations(1)
        s.step(T)
        cg, fg = s.moments(0), s.moments(1)
    pc = orc_params(c)
    r = O.refine([int(round(x - 0.25)) for x in cc.lr_loc], cc.lr_size, float(np.float32(c.omega)),
                 cs=float(np.float32(c.cs)), temporal=1, eq_kind=O.EQ_BGK_MEQ)
    pf = O.params(*cc.lr_size, bc=(O.CLAMP,) * 3, collision=pc.collision, halfway=1, omega=r.omega_f,
                  cs=pc.cs, dpdy=0.5 * pc.dpdy)
    u_in = O.inlet_profile(c.ny, c.nz, c.umax, c.inlet_amp, c.seed)
    ca, fa = O.refine_run(pc, pf, r, c0, f0, T, cmask=cm, fmask=fm, u_in=u_in)
    co = O.macro(pc, ca)
    fo = O.macro(O.params(*cc.lr_size), fa)
    inner = np.zeros_like(fm, dtype=bool)
    inner[2:-2, 2:-2, 2:-2] = True
    for name, got, ref, sel in (("coarse", cg, co, cm == O.FLUID), ("fine", fg, fo, (fm == O.FLUID) & inner)):
        err = rel_vel([g[sel] for g in got[1:]], [x[sel] for x in ref[1:]])
        say(f"  {name}: velocity error vs the FP32 oracle {err:.2e} (bar {tol})")
        assert err < tol


@pytest.mark.parametrize("temporal", [True, False])
def test_refined_channel_fp64_end_to_end(temporal):
    """FP64 populations with the interface in FP64 too (moments, CF data,
    trilinear/ORDER-2 interpolation, LES rescale): 10 coarse steps against
    the oracle's double instance of the interface from the same FP64 state,
    to 1e-11 -- FP64 rounding, where an FP32 step anywhere on the path would
    show at 1e-7."""
    say("refine FP64 end to end, temporal", temporal)
    cc = small_case(temporal)
    cc.storage = "f64"
    c = cc.to_c()
    T = 10
    with L.Solver(c) as s:
        s.set_geometry(None)
        cm, fm = to_orc(s.get_mask(0)), to_orc(s.get_mask(1))
        s.init_equilibrium(L.INIT_MRT)
        c0, f0 = s.populations(0, f64=True), s.populations(1, f64=True)
        s.step(T)
        cg, fg = s.moments(0), s.moments(1)
    pc = orc_params(c)
    r = O.refine([int(round(x - 0.25)) for x in cc.lr_loc], cc.lr_size, float(np.float32(c.omega)),
                 cs=float(np.float32(c.cs)), temporal=int(temporal), eq_kind=O.EQ_BGK_MEQ)
    pf = O.params(*cc.lr_size, bc=(O.CLAMP,) * 3, collision=pc.collision, halfway=1, omega=r.omega_f,
                  cs=pc.cs, dpdy=0.5 * pc.dpdy)
    u_in = O.inlet_profile(c.ny, c.nz, c.umax, c.inlet_amp, c.seed)
    ca, fa = O.refine_run(pc, pf, r, c0, f0, T, cmask=cm, fmask=fm, u_in=u_in)
    assert ca.dtype == np.float64
    co = O.macro(pc, ca)
    fo = O.macro(O.params(*cc.lr_size), fa)
    inner = np.zeros_like(fm, dtype=bool)
    inner[2:-2, 2:-2, 2:-2] = True
    for name, got, ref, sel in (("coarse", cg, co, cm == O.FLUID), ("fine", fg, fo, (fm == O.FLUID) & inner)):
        err = max(rel(got[0][sel], ref[0][sel]), rel_vel([g[sel] for g in got[1:]], [x[sel] for x in ref[1:]]))
        say(f"  {name}: error vs the FP64 oracle {err:.2e}")
        assert err < 1e-11


def test_fine_mask_is_geometry_at_fine_nodes():
    cc = small_case()
    c = cc.to_c()
    with L.Solver(c) as s:
        s.set_geometry(None)
        fm = s.get_mask(1)
    nfx, nfy, nfz = cc.lr_size
    k, j, i = np.meshgrid(np.arange(nfz), np.arange(nfy), np.arange(nfx), indexing="ij")
    X = np.float32(cc.lr_loc[0]) + np.float32(0.5) * i.astype(np.float32)
    Y = np.float32(cc.lr_loc[1]) + np.float32(0.5) * j.astype(np.float32)
    solid = (np.abs(X - np.float32(cc.obst1[0])) < np.float32(cc.obst1_r)) & \
            (np.abs(Y - np.float32(cc.obst1[1])) < np.float32(cc.obst1_r))
    exp = np.where(solid, L.NODE_SOLID_FORCE, L.NODE_FLUID).astype(np.uint8)
    assert np.array_equal(fm, exp)
    assert solid.sum() > 0


def test_full_size_refined_channel_runs_and_tracks_coarse():
    """The refined-channel workload (coarse 256x128x128, fine patch
    128x96x96 around the cylinder at a seeded offset), 200 coarse steps:
    finite everywhere, and on the patch's inner footprint the fine and
    coarse velocity fields agree to the discretisation level."""
    cc = cfgmod.refined_channel()
    say("refined channel", cc.lr_loc)
    with L.Solver(cc.to_c()) as s:
        s.set_geometry(None)
        s.init_equilibrium(L.INIT_MRT)
        ms = s.step_timed(200)
        rho_c, u_c, v_c, w_c = s.moments(0)
        rho_f, u_f, v_f, w_f = s.moments(1)
    cells = 256 * 128 * 128 + 2 * 128 * 96 * 96
    say(f"  {ms / 200:.3f} ms/coarse step, {cells / (ms / 200 * 1e-3) / 1e9:.2f} GLUPS (coarse + 2 x fine)")
    assert np.isfinite(u_c).all() and np.isfinite(u_f).all()
    assert np.abs(u_c).max() < 0.2 and np.abs(u_f).max() < 0.2
    a = [int(round(x - 0.25)) for x in cc.lr_loc]
    # coarse node (a+i) sits between fine nodes 2i-1 and 2i: compare the mean
    fx = 0.5 * (u_f[:, :, 1:-1:2] + u_f[:, :, 2::2])
    fx = 0.5 * (fx[:, 1:-1:2, :] + fx[:, 2::2, :])
    fx = 0.5 * (fx[1:-1:2] + fx[2::2])
    n = fx.shape
    cx = u_c[a[2] + 1:a[2] + 1 + n[0], a[1] + 1:a[1] + 1 + n[1], a[0] + 1:a[0] + 1 + n[2]]
    assert np.abs(fx - cx).mean() < 0.1 * np.abs(cx).mean()
