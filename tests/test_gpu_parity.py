"""HIP path (liblbm.so through its C ABI) against the oracle on the same
inputs, against the reference program's golden output, and against
size-independent properties at full size.

Tolerances: FP32 runs agree with the FP32 oracle within 1e-5 relative to the
field's scale (rho, u) after up to 1000 steps; single-step checks within
2e-6; FP64 storage within 1e-12.
"""
import os
import sys

import numpy as np
import pytest

import lbm_multigpu_amd as L
import oracle_lib as O
from lbm_multigpu_amd import config as cfgmod
from test_oracle import cpu3d_mask

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MODEL_TO_ORC = {L.MODEL_BGK: O.BGK, L.MODEL_MRT_CPU3D: O.MRT_CPU3D, L.MODEL_MRT_LES: O.MRT_LES,
                L.MODEL_MRT_DHUMIERES: O.MRT_DHUM}
# the library's node codes and the oracle's are the same numbers (both
# follow include/lbm.h); the table keeps the correspondence explicit
NODE_TO_ORC = {L.NODE_FLUID: O.FLUID, L.NODE_SOLID: O.SOLID, L.NODE_SOLID_FORCE: O.SOLID_FORCE,
               L.NODE_INLET_WEST: O.INLET_WEST, L.NODE_OUTLET_EAST: O.OUTLET_EAST,
               L.NODE_LID_WEST_REG: O.LID_WEST_REG, L.NODE_SYM_XMIN: O.SYM_XMIN, L.NODE_SYM_XMAX: O.SYM_XMAX,
               L.NODE_OUTLET_NORTH: O.OUTLET_NORTH, L.NODE_INLET_SOUTH: O.INLET_SOUTH, L.NODE_RECYCLE: O.RECYCLE,
               L.NODE_VEL_FROM_BELOW: O.VEL_FROM_BELOW, L.NODE_GHOST_W: O.GHOST_W, L.NODE_GHOST_E: O.GHOST_E,
               L.NODE_GHOST_S: O.GHOST_S, L.NODE_GHOST_N: O.GHOST_N, L.NODE_GHOST_DYN: O.GHOST_DYN}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def orc_params(c):
    """Oracle parameters of a library configuration: the precursor wrap and
    the ghost rows as the library derives them, and the south inlet table
    (the library's generator over (x, z))."""
    wrap = c.dyn_y1 if c.dyn_mode == 1 else -1
    gy = tuple(min(max(v, 0), c.ny - 1) for v in (c.dyn_y1 - 1, c.dyn_y2, 1))
    v_in = O.inlet_profile(c.nx, c.nz, c.umax, c.inlet_amp, c.seed)
    p = O.params(c.nx, c.ny, c.nz, bc=tuple(c.bc), collision=MODEL_TO_ORC[c.model], halfway=c.halfway,
                 omega=float(np.float32(c.omega)), cs=float(np.float32(c.cs)), dpdy=float(np.float32(c.dpdy)),
                 u_lid=float(np.float32(c.umax)), dyn_y1=c.dyn_y1 if c.dyn_mode == 0 else -1, y_wrap=wrap,
                 ghost_y=gy, v_in=v_in)
    for k in range(19):
        p.s_dhum[k] = float(np.float32(c.s_dhum[k]))
    return p


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)


def rel_vel(got, ref):
    """Velocity error relative to the velocity field's scale max|u| (the
    vector magnitude), over all three components."""
    scale = np.sqrt(sum(r.astype(np.float64) ** 2 for r in ref)).max()
    return max(np.abs(g - r).max() for g, r in zip(got, ref)) / scale


def random_state(rng, shape, amp=1e-3, umax=0.05, dtype=np.float32):
    nz, ny, nx = shape
    u = rng.uniform(-umax, umax, (3, nz, ny, nx))
    f = np.empty((19, nz, ny, nx), np.float64)
    C = np.array([(0, 0, 0), (1, 0, 0), (0, 1, 0), (-1, 0, 0), (0, -1, 0), (1, 1, 0), (-1, 1, 0), (-1, -1, 0),
                  (1, -1, 0), (0, 0, 1), (1, 0, 1), (0, 1, 1), (-1, 0, 1), (0, -1, 1), (0, 0, -1), (1, 0, -1),
                  (0, 1, -1), (-1, 0, -1), (0, -1, -1)])
    W = np.array([1 / 3] + [1 / 18] * 4 + [1 / 36] * 4 + [1 / 18] + [1 / 36] * 4 + [1 / 18] + [1 / 36] * 4)
    usq = (u ** 2).sum(0)
    for i in range(19):
        cu = C[i, 0] * u[0] + C[i, 1] * u[1] + C[i, 2] * u[2]
        f[i] = W[i] * (1.0 + 3 * cu + 4.5 * cu * cu - 1.5 * usq)
    f += rng.uniform(-amp, amp, f.shape) * W[:, None, None, None]
    return f.astype(dtype)


@pytest.mark.parametrize("model", [L.MODEL_BGK, L.MODEL_MRT_CPU3D, L.MODEL_MRT_LES, L.MODEL_MRT_DHUMIERES])
@pytest.mark.parametrize("les", [0.0, 0.01])
def test_single_step_each_model(model, les):
    if model == L.MODEL_MRT_CPU3D and les:
        pytest.skip("no LES in the CPU_3D operator")
    say("single step model", model, "cs", les)
    c = L.default_config()
    c.nx, c.ny, c.nz = 70, 9, 7   # nx not a multiple of 64: exercises the padded rows
    c.model, c.omega, c.cs, c.dpdy = model, 1.7, les, (1e-5 if model != L.MODEL_MRT_CPU3D else 0.0)
    rng = np.random.default_rng(model * 10 + int(les * 100))
    f = random_state(rng, (c.nz, c.ny, c.nx))
    with L.Solver(c) as s:
        s.set_populations(f)
        s.step(1)
        g = s.populations()
    ref, _ = O.step(orc_params(c), f)
    assert np.abs(g - ref).max() < 2e-6 * np.abs(ref).max()


def test_mrt_form_fill_matches_oracle():
    """initialize's fill (LBM_INIT_MRT: mrt_meq then InvertMoments,
    MRTtest/DynInlet3.cu:1186-1216) from seeded rho and u fields against the
    oracle's orc_feq_mrt node by node; solid nodes get u = 0, as there."""
    say("MRT-form fill")
    c = L.default_config()
    c.nx, c.ny, c.nz = 70, 9, 7
    c.model, c.omega = L.MODEL_MRT_LES, 1.7
    rng = np.random.default_rng(31)
    shp = (c.nz, c.ny, c.nx)
    rho = rng.uniform(0.97, 1.03, shp).astype(np.float32)
    u, v, w = (rng.uniform(-0.06, 0.06, shp).astype(np.float32) for _ in range(3))
    mask = np.zeros(shp, np.uint8)
    mask[:, 3:5, 20:24] = L.NODE_SOLID_FORCE
    with L.Solver(c) as s:
        s.set_geometry(mask)
        s.init_equilibrium(L.INIT_MRT, rho, u, v, w)
        g = s.populations()
    ref = np.empty_like(g)
    for idx in np.ndindex(*shp):
        solid = mask[idx] != 0
        uu = (0.0, 0.0, 0.0) if solid else (float(u[idx]), float(v[idx]), float(w[idx]))
        ref[(slice(None),) + idx] = O.feq_mrt(float(rho[idx]), *uu)
    assert np.abs(g - ref).max() < 2e-7 * np.abs(ref).max()


def test_taylor_green_1000_steps_matches_oracle():
    """32^3 periodic BGK tau=0.6 Taylor-Green, 1000 steps, GPU vs oracle.

    Two FP32 implementations of the same step differ by rounding (FMA
    contraction, summation order) that accumulates over 1000 steps to ~3e-7
    in u while the vortex decays 13x, so the 1e-5 bar is taken relative to
    the initial velocity scale u0 (and to rho's scale).  The GPU's FP32 error
    against the FP64 oracle must also be no larger than (2x) the FP32
    oracle's own error against FP64."""
    n, tau, u0, T = 32, 0.6, 0.05, 1000
    say("TGV parity", n, T)
    c = L.default_config()
    c.nx = c.ny = c.nz = n
    c.omega = 1 / tau
    with L.Solver(c) as s:
        s.init_taylor_green(u0)
        f0 = s.populations()
        s.step(T)
        rho, u, v, w = s.moments()
    p = orc_params(c)
    fo, _ = O.run(p, f0, T)
    r2, u2, v2, w2 = O.macro(p, fo)
    assert rel(rho, r2) < 1e-5
    assert np.abs(u - u2).max() < 1e-5 * u0 and np.abs(v - v2).max() < 1e-5 * u0
    assert np.abs(w).max() < 1e-6
    fd, _ = O.run(p, f0.astype(np.float64), T)
    r3, u3, v3, _ = O.macro(p, fd)
    err_gpu = max(np.abs(u - u3).max(), np.abs(v - v3).max())
    err_cpu = max(np.abs(u2 - u3).max(), np.abs(v2 - v3).max())
    say(f"  |u_gpu - u_fp64| = {err_gpu:.3e}, |u_oracle32 - u_fp64| = {err_cpu:.3e}")
    assert err_gpu <= 2 * err_cpu + 1e-8


def test_taylor_green_64_analytic_decay():
    """The 64^3 Taylor-Green workload (BGK tau=0.6, u0=0.05, k=2pi/64, 1000
    steps): kinetic energy follows exp(-4 nu k^2 t) to within 1%."""
    n, tau, u0, T = 64, 0.6, 0.05, 1000
    say("TGV decay", n, T)
    c = L.default_config()
    c.nx = c.ny = c.nz = n
    c.omega = 1 / tau
    with L.Solver(c) as s:
        s.init_taylor_green(u0)
        _, u, v, _ = s.moments()
        ke0 = (u ** 2 + v ** 2).sum()
        s.step(T)
        _, u, v, w = s.moments()
        mm = s.mass_momentum()
    k = 2 * np.pi / n
    nu = (tau - 0.5) / 3
    ratio = (u ** 2 + v ** 2).sum() / ke0
    assert abs(ratio / np.exp(-4 * nu * k * k * T) - 1) < 0.01
    # FP32 weights (1/3, 1/18, 1/36 rounded, as the reference's literals) sum to
    # 1 + 1.5e-8, so BGK adds ~omega*1.5e-8 of mass per step: <= 3e-5 here
    assert abs(mm[0] - n ** 3) < 3e-5 * n ** 3
    c.storage = L.STORE_F64
    with L.Solver(c) as s:
        s.init_taylor_green(u0)
        m0 = s.mass_momentum()
        s.step(T)
        m1 = s.mass_momentum()
    assert abs(m1[0] - m0[0]) < 1e-11 * m0[0] and np.abs(m1[1:]).max() < 1e-9


def test_cpu3d_case_matches_reference_program():
    """The reference program's own case (CPU_3D/CPU_3D.cpp, 251x43x43,
    clamped pulls, node bounce-back, regularised moving x=0 plane, MRT with
    m13/m14 at rate 1), 100 steps on the GPU, against the reference's
    output.dat (golden) and the oracle; masks bit-exact."""
    say("cpu3d case")
    g = np.load(os.path.join(GOLD, "cpu3d_t100.npz"))
    nx, ny, nz, tmax = (int(d) for d in g["dims"])
    umax, omega = float(g["consts"][0]), float(g["consts"][3])
    cc = cfgmod.cpu3d_case()
    assert abs(np.float32(cc.omega) - np.float32(omega)) == 0
    c = cc.to_c()
    with L.Solver(c) as s:
        s.set_geometry(None)
        m = s.get_mask()
        exp = cpu3d_mask(nx, ny, nz)
        assert np.array_equal(np.vectorize(NODE_TO_ORC.get)(m).astype(np.uint8), exp)
        s.init_equilibrium(L.INIT_BGK)
        f0 = s.populations()
        s.step(tmax)
        rho, u, v, w = s.moments()
    zs = g["zs"]
    for got, ref in ((rho[zs], g["rho"]), (u[zs], g["u"]), (v[zs], g["v"])):
        assert np.abs(got - ref).max() <= 1e-5 * np.abs(ref).max()
    fo, _ = O.run(orc_params(c), f0, tmax, mask=exp)
    r2, u2, v2, w2 = O.macro(orc_params(c), fo)
    assert rel(rho, r2) < 1e-5 and rel_vel((u, v, w), (u2, v2, w2)) < 1e-5


def test_cpu3d_case_1000_steps_matches_oracle():
    """The same case run 1000 steps (the reference program stops at its
    TMAX of 100, where the oracle is pinned to its output above): GPU FP32
    against the FP32 oracle, density and velocity within 1e-5 relative; the
    FP64 oracle shows how far both FP32 runs are from exact arithmetic."""
    say("cpu3d case, 1000 steps")
    cc = cfgmod.cpu3d_case()
    c = cc.to_c()
    T = 1000
    with L.Solver(c) as s:
        s.set_geometry(None)
        exp = np.vectorize(NODE_TO_ORC.get)(s.get_mask()).astype(np.uint8)
        s.init_equilibrium(L.INIT_BGK)
        f0 = s.populations()
        s.step(T)
        rho, u, v, w = s.moments()
    p = orc_params(c)
    fo, _ = O.run(p, f0, T, mask=exp)
    r2, u2, v2, w2 = O.macro(p, fo)
    fo64, _ = O.run(p, f0.astype(np.float64), T, mask=exp)
    r6, u6, v6, w6 = O.macro(p, fo64)
    e_rho, e_u = rel(rho, r2), rel_vel((u, v, w), (u2, v2, w2))
    say(f"  GPU vs FP32 oracle: rho {e_rho:.2e}, u {e_u:.2e}; vs FP64 oracle: u GPU {rel_vel((u, v, w), (u6, v6, w6)):.2e},"
        f" oracle FP32 {rel_vel((u2, v2, w2), (u6, v6, w6)):.2e}")
    assert e_rho < 1e-5 and e_u < 1e-5


@pytest.mark.parametrize("rates", ["lrmgpu", "dhumieres"])
def test_channel_mrt_les_matches_oracle(rates):
    """Channel with a square cylinder, halfway walls at y/z, West_Extrap inlet
    with the seeded perturbation, East_Extrap outlet, MRT + Smagorinsky:
    200 steps GPU vs oracle, fields and per-step momentum-exchange forces."""
    say("channel", rates)
    cc = cfgmod.channel(nx=96, ny=32, nz=24, side=8, x_center=32, tau=0.52, rates=rates, force_steps=200)
    c = cc.to_c()
    T = 200
    with L.Solver(c) as s:
        s.set_geometry(None)
        mask = np.vectorize(NODE_TO_ORC.get)(s.get_mask()).astype(np.uint8)
        s.init_equilibrium(L.INIT_BGK)
        f0 = s.populations()
        s.step(T)
        rho, u, v, w = s.moments()
        F = s.forces()
    u_in = O.inlet_profile(c.ny, c.nz, c.umax, c.inlet_amp, c.seed)
    fo, Fo = O.run(orc_params(c), f0, T, mask=mask, u_in=u_in, forces=True)
    r2, u2, v2, w2 = O.macro(orc_params(c), fo)
    fl = mask == O.FLUID
    assert rel(rho[fl], r2[fl]) < 1e-5
    assert rel_vel((u[fl], v[fl], w[fl]), (u2[fl], v2[fl], w2[fl])) < 1e-5
    # steps t > STARTF = 0 are recorded into row t; row 0 stays zero
    assert F.shape == (T, 3) and np.all(F[0] == 0)
    assert np.abs(F[1:] - Fo[1:]).max() < 1e-4 * np.abs(Fo).max()
    assert np.abs(Fo).max() > 0


def test_halfway_closed_box_conserves_mass_fp64_sums():
    say("closed box")
    c = L.default_config()
    c.nx, c.ny, c.nz = 64, 20, 12
    c.bc[0] = c.bc[1] = c.bc[2] = L.BC_WALL
    c.model, c.halfway, c.omega, c.cs = L.MODEL_MRT_LES, 1, 1.8, 0.01
    mask = np.zeros((c.nz, c.ny, c.nx), np.uint8)
    mask[:, 8:12, 20:24] = L.NODE_SOLID_FORCE
    rng = np.random.default_rng(11)
    f = random_state(rng, (c.nz, c.ny, c.nx), amp=1e-2, umax=0.02)
    with L.Solver(c) as s:
        s.set_geometry(mask)
        s.set_populations(f)
        m0 = s.mass_momentum()
        s.step(500)
        m1 = s.mass_momentum()
    assert abs(m1[0] - m0[0]) < 2e-6 * m0[0]


def test_fp64_storage_matches_fp64_oracle():
    say("fp64 storage")
    cc = cfgmod.channel(nx=64, ny=16, nz=16, side=6, x_center=20, tau=0.55, rates="dhumieres", storage="f64")
    c = cc.to_c()
    with L.Solver(c) as s:
        s.set_geometry(None)
        mask = np.vectorize(NODE_TO_ORC.get)(s.get_mask()).astype(np.uint8)
        s.init_equilibrium(L.INIT_MRT)
        f0 = s.populations(f64=True)
        s.step(50)
        g = s.populations(f64=True)
    u_in = O.inlet_profile(c.ny, c.nz, c.umax, c.inlet_amp, c.seed).astype(np.float64)
    fo, _ = O.run(orc_params(c), f0, 50, mask=mask, u_in=u_in)
    fl = np.broadcast_to(mask == O.FLUID, g.shape)
    assert np.abs(g[fl] - fo[fl]).max() < 1e-12


@pytest.mark.parametrize("storage,tol", [("f16", 1e-3), ("bf16", 1e-2)])
def test_16bit_storage_error_vs_fp64(storage, tol):
    """Shifted-deviation 16-bit storage with FP32 collide vs an FP64 run of the
    same case: velocity error relative to the field's scale."""
    say("16-bit", storage)
    out = {}
    for st in ("f64", storage):
        c = cfgmod.channel(nx=128, ny=32, nz=32, side=8, x_center=32, storage=st).to_c()
        with L.Solver(c) as s:
            s.set_geometry(None)
            s.init_equilibrium(L.INIT_BGK)
            s.step(300)
            out[st] = s.moments()
    err = rel_vel(out[storage][1:], out["f64"][1:])
    say(f"  {storage}: velocity error vs FP64 {err:.3e} (bar {tol})")
    assert err < tol


W19 = np.array([12] + [2] * 4 + [1] * 4 + [2] + [1] * 4 + [2] + [1] * 4, np.float32) / np.float32(36)
CX19 = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1, 0, 1, 0, -1, 0, 0, 1, 0, -1, 0], np.float32)


def storage_offsets(umax):
    """Offsets the 16-bit formats subtract before rounding: the equilibrium
    populations at rho = 1, u = (umax, 0, 0) (float32 arithmetic)."""
    U = np.float32(umax)
    cu = CX19 * U
    return W19 * (np.float32(1) + np.float32(3) * cu + np.float32(4.5) * cu * cu - np.float32(1.5) * U * U)


def on_16bit_grid(f, storage, sh):
    """Round f - offset to FP16 / BF16 (nearest even) and add the offset
    back; also the spacing of that grid at each value."""
    d = (f - sh[:, None, None, None]).astype(np.float32)
    if storage == "f16":
        h = d.astype(np.float16)
        q = h.astype(np.float32)
        ulp = np.spacing(np.abs(h)).astype(np.float32)
    else:
        b = d.view(np.uint32).astype(np.uint64)
        b = ((b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
        q = b.view(np.float32)
        _, e = np.frexp(np.abs(q))
        ulp = np.ldexp(np.float32(1), e - 8).astype(np.float32)
    return q + sh[:, None, None, None], ulp


def _bc_case(case):
    """(config, mask) of a 16-bit parity case: odd and even widths, one and
    several 128-node segments, every x boundary condition and node kind."""
    c = L.default_config()
    c.umax = 0.05
    c.omega, c.cs, c.dpdy = 1.6, 0.01, 1e-5
    c.model = L.MODEL_MRT_LES
    if case == "periodic_odd":
        c.nx, c.ny, c.nz = 37, 6, 5
        mask = np.zeros((c.nz, c.ny, c.nx), np.uint8)
    elif case == "walls_halfway_solid":
        c.nx, c.ny, c.nz = 130, 7, 6
        c.bc[0], c.bc[1], c.bc[2] = L.BC_WALL, L.BC_WALL, L.BC_WALL
        c.halfway, c.force_steps = 1, 8
        mask = np.zeros((c.nz, c.ny, c.nx), np.uint8)
        mask[1:4, 2:5, 63:67] = L.NODE_SOLID_FORCE
        mask[1:3, 1:3, 0:2] = L.NODE_SOLID
    elif case == "inlet_outlet_node_bb":
        c.nx, c.ny, c.nz = 69, 8, 6
        c.bc[0], c.bc[1], c.bc[2] = L.BC_CLAMP, L.BC_PERIODIC, L.BC_CLAMP
        c.model, c.force_steps = L.MODEL_BGK, 8
        mask = np.zeros((c.nz, c.ny, c.nx), np.uint8)
        mask[:, :, 0] = L.NODE_INLET_WEST
        mask[:, :, -1] = L.NODE_OUTLET_EAST
        mask[2:4, 3:6, 30:33] = L.NODE_SOLID_FORCE
        mask[0, :, 1:68] = L.NODE_SOLID
    elif case == "inflow_family":
        from test_gpu_inflow import inflow_box
        c, mask = inflow_box()
        c.umax = 0.05
        return c, mask
    else:  # symmetry planes
        c.nx, c.ny, c.nz = 256, 6, 5
        c.bc[0], c.bc[1], c.bc[2] = L.BC_CLAMP, L.BC_WALL, L.BC_WALL
        c.halfway = 1
        c.model = L.MODEL_MRT_DHUMIERES
        mask = np.zeros((c.nz, c.ny, c.nx), np.uint8)
        mask[:, :, 0] = L.NODE_SYM_XMIN
        mask[:, :, -1] = L.NODE_SYM_XMAX
    return c, mask


@pytest.mark.parametrize("storage", ["f16", "bf16"])
@pytest.mark.parametrize("case", ["periodic_odd", "walls_halfway_solid", "inlet_outlet_node_bb", "symmetry",
                                  "inflow_family"])
def test_16bit_steps_match_oracle_on_the_16bit_grid(storage, case):
    """Three single steps of 16-bit storage, each from the GPU's own previous
    state (as stored: 16-bit value plus offset), against the FP32 oracle step
    of that state rounded onto the same 16-bit grid.  The GPU's collide
    differs from the oracle's by FP32 rounding only (fast reciprocal and
    square root in the 16-bit kernels; the FP32 single-step bar is 2e-6 of
    max|f|), so a value may round to a neighbouring grid point: allowed by
    one grid spacing plus that FP32 bar (which matters where the deviation is
    tiny and BF16's grid is fine), on under 2 % of the values.  A value
    pulled from a wrong neighbour or boundary misses by ~1e-3."""
    say("16-bit steps", storage, case)
    c, mask = _bc_case(case)
    c.storage = {"f16": 1, "bf16": 2}[storage]  # LBM_STORE_F16 / LBM_STORE_BF16
    sh = storage_offsets(c.umax)
    rng = np.random.default_rng(5)
    f = random_state(rng, mask.shape, amp=1e-2)
    omask = np.vectorize(NODE_TO_ORC.get)(mask).astype(np.uint8)
    u_in = O.inlet_profile(c.ny, c.nz, c.umax, c.inlet_amp, c.seed)
    solid = np.isin(mask, (L.NODE_SOLID, L.NODE_SOLID_FORCE))
    keep = ~solid if c.halfway else np.ones_like(solid)
    keep = np.broadcast_to(keep, (19,) + mask.shape)
    with L.Solver(c) as s:
        s.set_geometry(mask)
        s.set_populations(f)
        for t in range(3):
            g0 = s.populations()
            s.step(1)
            g = s.populations()
            fo, Fo = O.run(orc_params(c), g0, 1, mask=omask, u_in=u_in, forces=True)
            q, ulp = on_16bit_grid(fo, storage, sh)
            err = np.abs(g - q)[keep]
            off = err > 0
            tol32 = 2e-6 * np.abs(fo).max()
            say(f"  step {t}: {off.mean():.2e} of values off the oracle's grid point, max {err.max():.2e} "
                f"(bar: one spacing + {tol32:.1e})")
            assert (err <= ulp[keep] * 1.0001 + tol32).all()
            assert off.mean() < 0.02
        if c.force_steps:
            F = s.forces()
            assert np.abs(Fo).max() > 0
            assert np.abs(F[2] - Fo[0]).max() < 1e-5 * np.abs(Fo[0]).max()


def test_bad_mask_code_rejected():
    c = L.default_config()
    c.nx, c.ny, c.nz = 8, 4, 4
    with L.Solver(c) as s:
        bad = np.full((4, 4, 8), L.NODE_GHOST_DYN + 1, np.uint8)
        with pytest.raises(L.LBMError) as e:
            s.set_geometry(bad)
        assert e.value.code == -22
        with pytest.raises(L.LBMError):
            s.step(1)  # not initialised


def test_symmetry_planes_match_oracle():
    """x = 0 and x = nx-1 symmetry planes (xsymmetry_bot/top, im 26/25) with
    halfway walls in y and z, so the planes' edges and corners take their
    special cases: 30 steps of MRT-LES from a random state, GPU vs oracle."""
    say("symmetry planes")
    c = L.default_config()
    c.nx, c.ny, c.nz = 40, 12, 10
    c.bc[0], c.bc[1], c.bc[2] = L.BC_CLAMP, L.BC_WALL, L.BC_WALL
    c.halfway = 1
    c.model, c.omega, c.cs = L.MODEL_MRT_LES, 1.6, 0.01
    mask = np.zeros((c.nz, c.ny, c.nx), np.uint8)
    mask[:, :, 0] = L.NODE_SYM_XMIN
    mask[:, :, -1] = L.NODE_SYM_XMAX
    rng = np.random.default_rng(11)
    f = random_state(rng, mask.shape)
    with L.Solver(c) as s:
        s.set_geometry(mask)
        assert np.array_equal(s.get_mask(), mask)
        s.set_populations(f)
        s.step(30)
        g = s.populations()
    omask = np.vectorize(NODE_TO_ORC.get)(mask).astype(np.uint8)
    ref, _ = O.run(orc_params(c), f, 30, mask=omask)
    assert np.abs(g - ref).max() < 1e-5 * np.abs(ref).max()
