"""Pins the oracle (oracle/liboracle.so) before it is trusted as the checker.

Each check is chosen so that a plausible mistake in the restatement (a
dropped term, a wrong sign or index, a transposed operand) fails it:
  * the reference's own CPU program, CPU_3D/CPU_3D.cpp, through the golden
    fixtures it produced (node functions on seeded inputs and a full 100-step
    run of its own case), and directly through oracle/_ref/ when that is built;
  * closed forms: orthogonality and norms of the moment matrix, conservation,
    equilibrium fixed points, the Smagorinsky tau(Q) law, the streaming
    permutation by brute force, Taylor-Green decay.
"""
import os

import numpy as np
import pytest

import oracle_lib as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
C = np.array([(0, 0, 0), (1, 0, 0), (0, 1, 0), (-1, 0, 0), (0, -1, 0), (1, 1, 0), (-1, 1, 0), (-1, -1, 0),
              (1, -1, 0), (0, 0, 1), (1, 0, 1), (0, 1, 1), (-1, 0, 1), (0, -1, 1), (0, 0, -1), (1, 0, -1),
              (0, 1, -1), (-1, 0, -1), (0, -1, -1)])


def rand_pops(rng, n, amp=2e-3, dtype=np.float32):
    out = np.empty((n, 19), dtype)
    for k in range(n):
        u = rng.uniform(-0.06, 0.06, 3)
        out[k] = O.feq_bgk(float(rng.uniform(0.97, 1.03)), *u, dtype=dtype) + rng.uniform(-amp, amp, 19)
    return out


def test_moment_matrix_orthogonal_with_reference_norms():
    """M M^T = diag(19, 2394, 252, 10, 40, ...): these norms fix InvertMoments'
    coefficients 1/19, -11/2394, 1/10, 1/40 ... (MRTtest/DynInlet3.cu:409-435)."""
    M = O.moment_matrix()
    G = M @ M.T
    norms = [19, 2394, 252, 10, 40, 10, 40, 10, 40, 36, 72, 12, 24, 4, 4, 4, 8, 8, 8]
    assert np.array_equal(G, np.diag(norms).astype(float))
    # the reference's f0 row of InvertMoments: 1/19 m0 - 30/2394 m1 + 12/252 m2
    Minv = M.T / np.array(norms)
    assert np.allclose(Minv[0, :3], [0.052631579, -0.012531328, 0.047619048], atol=1e-9)
    assert np.allclose(Minv[5, [1, 2, 4, 13, 16, 17]], [0.0033416876, 0.003968254, 0.025, 0.25, 0.125, -0.125], atol=1e-9)


def test_node_functions_match_reference_fixtures():
    """bgk_collide, mrt_collide and DirichletWest_Regularized against the
    reference's own outputs (tests/golden/cpu3d_node_vectors.npz)."""
    g = np.load(os.path.join(GOLD, "cpu3d_node_vectors.npz"))
    fin, om = g["fin"], g["omega"]
    for name, coll in (("bgk", O.BGK), ("mrt", O.MRT_CPU3D)):
        got = np.stack([O.collide(O.params(collision=coll, omega=float(om[k])), fin[k]) for k in range(len(fin))])
        err = np.abs(got - g[name]).max()
        assert err < 3e-7, (name, err)
    lid = fin.copy()
    for k in range(len(fin)):
        O.lib().orc_lid_west_reg_f32(lid[k], np.float32(g["umax"][0]))
    assert np.abs(lid - g["lid"]).max() < 1e-7


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (reference tree absent)")
def test_node_functions_match_reference_build():
    rng = np.random.default_rng(7)
    r = O.ref()
    f = rand_pops(rng, 200)
    for k in range(len(f)):
        om = np.float32(rng.uniform(0.6, 1.9))
        a = f[k].copy()
        r.ref_mrt_collide(a, om)
        b = O.collide(O.params(collision=O.MRT_CPU3D, omega=float(om)), f[k])
        assert np.abs(a - b).max() < 3e-7


def test_full_run_matches_reference_program():
    """The reference program run as is (251x43x43, 100 steps, output.dat at
    6 significant digits) against the oracle on the same case."""
    g = np.load(os.path.join(GOLD, "cpu3d_t100.npz"))
    nx, ny, nz, tmax = (int(d) for d in g["dims"])
    umax, omega = float(g["consts"][0]), float(g["consts"][3])
    mask = cpu3d_mask(nx, ny, nz)
    f0 = O.feq_bgk(1.0, umax, 0.0, 0.0)
    f = np.ascontiguousarray(np.broadcast_to(f0[:, None, None, None], (19, nz, ny, nx)))
    p = O.params(nx, ny, nz, bc=(O.CLAMP,) * 3, collision=O.MRT_CPU3D, omega=omega, u_lid=umax)
    fo, _ = O.run(p, f, tmax, mask=mask)
    rho, u, v, w = O.macro(p, fo)
    zs = g["zs"]
    # 1e-5 relative to the field's scale (output.dat itself holds 6 significant
    # digits, i.e. up to 5e-6 relative rounding)
    for got, ref in ((rho[zs], g["rho"]), (u[zs], g["u"]), (v[zs], g["v"])):
        assert np.abs(got - ref).max() <= 1e-5 * np.abs(ref).max()
    assert np.allclose(rho.sum(axis=(1, 2)), g["plane_rho"], rtol=1e-6)
    assert np.allclose(u.sum(axis=(1, 2)), g["plane_u"], rtol=1e-4, atol=1e-4)


def cpu3d_mask(nx, ny, nz, ox=50.5, oy=20.5, r=5.0):
    """ImageFcn of CPU_3D/CPU_3D.cpp:39-75, evaluated in float like the reference."""
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    dx = x.astype(np.float32) - np.float32(ox)
    dy = y.astype(np.float32) - np.float32(oy)
    m = np.zeros((nz, ny, nx), np.uint8)
    wall = (x == nx - 1) | (y == 0) | (y == ny - 1) | (z == 0) | (z == nz - 1)
    m[wall] = O.SOLID
    m[x == 0] = O.LID_WEST_REG
    m[(dx * dx + dy * dy) < np.float32(r) * np.float32(r)] = O.SOLID_FORCE
    return m


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (reference tree absent)")
def test_cpu3d_mask_matches_reference_imagefcn():
    r = O.ref()
    m = cpu3d_mask(251, 43, 43)
    code = {0: O.FLUID, 1: O.SOLID, 10: O.SOLID_FORCE, 53: O.LID_WEST_REG}
    for z in (0, 1, 21, 42):
        for y in range(43):
            for x in range(251):
                assert m[z, y, x] == code[r.ref_image(x, y, z)]


@pytest.mark.parametrize("coll", [O.BGK, O.MRT_CPU3D, O.MRT_LES, O.MRT_DHUM])
@pytest.mark.parametrize("cs", [0.0, 0.01])
def test_collide_conserves_mass_and_momentum(coll, cs):
    if coll == O.MRT_CPU3D and cs:
        pytest.skip("CPU_3D operator has no LES")
    rng = np.random.default_rng(3)
    for f in rand_pops(rng, 50, dtype=np.float64):
        g = O.collide(O.params(collision=coll, omega=1.7, cs=cs), f)
        assert abs(g.sum() - f.sum()) < 1e-14
        assert np.abs(C.T @ g - C.T @ f).max() < 1e-14


def test_equilibrium_is_fixed_point():
    """feq_BGK is invariant under BGK at any omega; the LES-family MRT leaves its
    own equilibrium invariant (bgk_meq with meq10 = -meq9/6 as coded)."""
    rng = np.random.default_rng(4)
    for _ in range(20):
        u = rng.uniform(-0.08, 0.08, 3)
        feq = O.feq_bgk(1.01, *u, dtype=np.float64)
        g = O.collide(O.params(collision=O.BGK, omega=1.3), feq)
        assert np.abs(g - feq).max() < 1e-15
        # MRT_LES at omega=1 relaxes every moment fully: output moments are bgk_meq
        f = rand_pops(rng, 1, dtype=np.float64)[0]
        out = O.collide(O.params(collision=O.MRT_LES, omega=1.0), f)
        m = O.moment_matrix() @ out
        again = O.collide(O.params(collision=O.MRT_LES, omega=1.7), out)
        assert np.abs(again - out).max() < 1e-14
        # and it differs from BGK only along the 3pi_xx direction (k = 10)
        bgk = O.collide(O.params(collision=O.BGK, omega=1.0), f)
        d = O.moment_matrix() @ (out - bgk)
        assert np.abs(np.delete(d, 10)).max() < 1e-13
        m9 = m[9]
        assert abs(d[10] - (-m9 / 6 + m9 / 2)) < 1e-13


def test_mrt_with_all_rates_equal_is_bgk_up_to_pixx():
    """MRT_DHUM with every rate = omega equals BGK except along m10 (where the
    reference's meq10 differs from the BGK projection)."""
    rng = np.random.default_rng(5)
    M = O.moment_matrix()
    for f in rand_pops(rng, 20, dtype=np.float64):
        p = O.params(collision=O.MRT_DHUM, omega=1.4)
        for k in range(19):
            p.s_dhum[k] = 1.4
        a = O.collide(p, f)
        b = O.collide(O.params(collision=O.BGK, omega=1.4), f)
        d = M @ (a - b)
        assert np.abs(np.delete(d, 10)).max() < 1e-13


def test_smagorinsky_tau_law():
    """Q = 0 gives tau0; otherwise tau = tau0/2 + sqrt(tau0^2 + 18 sqrt2 CS Q)/2
    with Q = |Pi_neq| (MRTtest/DynInlet3.cu:722-734) computed here from the
    definition Pi_neq = sum c c (f - feq)."""
    import ctypes
    rng = np.random.default_rng(6)
    p = O.params(collision=O.MRT_LES, omega=1.6, cs=0.01)
    om = ctypes.c_double()
    q = ctypes.c_double()
    feq = O.feq_bgk(1.0, 0.03, -0.01, 0.02, dtype=np.float64)
    O.lib().orc_les_omega_f64(ctypes.byref(p), feq, ctypes.byref(om), ctypes.byref(q))
    assert abs(q.value) < 1e-15 and abs(om.value - 1.6) < 1e-14
    for f in rand_pops(rng, 20, dtype=np.float64):
        rho, j = f.sum(), C.T @ f
        feq = O.feq_bgk(rho, *j, dtype=np.float64)
        pi = np.einsum("ia,ib,i->ab", C, C, f - feq)
        Q = np.sqrt((pi * pi).sum())
        tau0 = 1 / 1.6
        tau = 0.5 * tau0 + 0.5 * np.sqrt(tau0 ** 2 + 18 * np.sqrt(2) * 0.01 * Q)
        O.lib().orc_les_omega_f64(ctypes.byref(p), f, ctypes.byref(om), ctypes.byref(q))
        assert abs(q.value - Q) < 1e-15 and abs(om.value - 1 / tau) < 1e-13


def test_streaming_permutation_brute_force():
    """With omega = 0 (no collision) one periodic step moves f_i by c_i exactly."""
    n = (4, 5, 3)
    rng = np.random.default_rng(8)
    f = rng.uniform(0, 1, (19, n[2], n[1], n[0]))
    p = O.params(*n, bc=(O.PERIODIC,) * 3, collision=O.BGK, omega=0.0)
    g, _ = O.step(p, f)
    for i in range(19):
        cx, cy, cz = C[i]
        assert np.array_equal(g[i], np.roll(f[i], shift=(cz, cy, cx), axis=(0, 1, 2)))


def test_clamp_and_wall_streaming():
    n = (4, 3, 3)
    rng = np.random.default_rng(9)
    f = rng.uniform(0, 1, (19, n[2], n[1], n[0]))
    g, _ = O.step(O.params(*n, bc=(O.CLAMP, O.WALL, O.PERIODIC), collision=O.BGK, omega=0.0), f)
    opp = [0, 3, 4, 1, 2, 7, 8, 5, 6, 14, 17, 18, 15, 16, 9, 12, 13, 10, 11]
    for i in range(19):
        cx, cy, cz = C[i]
        for z in range(n[2]):
            for y in range(n[1]):
                for x in range(n[0]):
                    sy = y - cy
                    if sy < 0 or sy >= n[1]:
                        exp = f[opp[i], z, y, x]
                    else:
                        sx = min(max(x - cx, 0), n[0] - 1)
                        exp = f[i, (z - cz) % n[2], sy, sx]
                    assert g[i, z, y, x] == exp


def test_halfway_closed_box_conserves_mass():
    n = (6, 5, 4)
    rng = np.random.default_rng(10)
    f = np.stack([O.feq_bgk(1.0, *rng.uniform(-0.03, 0.03, 3), dtype=np.float64) for _ in range(np.prod(n))], 1)
    f = f.reshape(19, n[2], n[1], n[0]) + rng.uniform(0, 1e-3, (19, n[2], n[1], n[0]))
    mask = np.zeros((n[2], n[1], n[0]), np.uint8)
    mask[1:3, 2, 2] = O.SOLID_FORCE
    p = O.params(*n, bc=(O.WALL,) * 3, collision=O.MRT_LES, halfway=1, omega=1.2, cs=0.01)
    fluid = mask == 0
    m0 = f[:, fluid].sum()
    g, _ = O.run(p, f, 20, mask=mask)
    assert abs(g[:, fluid].sum() - m0) < 1e-11 * m0


def test_extrap_conditions_keep_uniform_equilibrium():
    """A uniform BGK equilibrium with u = u_in and rho = 1 is returned unchanged
    by West_Extrap(u_in) and East_Extrap(1)."""
    feq = O.feq_bgk(1.0, 0.05, 0.0, 0.0, dtype=np.float64)
    a = feq.copy()
    O.lib().orc_west_extrap_f64(a, 0.05)
    b = feq.copy()
    O.lib().orc_east_extrap_f64(b, 1.0)
    assert np.abs(a - feq).max() < 1e-16 and np.abs(b - feq).max() < 1e-15
    f = feq + np.linspace(-1e-3, 1e-3, 19)
    a = f.copy()
    O.lib().orc_west_extrap_f64(a, 0.04)
    j = C.T @ a
    assert abs(a.sum() - f.sum()) < 1e-15 and abs(j[0] - 0.04) < 1e-15 and abs(j[1]) < 1e-15


def test_taylor_green_decay_matches_analytic():
    """z-extruded 2-D Taylor-Green vortex: KE(t) = KE(0) exp(-4 nu k^2 t),
    nu = (tau - 1/2)/3 (an exact Navier-Stokes solution)."""
    n, tau, u0, T = 32, 0.6, 0.03, 200
    k = 2 * np.pi / n
    x = np.arange(n)
    X, Y = np.meshgrid(x, x, indexing="xy")
    u = u0 * np.cos(k * X) * np.sin(k * Y)
    v = -u0 * np.sin(k * X) * np.cos(k * Y)
    f = np.empty((19, 2, n, n))
    for yy in range(n):
        for xx in range(n):
            f[:, :, yy, xx] = O.feq_bgk(1.0, u[yy, xx], v[yy, xx], 0.0, dtype=np.float64)[:, None]
    p = O.params(n, n, 2, collision=O.BGK, omega=1 / tau)
    ke0 = (u ** 2 + v ** 2).sum()
    g, _ = O.run(p, f, T)
    _, uu, vv, _ = O.macro(p, g)
    ke = (uu[0] ** 2 + vv[0] ** 2).sum()
    nu = (tau - 0.5) / 3
    assert abs(ke / ke0 / np.exp(-4 * nu * k * k * T) - 1) < 0.02


def test_inlet_profile_hash_properties():
    prof = O.inlet_profile(16, 8, 0.05, 1e-3, 0)
    assert prof.dtype == np.float32 and prof.shape == (128,)
    assert np.all(np.abs(prof - np.float32(0.05)) <= np.float32(1e-3))
    assert len(np.unique(prof)) > 120
    assert not np.array_equal(prof, O.inlet_profile(16, 8, 0.05, 1e-3, 1))
    # r = (h >> 8) 2^-24, so 2r-1 is exact and the table is a pure function of the hash
    h = np.array([O.lib().orc_hash_u32(0, c) for c in range(128)], np.uint64)
    r = (h >> np.uint64(8)).astype(np.float32) * np.float32(1 / 16777216)
    t = np.float32(2) * r - np.float32(1)
    assert np.array_equal(prof, np.float32(0.05) + np.float32(1e-3) * t)


def test_stats_update_closed_forms():
    """Running statistics restatement: unit weight gives the arithmetic mean
    of the sampled momenta, then the mean squared deviation from that frozen
    mean; a constant sequence stays constant under the fine-level weight 1/2;
    solid nodes are left alone."""
    p = O.params(2, 1, 1)
    rng = np.random.default_rng(4)
    us = rng.normal(size=(10, 3)).astype(np.float32)
    stat = np.zeros((6, 1, 1, 2), np.float32)
    mask = np.array([O.FLUID, O.SOLID], np.uint8).reshape(1, 1, 2)

    def pops(u):
        f = np.zeros((19, 1, 1, 2), np.float32)
        f[1], f[2], f[9] = u[0], u[1], u[2]  # c_1 = +x, c_2 = +y, c_9 = +z
        return f

    for t in range(6):
        O.stats_update(p, pops(us[t]), stat, 1, float(t), 1.0, mask=mask)
    mean = us[:6].astype(np.float64).mean(0)
    assert np.allclose(stat[:3, 0, 0, 0], mean, rtol=1e-5, atol=1e-6)
    for t in range(6, 10):
        O.stats_update(p, pops(us[t]), stat, 2, float(t - 6), 1.0, mask=mask)
    dev = ((us[6:].astype(np.float64) - stat[:3, 0, 0, 0].astype(np.float64)) ** 2).mean(0)
    assert np.allclose(stat[3:, 0, 0, 0], dev, rtol=1e-5)
    assert np.all(stat[:, 0, 0, 1] == 0)
    st2 = np.zeros_like(stat)
    c = np.array([0.25, -0.5, 0.125], np.float32)
    for n in range(5):
        for _ in range(2):
            O.stats_update(p, pops(c), st2, 1, float(n), 0.5)
    assert np.array_equal(st2[:3, 0, 0, 0], c)


def test_symmetry_plane_is_a_mirror_inside_the_plane():
    """Away from the plane's edges, xsymmetry_bot/top make the populations
    entering through the plane the mirror images (c_x -> -c_x) of the ones
    leaving it, leave every other population alone, and so zero the normal
    momentum.  (The edge and corner statements of the reference are kept as
    coded; no reference output pins them.)"""
    rng = np.random.default_rng(3)
    C = np.array([(0, 0, 0), (1, 0, 0), (0, 1, 0), (-1, 0, 0), (0, -1, 0), (1, 1, 0), (-1, 1, 0), (-1, -1, 0),
                  (1, -1, 0), (0, 0, 1), (1, 0, 1), (0, 1, 1), (-1, 0, 1), (0, -1, 1), (0, 0, -1), (1, 0, -1),
                  (0, 1, -1), (-1, 0, -1), (0, -1, -1)])
    mirror = [int(np.flatnonzero((C == C[i] * [-1, 1, 1]).all(1))[0]) for i in range(19)]
    for side, into in (("bot", 1), ("top", -1)):
        f = rng.uniform(0.01, 0.1, 19)
        g = O.xsym(f, side, 5, 4, 12, 10)
        for i in range(19):
            if C[i][0] == into:
                assert g[i] == g[mirror[i]] == f[mirror[i]]
            else:
                assert g[i] == f[i]
        assert abs((C[:, 0] * g).sum()) < 1e-15


def test_row_limited_body_force():
    """dyn_y1 (MRTtest/DynInlet3.cu:1744, "if(y>DYNY1) dpdy = 0.f"): rows up to
    dyn_y1 step exactly as with the force everywhere, rows past it exactly
    as with no force."""
    rng = np.random.default_rng(3)
    nx, ny, nz = 8, 10, 6
    f = rand_pops(rng, nx * ny * nz).T.reshape(19, nz, ny, nx).copy()
    kw = dict(collision=O.MRT_LES, omega=1.6, cs=0.01)
    full, _ = O.step(O.params(nx, ny, nz, dpdy=1e-4, **kw), f)
    none, _ = O.step(O.params(nx, ny, nz, dpdy=0.0, **kw), f)
    lim, _ = O.step(O.params(nx, ny, nz, dpdy=1e-4, dyn_y1=4, **kw), f)
    assert np.array_equal(lim[:, :, :5], full[:, :, :5])
    assert np.array_equal(lim[:, :, 5:], none[:, :, 5:])
    assert not np.array_equal(full, none)


def test_plane_v_sum_against_momentum_definition():
    """AverageV's restated sum equals sum over the plane of sum_i c_iy f_i with
    c_y from the velocity table (FP64), so the ten hard-coded indices and
    signs are pinned."""
    rng = np.random.default_rng(4)
    f = rng.standard_normal((19, 5, 7, 9)).astype(np.float32)
    for y in (0, 3, 6):
        ref = float(np.einsum("i,izx->", C[:, 1].astype(np.float64), f[:, :, y, :].astype(np.float64)))
        assert abs(O.plane_v_sum(f, y) - ref) < 1e-5 * np.abs(f).sum()


def test_inlet_control_law():
    """P-controller (MRTtest/DynInlet3.cu:3159-3165): at Av_V = UMAX dpdy is
    held; at Av_V = 2 UMAX it moves by KP |DPDY|; KP = 0 holds it."""
    nx, nz, umax = 10, 4, 0.05
    plane = (nx - 2) * nz
    av, d = O.inlet_control_update(2e-5, umax * plane, nx, nz, umax, 0.5, 2e-5)
    assert abs(av - umax) < 1e-8 and d == np.float32(2e-5)
    av, d = O.inlet_control_update(2e-5, 2 * umax * plane, nx, nz, umax, 0.5, -2e-5)
    assert abs(d - (2e-5 + 0.5 * 2e-5)) < 1e-11
    av, d = O.inlet_control_update(3e-5, 2 * umax * plane, nx, nz, umax, 0.0, 2e-5)
    assert d == np.float32(3e-5)
