"""Pins for the parts of the oracle that no reference output covers: the
LES-family per-moment relaxation rates, AddForce, the halfway wall, the
x-symmetry planes, the Extrap conditions and the MRT-form initial fill.

Each check comes from mathematics the oracle's code does not restate, so a
plausible mistake (a moment relaxed at the wrong rate, a sign in the force,
a wall half a cell off, a mirrored index) fails it:
  * shear waves in six orientations decay at nu = (1/omega - 1/2)/3, one per
    stress moment (m13, m14, m15 by the axis-aligned waves, m9 and m11 by
    the diagonal ones) -- a moment relaxed at rate 1 instead of omega shows
    as a 4x viscosity at omega = 1.6;
  * each non-conserved moment leaves the collide at m - s_k (m - meq_k), with
    meq_k the collide's own full-relaxation output and s_k the preset's rate;
  * a body force changes only the y momentum, by exactly -dpdy;
  * a halfway-wall Poiseuille channel at the one relaxation time where the
    scheme is exact for a parabola reproduces g/(2 nu) (x+1/2)(H-x-1/2),
    walls at -1/2 and H-1/2;
  * a box with symmetry planes at x = 0 and x = nx-1 streams exactly as a
    periodic box twice as long started mirror-symmetric, except for the
    plane's y-edge statements, whose deviations are listed;
  * the Extrap conditions and the MRT-form fill, through the moments of
    their outputs.
"""
import ctypes

import numpy as np
import pytest

import oracle_lib as O

C = np.array([(0, 0, 0), (1, 0, 0), (0, 1, 0), (-1, 0, 0), (0, -1, 0), (1, 1, 0), (-1, 1, 0), (-1, -1, 0),
              (1, -1, 0), (0, 0, 1), (1, 0, 1), (0, 1, 1), (-1, 0, 1), (0, -1, 1), (0, 0, -1), (1, 0, -1),
              (0, 1, -1), (-1, 0, -1), (0, -1, -1)])
W = np.array([1 / 3] + [1 / 18] * 4 + [1 / 36] * 4 + [1 / 18] + [1 / 36] * 4 + [1 / 18] + [1 / 36] * 4)
# d'Humieres 2002 rates of the reference (MRTtest/DynInlet3.cu:774-783) per moment
DHUM = {1: 1.19, 2: 1.4, 4: 1.2, 6: 1.2, 8: 1.2, 10: 1.4, 12: 1.4, 16: 1.98, 17: 1.98, 18: 1.98}
STRESS = (9, 11, 13, 14, 15)
CONSERVED = (0, 3, 5, 7)


def dhum_params(omega, **kw):
    p = O.params(collision=O.MRT_DHUM, omega=omega, **kw)
    for k, s in DHUM.items():
        p.s_dhum[k] = s
    return p


def rand_pops(rng, n):
    out = np.empty((n, 19))
    for k in range(n):
        out[k] = O.feq_bgk(float(rng.uniform(0.97, 1.03)), *rng.uniform(-0.06, 0.06, 3), dtype=np.float64)
        out[k] += rng.uniform(-2e-3, 2e-3, 19) * W
    return out


def shear_wave_viscosity(p_of, kdir, udir, n=32, t0=40, T=600):
    """Measured kinematic viscosity of a shear wave u = U udir sin(k.x),
    k = 2 pi kdir / n (udir perpendicular to kdir), from its amplitude
    between steps t0 and t0 + T on a periodic box (FP64 oracle)."""
    kdir = np.array(kdir)
    udir = np.array(udir, float) / np.linalg.norm(udir)
    nx, ny, nz = (n if kdir[a] else 1 for a in range(3))
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    k = 2 * np.pi / n
    ph = k * (kdir[0] * x + kdir[1] * y + kdir[2] * z)
    f = np.empty((19, nz, ny, nx))
    for idx in np.ndindex(nz, ny, nx):
        f[(slice(None),) + idx] = O.feq_bgk(1.0, *(1e-3 * udir * np.sin(ph[idx])), dtype=np.float64)
    p = p_of(nx, ny, nz)

    def amp(g):
        _, u, v, w = O.macro(p, g)
        proj = u * udir[0] + v * udir[1] + w * udir[2]
        return (proj * np.sin(ph)).sum() / (np.sin(ph) ** 2).sum()

    f, _ = O.run(p, f, t0)
    a1 = amp(f)
    f, _ = O.run(p, f, T)
    a2 = amp(f)
    return np.log(a1 / a2) / (k * k * (kdir ** 2).sum() * T)


WAVES = [((0, 1, 0), (1, 0, 0)),   # u_x(y): P_xy, m13
         ((0, 0, 1), (0, 1, 0)),   # u_y(z): P_yz, m14
         ((0, 0, 1), (1, 0, 0)),   # u_x(z): P_xz, m15
         ((1, 1, 0), (1, -1, 0)),  # P_xx - P_yy: m9 and m11
         ((0, 1, 1), (0, 1, -1)),  # P_yy - P_zz: m11
         ((1, 0, 1), (1, 0, -1))]  # P_xx - P_zz: m9 and m11


@pytest.mark.parametrize("preset", ["les", "dhum", "bgk"])
@pytest.mark.parametrize("kdir,udir", WAVES)
def test_shear_wave_decays_at_the_viscosity_of_omega(preset, kdir, udir):
    """nu = (1/omega - 1/2)/3 within 1 % (the lattice's O(k^2) dispersion at
    32 nodes per wavelength is 0.3 %), for every stress component."""
    om = 1.6
    mk = {"les": lambda nx, ny, nz: O.params(nx, ny, nz, collision=O.MRT_LES, omega=om),
          "dhum": lambda nx, ny, nz: dhum_params(om, nx=nx, ny=ny, nz=nz),
          "bgk": lambda nx, ny, nz: O.params(nx, ny, nz, collision=O.BGK, omega=om)}[preset]
    nu = (1 / om - 0.5) / 3
    assert abs(shear_wave_viscosity(mk, kdir, udir) / nu - 1) < 0.01


def test_cpu3d_preset_shows_its_rate_one_moments_in_the_shear_waves():
    """The CPU_3D operator (CPU_3D/CPU_3D.cpp:246-247) relaxes m13 and m14 at
    rate 1: the waves carried by P_xy and P_yz decay at nu(omega=1) = 1/6,
    the others at nu(omega)."""
    om = 1.6
    mk = lambda nx, ny, nz: O.params(nx, ny, nz, collision=O.MRT_CPU3D, omega=om)  # noqa: E731
    nu, nu1 = (1 / om - 0.5) / 3, 1 / 6
    got = [shear_wave_viscosity(mk, kd, ud) for kd, ud in WAVES]
    assert abs(got[0] / nu1 - 1) < 0.01 and abs(got[1] / nu1 - 1) < 0.01
    for g in got[2:]:
        assert abs(g / nu - 1) < 0.01


@pytest.mark.parametrize("preset", ["les", "dhum"])
def test_each_moment_relaxes_at_its_preset_rate(preset):
    """At omega = 1.6 every non-conserved moment of the output is
    m - s_k (m - meq_k): s_k = omega on the five stress moments
    (MRTtest/DynInlet3.cu:767-789), 1 (MRT_LES) or the d'Humieres rate
    (MRT_DHUM, :774-783) on the others; meq_k is the collide's own
    output at full relaxation (MRT_LES at omega = 1, whose equilibrium
    moments are pinned against the BGK equilibrium elsewhere).  Conserved
    moments are unchanged."""
    M = O.moment_matrix()
    rng = np.random.default_rng(21)
    om = 1.6
    p = O.params(collision=O.MRT_LES, omega=om) if preset == "les" else dhum_params(om)
    for f in rand_pops(rng, 30):
        m = M @ f
        meq = M @ O.collide(O.params(collision=O.MRT_LES, omega=1.0), f)
        out = M @ O.collide(p, f)
        for k in range(19):
            if k in CONSERVED:
                s = 0.0
            elif k in STRESS:
                s = om
            else:
                s = 1.0 if preset == "les" else DHUM[k]
            assert abs(out[k] - (m[k] - s * (m[k] - meq[k]))) < 1e-13, (preset, k)


@pytest.mark.parametrize("coll", [O.BGK, O.MRT_LES, O.MRT_DHUM])
def test_add_force_changes_only_y_momentum_by_minus_dpdy(coll):
    """AddForce (MRTtest/DynInlet3.cu:229-253): with and without dpdy the
    outputs differ by 3 w_i c_iy dpdy in each direction, so mass and the x
    and z momenta are unchanged and the y momentum moves by exactly -dpdy
    (sum_i 3 w_i c_iy^2 = 1)."""
    rng = np.random.default_rng(22)
    dpdy = 3.7e-5
    for f in rand_pops(rng, 10):
        kw = dict(collision=coll, omega=1.3, cs=0.01)
        a = O.collide(O.params(dpdy=dpdy, **kw), f)
        b = O.collide(O.params(**kw), f)
        d = a - b
        # (1e-16: the rounding of differences of O(0.1) populations; a wrong
        # sign, weight or direction is off by O(dpdy) = 1e-5)
        assert abs(d.sum()) < 1e-16
        j = C.T @ d
        assert abs(j[0]) < 1e-16 and abs(j[2]) < 1e-16 and abs(j[1] + dpdy) < 1e-16
        # the change lies along w_i c_iy: d = -dpdy * 3 w_i c_iy
        assert np.abs(d + 3 * W * C[:, 1] * dpdy).max() < 1e-16


def test_halfway_wall_poiseuille_is_exact_at_lambda_three_sixteenths():
    """Flow along y between halfway walls at the x faces, driven by the body
    force g = -dpdy.  With the ghost moments at rate 1 (MRT_LES) and
    tau = 7/8 the product (tau - 1/2)(1 - 1/2) is 3/16, where link
    bounce-back puts the wall exactly half a cell beyond the last node and
    the parabola is reproduced exactly: the stored momentum (after the
    step's force) is g/(2 nu) (x + 1/2)(H - x - 1/2) + g/2 with
    nu = (tau - 1/2)/3.  A wall off by a fraction of a cell, a wrong
    viscosity or force scale moves it by far more than the 1e-9 bar."""
    H, tau, g = 16, 0.875, 1e-6
    nu = (tau - 0.5) / 3
    p = O.params(H, 1, 1, bc=(O.WALL, O.PERIODIC, O.PERIODIC), collision=O.MRT_LES, halfway=1,
                 omega=1 / tau, dpdy=-g)
    f = np.ascontiguousarray(np.broadcast_to(O.feq_bgk(1.0, 0, 0, 0, dtype=np.float64)[:, None, None, None],
                                             (19, 1, 1, H)))
    f, _ = O.run(p, f, 12000)
    _, u, v, w = O.macro(p, f)
    x = np.arange(H)
    exact = g / (2 * nu) * (x + 0.5) * (H - x - 0.5) + g / 2
    assert np.abs(v[0, 0] - exact).max() < 1e-9 * exact.max()
    assert np.abs(u).max() < 1e-14 and np.abs(w).max() < 1e-14  # sums of O(0.1) values
    # at another tau the same scheme slips: the bar above is not loose
    p.omega = 1 / 0.6
    f2 = np.ascontiguousarray(np.broadcast_to(O.feq_bgk(1.0, 0, 0, 0, dtype=np.float64)[:, None, None, None],
                                              (19, 1, 1, H)))
    f2, _ = O.run(p, f2, 20000)
    nu2 = 0.1 / 3
    exact2 = g / (2 * nu2) * (x + 0.5) * (H - x - 0.5) + g / 2
    assert np.abs(O.macro(p, f2)[2][0, 0] - exact2).max() > 1e-4 * exact2.max()


def _mirror_index(sx, sy, sz):
    return [int(np.flatnonzero((C == C[i] * [sx, sy, sz]).all(1))[0]) for i in range(19)]


def _xsym_vs_mirror(ybc, zbc, collision=O.BGK, omega=0.0, n=(7, 6, 5), seed=1):
    """One step of an n-box with symmetry planes at x = 0 (xsymmetry_bot) and
    x = nx-1 (xsymmetry_top) against a periodic box of 2(nx-1) nodes in x
    whose state is the box's mirrored about both planes.  Returns |diff|
    over the n-box, [19][nz][ny][nx]."""
    N, ny, nz = n
    P = 2 * (N - 1)
    mx = _mirror_index(-1, 1, 1)
    rng = np.random.default_rng(seed)
    g = np.empty((19, nz, ny, N))
    for idx in np.ndindex(nz, ny, N):
        g[(slice(None),) + idx] = O.feq_bgk(1.0, *rng.uniform(-0.05, 0.05, 3), dtype=np.float64)
    g += rng.uniform(-1e-3, 1e-3, g.shape) * W[:, None, None, None]
    for x in (0, N - 1):  # the planes are their own mirror images
        g[:, :, :, x] = 0.5 * (g[:, :, :, x] + g[mx][:, :, :, x])
    full = np.empty((19, nz, ny, P))
    full[:, :, :, :N] = g
    for j in range(N, P):
        full[:, :, :, j] = g[mx][:, :, :, P - j]
    mask = np.zeros((nz, ny, N), np.uint8)
    mask[:, :, 0], mask[:, :, N - 1] = O.SYM_XMIN, O.SYM_XMAX
    kw = dict(collision=collision, omega=omega, halfway=1, cs=0.01 if collision != O.BGK else 0.0)
    a, _ = O.step(O.params(N, ny, nz, bc=(O.CLAMP, ybc, zbc), **kw), g, mask=mask)
    b, _ = O.step(O.params(P, ny, nz, bc=(O.PERIODIC, ybc, zbc), **kw), full)
    return np.abs(a - b[:, :, :, :N])


@pytest.mark.parametrize("ybc,zbc", [(O.PERIODIC, O.PERIODIC), (O.WALL, O.WALL), (O.WALL, O.PERIODIC)])
def test_symmetry_planes_are_mirror_domains_off_the_y_edges(ybc, zbc):
    """Away from the plane's y = 0 and y = ny-1 rows the symmetry planes
    (MRTtest/DynInlet3.cu:1020-1137) stream exactly as the mirrored
    domain -- pure streaming (BGK at omega = 0) and with the MRT-LES
    collide -- whatever the y/z treatment, z edges included."""
    for coll, om in ((O.BGK, 0.0), (O.MRT_LES, 1.6)):
        d = _xsym_vs_mirror(ybc, zbc, coll, om)
        assert d[:, :, 1:-1, :].max() == 0.0, (coll, ybc, zbc)
        assert d[:, :, :, 1:-1].max() == 0.0  # every node off the planes


# Populations (after streaming, before the collide) in which the as-coded
# plane edges depart from the mirrored domain, keyed by (plane, y edge, z
# position); the statements responsible are listed in the design notes.
EDGE_DEVIATIONS = {
    ("bot", "y0", "zi"): [2, 11, 16], ("bot", "y1", "zi"): [4, 13, 18],
    ("top", "y0", "zi"): [2, 5, 6, 11, 16], ("top", "y1", "zi"): [4, 7, 8, 13, 18],
    ("bot", "y0", "z0"): [2, 5, 6, 9, 10, 11, 12, 13, 16], ("bot", "y0", "z1"): [4, 5, 6, 11, 14, 15, 16, 17, 18],
    ("bot", "y1", "z0"): [4, 7, 8, 9, 10, 11, 12, 13, 18], ("bot", "y1", "z1"): [4, 7, 8, 13, 14, 15, 16, 17, 18],
    ("top", "y0", "z0"): [2, 5, 6, 9, 10, 11, 12, 13, 16], ("top", "y0", "z1"): [2, 5, 6, 11, 14, 15, 16, 17, 18],
    ("top", "y1", "z0"): [4, 7, 8, 9, 10, 11, 12, 13, 18], ("top", "y1", "z1"): [4, 7, 8, 13, 14, 15, 16, 17, 18],
}


@pytest.mark.parametrize("ybc,zbc", [(O.PERIODIC, O.PERIODIC), (O.WALL, O.WALL)])
def test_symmetry_plane_edges_depart_from_the_mirror_as_coded(ybc, zbc):
    """At the y edges the reference's extra copies (a y-mirror through the
    node, plus the corner statements) replace what the x-mirrored domain
    streams there; the deviating populations are exactly these, whatever
    the y/z treatment."""
    d = _xsym_vs_mirror(ybc, zbc)
    N, ny, nz = 7, 6, 5
    for (plane, ye, zp), pops in EDGE_DEVIATIONS.items():
        x = 0 if plane == "bot" else N - 1
        y = 0 if ye == "y0" else ny - 1
        zs = {"zi": range(1, nz - 1), "z0": [0], "z1": [nz - 1]}[zp]
        for z in zs:
            got = sorted(int(i) for i in np.flatnonzero(d[:, z, y, x] > 0))
            assert got == pops, (plane, ye, zp, z, got)


def test_extrap_conditions_set_only_their_prescribed_moments():
    """West_Extrap (MRTtest/DynInlet3.cu:973-1017) returns populations whose
    moments equal the input's except j = (u_in, 0, 0); East_Extrap
    (:926-971) except rho = rho_out (all through the pinned moment matrix)."""
    M = O.moment_matrix()
    rng = np.random.default_rng(23)
    for f in rand_pops(rng, 20):
        m = M @ f
        a = f.copy()
        O.lib().orc_west_extrap_f64(a, 0.047)
        ma = M @ a
        exp = m.copy()
        exp[3], exp[5], exp[7] = 0.047, 0.0, 0.0
        assert np.abs(ma - exp).max() < 1e-14
        b = f.copy()
        O.lib().orc_east_extrap_f64(b, 1.013)
        mb = M @ b
        exp = m.copy()
        exp[0] = 1.013
        assert np.abs(mb - exp).max() < 1e-14


def test_mrt_fill_has_the_incompressible_dhumieres_equilibrium_moments():
    """initialize's fill (mrt_meq then InvertMoments, MRTtest/DynInlet3.cu:
    329-345, 1186-1216): rho, j and the five stress moments equal those of
    the BGK equilibrium (the same Navier-Stokes stress), the energy moment
    e = -11 rho + 19 j^2 as in the BGK equilibrium's projection, the energy
    square eps = 475/63 j^2 (d'Humieres et al. 2002, w_eps = 0,
    w_epsj = -475/63), q = -2/3 j, and 3pi_xx, pi_ww, m16-18 zero."""
    M = O.moment_matrix()
    rng = np.random.default_rng(24)
    for _ in range(20):
        rho = float(rng.uniform(0.97, 1.03))
        u = rng.uniform(-0.06, 0.06, 3)
        mm = M @ O.feq_mrt(rho, *u, dtype=np.float64)
        mb = M @ O.feq_bgk(rho, *u, dtype=np.float64)
        j2 = (u * u).sum()
        for k in (0, 3, 5, 7) + STRESS + (1,):
            assert abs(mm[k] - mb[k]) < 1e-14, k
        assert abs(mm[2] - 475 / 63 * j2) < 1e-14
        assert np.abs(mm[[4, 6, 8]] + 2 / 3 * u).max() < 1e-15
        assert np.abs(mm[[10, 12, 16, 17, 18]]).max() < 1e-15


def test_smagorinsky_rate_reaches_the_collide():
    """With CS > 0 the stress moments relax at the tau(Q) law's rate, not
    at omega: the LES collide equals the no-LES collide at the rate the
    law gives for that input (MRTtest/DynInlet3.cu:722-735, 759-789)."""
    rng = np.random.default_rng(25)
    for f in rand_pops(rng, 10):
        p = O.params(collision=O.MRT_LES, omega=1.6, cs=0.01)
        om = ctypes.c_double()
        O.lib().orc_les_omega_f64(ctypes.byref(p), f, ctypes.byref(om), None)
        assert om.value < 1.6 - 1e-6
        a = O.collide(p, f)
        b = O.collide(O.params(collision=O.MRT_LES, omega=om.value), f)
        assert np.abs(a - b).max() < 1e-16
