"""Pins the oracle's coarse-fine interface restatement (oracle/refine_impl.c,
following MRTtest/DynInlet3.cu:367-595 and :2202-2560) with closed forms:
  * a uniform equilibrium state is reproduced exactly by CF and FC;
  * a linear velocity field is reproduced exactly (trilinear weights, and the
    ORDER-2 correction vanishes for linear fields: every term is a second
    difference over the 8 corners);
  * moments linear in time give the exact midpoint at alpha = 1/2;
  * with CS = 0 the LES-aware rescale reduces to the plain
    SF_cf = LRFACTOR omega_c (1 - omega_f) / ((1 - omega_c) omega_f) of main()
    (MRTtest/DynInlet3.cu:2878), and FC's SF to its inverse;
  * StrainRate(m) = -3/2 (Pi_neq) (omega = 1 as coded) against the definition.
The float instance and the double instance (the FP64 yardstick of the refined
path, the same source compiled in double) are both held to these, and the two
agree on a refined run to float rounding.  The double instance keeps the
reference's decimal constants (-0.026315789 for -1/38 in StrainRate,
0.666666667, 7.53968254), so its closed-form residuals are set by their
truncation (1e-12 to 1e-10 here), not by double rounding.
"""
import numpy as np
import pytest

import oracle_lib as O

C = np.array([(0, 0, 0), (1, 0, 0), (0, 1, 0), (-1, 0, 0), (0, -1, 0), (1, 1, 0), (-1, 1, 0), (-1, -1, 0),
              (1, -1, 0), (0, 0, 1), (1, 0, 1), (0, 1, 1), (-1, 0, 1), (0, -1, 1), (0, 0, -1), (1, 0, -1),
              (0, 1, -1), (-1, 0, -1), (0, -1, -1)])


DTYPES = [np.float32, np.float64]
SUF = {np.float32: "f32", np.float64: "f64"}


def mrt_eq(rho, u, v, w, dtype=np.float32):
    """The fixed point of the LES-family collide: M^-1 bgk_meq (BGK feq with
    3pixx = -meq9/6), which CF/FC rebuild with eq_kind = EQ_BGK_MEQ."""
    f = O.feq_bgk(rho, u, v, w, dtype=np.float64)
    m9 = 2 * u * u - (v * v + w * w)
    col10 = O.moment_matrix()[10] / 72.0
    return (f + (-m9 / 6 + m9 / 2) * col10).astype(dtype)


def coarse_field(n, fn, dtype=np.float32):
    nx, ny, nz = n
    f = np.empty((19, nz, ny, nx), dtype)
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                f[:, z, y, x] = mrt_eq(*fn(x, y, z), dtype=dtype)
    return f


def test_phys_moments_and_strain_rate():
    rng = np.random.default_rng(1)
    L = O._bind_refine()
    M = O.moment_matrix()
    for _ in range(10):
        f = (O.feq_bgk(1.0, *rng.uniform(-0.05, 0.05, 3)) + rng.uniform(-1e-3, 1e-3, 19)).astype(np.float32)
        mom = np.zeros(9, np.float32)
        L.orc_phys_moments_f32(f, mom)
        m = M @ f.astype(np.float64)
        assert np.allclose(mom, m[[0, 3, 5, 7, 9, 11, 13, 14, 15]], atol=2e-7)
        S = np.zeros(6, np.float32)
        L.orc_strain_rate_f32(mom, S)
        j = C.T @ f.astype(np.float64)
        # Pi_neq deviatoric parts from the definition, with the MRT (mrt_meq) equilibrium
        P = np.einsum("ia,ib,i->ab", C, C, f.astype(np.float64))
        Pneq = P - np.outer(j, j)
        tr = np.trace(Pneq)
        dev = Pneq - tr / 3 * np.eye(3)
        # S = -1.5 * (deviatoric Pi_neq) in the reference's scaling
        assert np.allclose(S[[0, 1, 2]], -1.5 * np.diag(dev), atol=3e-7)
        assert np.allclose(S[[3, 4, 5]], -1.5 * np.array([Pneq[0, 1], Pneq[1, 2], Pneq[0, 2]]), atol=3e-7)


def setup(n=(20, 16, 16), a=(3, 2, 2), nf=(16, 12, 12), omega=1.6, fn=None, dtype=np.float32):
    pc = O.params(*n, bc=(O.CLAMP,) * 3, collision=O.MRT_LES, omega=omega)
    r = O.refine(a, nf, omega)
    fc = coarse_field(n, fn, dtype)
    return pc, r, fc


@pytest.mark.parametrize("dtype", DTYPES)
def test_cf_fc_reproduce_uniform_and_linear_fields(dtype):
    # the linear field's u u terms are quadratic in position, which the
    # trilinear moments miss by up to 9e-8 even in double; the uniform field is
    # exact up to the constants' truncation
    def lin(x, y, z):
        return (1.0 + 1e-4 * x, 0.04 + 1e-3 * y, -0.01 + 5e-4 * z, 0.002 * x - 0.001 * z)

    def lin_rho1(x, y, z):
        rho, u, v, w = lin(x, y, z)
        return (1.0, u, v, w)

    for fn, tol64 in ((lambda x, y, z: (1.01, 0.03, -0.01, 0.02), 3e-12), (lin_rho1, 1.5e-7)):
        tol = 3e-7 if dtype == np.float32 else tol64
        pc, r, fc = setup(fn=fn, dtype=dtype)
        M = O.extract(pc, r, fc)
        assert M.dtype == dtype
        nf = tuple(r.nf)
        ff = np.zeros((19, nf[2], nf[1], nf[0]), dtype)
        out = O.interp_cf(r, M, M, 0.0, ff)
        for k in range(nf[2]):
            for j in range(nf[1]):
                for i in range(nf[0]):
                    frame = not (2 <= i < nf[0] - 2 and 2 <= j < nf[1] - 2 and 2 <= k < nf[2] - 2)
                    if not frame:
                        assert np.all(out[:, k, j, i] == 0)
                        continue
                    X, Y, Z = r.a[0] + 0.25 + 0.5 * i, r.a[1] + 0.25 + 0.5 * j, r.a[2] + 0.25 + 0.5 * k
                    exp = mrt_eq(*fn(X, Y, Z), dtype=dtype)
                    assert np.abs(out[:, k, j, i] - exp).max() < tol
        # FC of the exact fine field gives the coarse equilibrium back
        fine = np.empty_like(ff)
        for k in range(nf[2]):
            for j in range(nf[1]):
                for i in range(nf[0]):
                    fine[:, k, j, i] = mrt_eq(*fn(r.a[0] + 0.25 + 0.5 * i, r.a[1] + 0.25 + 0.5 * j,
                                                  r.a[2] + 0.25 + 0.5 * k), dtype=dtype)
        back = O.interp_fc(pc, r, fine, np.zeros_like(fc))
        for z in range(r.a[2] + 2, r.a[2] + nf[2] // 2 - 1):
            for y in range(r.a[1] + 2, r.a[1] + nf[1] // 2 - 1):
                for x in range(r.a[0] + 2, r.a[0] + nf[0] // 2 - 1):
                    assert np.abs(back[:, z, y, x] - fc[:, z, y, x]).max() < tol
        assert np.all(back[:, :, :, : r.a[0] + 2] == 0)


def _moments_with_strain(u, S):
    """Physical moments (rho = 1, j = u) whose StrainRate is the symmetric,
    traceless S: the deviatoric non-equilibrium stress -S/1.5 written into
    3pixx, piww, pixy, piyz, pixz on top of the equilibrium's u u parts."""
    P = -S / 1.5
    ux, uy, uz = u
    return np.array([1.0, ux, uy, uz,
                     2 * P[0, 0] - P[1, 1] - P[2, 2] + 2 * ux * ux - (uy * uy + uz * uz),
                     P[1, 1] - P[2, 2] + uy * uy - uz * uz,
                     P[0, 1] + ux * uy, P[1, 2] + uy * uz, P[0, 2] + ux * uz])


def _order2_at(vel, sgrad, pts, dtype=np.float32):
    """Trilinear interpolation of the 8 corner moments of the unit cube plus
    the ORDER-2 correction, at fine positions pts; returns (got, exact) j."""
    L = O._bind_refine()
    corners = [np.array([k & 1, (k >> 1) & 1, (k >> 2) & 1], float) for k in range(8)]
    c = np.ascontiguousarray([_moments_with_strain(vel(X), sgrad(X)) for X in corners], dtype)
    got, exact = [], []
    for p in pts:
        w = np.array([np.prod([p[a] if X[a] else 1 - p[a] for a in range(3)]) for X in corners])
        mom = (w @ c.astype(np.float64)).astype(dtype)
        getattr(L, "orc_order2_correction_" + SUF[dtype])(c, mom, *(float(v) for v in p), 0)
        got.append(mom[1:4].copy())
        exact.append(vel(np.asarray(p, float)))
    return np.array(got), np.array(exact)


@pytest.mark.parametrize("dtype", DTYPES)
def test_order2_reproduces_divergence_free_quadratic_fields(dtype):
    """ORDER 2 (MRTtest/DynInlet3.cu:2312-2357): the trilinear momentum plus
    the correction built from the corners' strain rates and velocity
    differences reproduces every divergence-free quadratic velocity field
    exactly when the corners carry its exact strain rate.  Each field below
    exercises one coefficient (cy, cz, dx, dz, ex, ey through the off-
    diagonal rates and the velocity differences; cx, dy, ez through the
    diagonal ones), so a wrong sign or index in any of them fails here."""
    a = 0.02
    e = np.zeros((3, 3))

    def grad_field(Gfn):
        return lambda X: 0.5 * (Gfn(X) + Gfn(X).T)

    def G(entries):
        def f(X):
            g = e.copy()
            for (i, j), fn in entries.items():
                g[i, j] = fn(X)
            return g
        return f

    fields = [
        (lambda X: np.array([a * X[1] ** 2, 0, 0]), G({(0, 1): lambda X: 2 * a * X[1]})),
        (lambda X: np.array([a * X[2] ** 2, 0, 0]), G({(0, 2): lambda X: 2 * a * X[2]})),
        (lambda X: np.array([0, a * X[0] ** 2, 0]), G({(1, 0): lambda X: 2 * a * X[0]})),
        (lambda X: np.array([0, a * X[2] ** 2, 0]), G({(1, 2): lambda X: 2 * a * X[2]})),
        (lambda X: np.array([0, 0, a * X[0] ** 2]), G({(2, 0): lambda X: 2 * a * X[0]})),
        (lambda X: np.array([0, 0, a * X[1] ** 2]), G({(2, 1): lambda X: 2 * a * X[1]})),
        (lambda X: np.array([a * X[0] ** 2, -2 * a * X[0] * X[1], 0]),
         G({(0, 0): lambda X: 2 * a * X[0], (1, 0): lambda X: -2 * a * X[1], (1, 1): lambda X: -2 * a * X[0]})),
        (lambda X: np.array([0, a * X[1] ** 2, -2 * a * X[1] * X[2]]),
         G({(1, 1): lambda X: 2 * a * X[1], (2, 1): lambda X: -2 * a * X[2], (2, 2): lambda X: -2 * a * X[1]})),
        (lambda X: np.array([-2 * a * X[0] * X[2], 0, a * X[2] ** 2]),
         G({(2, 2): lambda X: 2 * a * X[2], (0, 2): lambda X: -2 * a * X[0], (0, 0): lambda X: -2 * a * X[2]})),
    ]
    pts = [(x, y, z) for x in (0.25, 0.75, 0.4) for y in (0.25, 0.75, 0.6) for z in (0.25, 0.75)]
    for vel, Gfn in fields:
        got, exact = _order2_at(vel, grad_field(Gfn), pts, dtype)
        assert np.abs(got - exact).max() < (2e-9 if dtype == np.float32 else 2e-10)
        # trilinear alone misses the curvature by up to a/4 x (1 - x) terms
        assert np.abs(exact).max() > 0


@pytest.mark.parametrize("dtype", DTYPES)
def test_order2_with_compressible_curvature_recovers_two_thirds(dtype):
    """u = a x^2 alone: the strain rate the corners can carry is deviatoric
    (StrainRate drops the trace), so S_xx = 2/3 du/dx and the correction
    supplies 2/3 of the curvature: the residual is exactly a x (1 - x) / 3."""
    a = 0.02

    def vel(X):
        return np.array([a * X[0] ** 2, 0, 0])

    def sgrad(X):
        g = np.diag([2 * a * X[0], 0.0, 0.0])
        return g - np.trace(g) / 3 * np.eye(3)

    pts = [(x, 0.5, 0.5) for x in (0.25, 0.75, 0.4)]
    got, exact = _order2_at(vel, sgrad, pts, dtype)
    tol = 2e-9 if dtype == np.float32 else 1e-10
    for p, g, ex in zip(pts, got, exact):
        assert abs((g[0] - ex[0]) - a * p[0] * (1 - p[0]) / 3) < tol
        assert abs(g[1]) < tol and abs(g[2]) < tol


def test_temporal_midpoint_is_exact():
    """alpha = 1/2 blends the stored coarse moments (linear in time) before
    the nonlinear rescale: equal to CF of the averaged moments, and different
    from averaging the two CF results."""
    pc, r, fc0 = setup(fn=lambda x, y, z: (1.0, 0.02, 0.0, 0.0))
    fc1 = coarse_field((20, 16, 16), lambda x, y, z: (1.0, 0.04 + 1e-3 * x, 0.01, 0.0))
    M0, M1 = O.extract(pc, r, fc0), O.extract(pc, r, fc1)
    Mm = (0.5 * M0 + 0.5 * M1).astype(np.float32)
    nf = tuple(r.nf)
    ff = np.zeros((19, nf[2], nf[1], nf[0]), np.float32)
    a = O.interp_cf(r, M0, M1, 0.5, ff)
    b = O.interp_cf(r, Mm, Mm, 0.0, ff)
    assert np.abs(a - b).max() < 1e-7
    c = 0.5 * O.interp_cf(r, M0, M0, 0.0, ff) + 0.5 * O.interp_cf(r, M1, M1, 0.0, ff)
    assert np.abs(a - c).max() > 1e-6


def test_les_rescale_reduces_to_reference_sf():
    """With CS = 0, CF's rescale multiplies the non-equilibrium second moments
    by SF_cf and FC's by 1/SF_cf (MRTtest/DynInlet3.cu:2878-2879)."""
    L = O._bind_refine()
    om = np.float32(1.7)
    om_f = np.float32(2) / (np.float32(1) + np.float32(2) * (np.float32(2) / om - np.float32(1)))
    sf = 0.5 * om * (1 - om_f) / ((1 - om) * om_f)
    rng = np.random.default_rng(3)
    M = O.moment_matrix()
    mom = np.array([1.0, 0.03, -0.01, 0.02, 0, 0, 0, 0, 0], np.float32)
    u, v, w = mom[1:4]
    eq = np.array([2 * u * u - (v * v + w * w), v * v - w * w, u * v, v * w, u * w])
    neq = rng.uniform(-1e-3, 1e-3, 5).astype(np.float32)
    mom[4:] = eq + neq
    for fn, s in ((L.orc_invert_phys_les_cf_f32, sf), (L.orc_invert_phys_les_fc_f32, 1 / sf)):
        f = np.zeros(19, np.float32)
        fn(f, mom, om if fn is L.orc_invert_phys_les_cf_f32 else om_f, 0.0, O.EQ_MRT_MEQ)
        m = M @ f.astype(np.float64)
        assert np.allclose(m[[0, 3, 5, 7]], mom[:4], atol=1e-7)
        assert np.allclose(m[[9, 11, 13, 14, 15]] - eq, s * neq, rtol=2e-4, atol=2e-8)
        # as coded: eps = +475/63 j^2 and 3pixx = piww = 0 (mrt_meq)
        assert abs(m[2] - 475 / 63 * (u * u + v * v + w * w)) < 1e-6 and abs(m[10]) < 1e-7 and abs(m[12]) < 1e-7


@pytest.mark.parametrize("dtype", DTYPES)
def test_refined_run_conserves_uniform_flow(dtype):
    """A uniform stream at the collide's fixed point stays exactly uniform
    through the refined region when the interface rebuilds populations with
    the collide's own equilibrium (eq_kind = EQ_BGK_MEQ); with mrt_meq as
    InvertPhysicalMoments codes it, the interface injects an eps jump of
    ~3 rho that streaming turns into density errors next to the frame."""
    n = (24, 16, 16)
    pc = O.params(*n, bc=(O.PERIODIC,) * 3, collision=O.MRT_LES, omega=1.5, cs=0.01)
    r = O.refine((4, 3, 3), (20, 16, 16), 1.5, cs=0.01)
    pf = O.params(*r.nf, bc=(O.CLAMP,) * 3, collision=O.MRT_LES, omega=r.omega_f, cs=0.01)
    fc = coarse_field(n, lambda x, y, z: (1.0, 0.03, 0.01, -0.005), dtype)
    ff = coarse_field(tuple(r.nf), lambda x, y, z: (1.0, 0.03, 0.01, -0.005), dtype)
    ca, fa = O.refine_run(pc, pf, r, fc, ff, 5)
    assert ca.dtype == dtype
    tol = 2e-6 if dtype == np.float32 else 3e-12
    assert np.abs(ca - fc).max() < tol and np.abs(fa - ff).max() < tol
    r.eq_kind = O.EQ_MRT_MEQ
    ca, fa = O.refine_run(pc, pf, r, fc, ff, 5)
    pfine = O.params(*r.nf)
    rho = O.macro(pfine, fa)[0]
    assert np.abs(rho - 1.0).max() > 1e-3


def test_double_interface_agrees_with_float_on_a_refined_run():
    """The double instance against the float one on 4 coarse steps of a
    random near-equilibrium state with walls, forcing and LES: FP32 rounding
    apart (the two are one source compiled twice, so a macro or constant
    slip in either shows here)."""
    rng = np.random.default_rng(5)
    n, a, nf = (24, 16, 16), (6, 4, 4), (16, 12, 12)
    pc = O.params(*n, bc=(O.CLAMP, O.WALL, O.WALL), collision=O.MRT_LES, omega=1 / 0.53, cs=0.01, halfway=1,
                  dpdy=1e-5)
    r = O.refine(a, nf, float(np.float32(1 / 0.53)), cs=0.01)
    pf = O.params(*nf, bc=(O.CLAMP,) * 3, collision=O.MRT_LES, omega=r.omega_f, cs=0.01, halfway=1, dpdy=5e-6)
    fc = coarse_field(n, lambda x, y, z: (1.0, 0.04, 0.0, 0.0), np.float64)
    ff = coarse_field(nf, lambda x, y, z: (1.0, 0.04, 0.0, 0.0), np.float64)
    fc *= 1 + rng.uniform(-0.02, 0.02, fc.shape)
    ff *= 1 + rng.uniform(-0.02, 0.02, ff.shape)
    c64, f64 = O.refine_run(pc, pf, r, fc, ff, 4)
    c32, f32 = O.refine_run(pc, pf, r, fc.astype(np.float32), ff.astype(np.float32), 4)
    assert c64.dtype == np.float64 and c32.dtype == np.float32
    scale = np.abs(c64).max()
    d = max(np.abs(c64 - c32).max(), np.abs(f64 - f32).max()) / scale
    assert 1e-9 < d < 2e-6
