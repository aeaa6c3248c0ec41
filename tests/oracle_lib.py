"""ctypes bindings of the oracle (oracle/liboracle.so) and of the reference's
own CPU solver built into oracle/_ref/libcpu3d_ref.so.

Test infrastructure: these are the checkers, never the thing measured.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libcpu3d_ref.so")
REF_EXE = os.path.join(ROOT, "oracle", "_ref", "cpu3d_ref")

BGK, MRT_CPU3D, MRT_LES, MRT_DHUM = 0, 1, 2, 3
PERIODIC, CLAMP, WALL = 0, 1, 2
(FLUID, SOLID, SOLID_FORCE, INLET_WEST, OUTLET_EAST, LID_WEST_REG, SYM_XMIN, SYM_XMAX, OUTLET_NORTH, INLET_SOUTH,
 RECYCLE, VEL_FROM_BELOW, GHOST_W, GHOST_E, GHOST_S, GHOST_N, GHOST_DYN) = range(17)


class OrcParams(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("bc", C.c_int * 3), ("collision", C.c_int), ("halfway", C.c_int),
                ("omega", C.c_double), ("cs", C.c_double), ("dpdy", C.c_double),
                ("u_lid", C.c_double), ("s_dhum", C.c_double * 19), ("dyn_y1", C.c_int),
                ("y_wrap", C.c_int), ("ghost_y", C.c_int * 3), ("v_in", C.c_void_p)]


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        _lib = C.CDLL(ORACLE_SO)
        f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
        f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
        for suf, fp in (("f32", f32p), ("f64", f64p)):
            getattr(_lib, "orc_collide_" + suf).argtypes = [C.POINTER(OrcParams), fp]
            getattr(_lib, "orc_moments_" + suf).argtypes = [fp, fp]
            getattr(_lib, "orc_step_" + suf).argtypes = [C.POINTER(OrcParams), C.c_void_p, C.c_void_p, fp, fp, C.c_void_p]
            getattr(_lib, "orc_run_" + suf).argtypes = [C.POINTER(OrcParams), C.c_void_p, C.c_void_p, fp, fp, C.c_int, C.c_void_p]
            getattr(_lib, "orc_macro_" + suf).argtypes = [C.POINTER(OrcParams), fp, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
            getattr(_lib, "orc_stats_update_" + suf).argtypes = [C.POINTER(OrcParams), C.c_void_p, fp, f32p, C.c_int,
                                                                 C.c_float, C.c_float]
            getattr(_lib, "orc_west_extrap_" + suf).argtypes = [fp, C.c_float if suf == "f32" else C.c_double]
            getattr(_lib, "orc_east_extrap_" + suf).argtypes = [fp, C.c_float if suf == "f32" else C.c_double]
            sct = C.c_float if suf == "f32" else C.c_double
            getattr(_lib, "orc_vel_extrap_" + suf).argtypes = [fp, sct, sct, sct]
            getattr(_lib, "orc_lid_west_reg_" + suf).argtypes = [fp, C.c_float if suf == "f32" else C.c_double]
            getattr(_lib, "orc_xsym_bot_" + suf).argtypes = [fp, C.c_int, C.c_int, C.c_int, C.c_int]
            getattr(_lib, "orc_xsym_top_" + suf).argtypes = [fp, C.c_int, C.c_int, C.c_int, C.c_int]
            sc = C.c_float if suf == "f32" else C.c_double
            getattr(_lib, "orc_feq_bgk_" + suf).argtypes = [sc, sc, sc, sc, fp]
            getattr(_lib, "orc_feq_mrt_" + suf).argtypes = [sc, sc, sc, sc, fp]
            getattr(_lib, "orc_les_omega_" + suf).argtypes = [C.POINTER(OrcParams), fp, C.POINTER(sc), C.POINTER(sc)]
        _lib.orc_default_params.argtypes = [C.POINTER(OrcParams)]
        _lib.orc_moment_matrix.argtypes = [f64p]
        _lib.orc_hash_u32.argtypes = [C.c_uint64, C.c_uint64]
        _lib.orc_hash_u32.restype = C.c_uint32
        _lib.orc_inlet_profile.argtypes = [C.c_int, C.c_int, C.c_float, C.c_float, C.c_uint64, f32p]
    return _lib


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        _ref = C.CDLL(REF_SO)
        f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
        _ref.ref_bgk_collide.argtypes = [f32p, C.c_float]
        _ref.ref_mrt_collide.argtypes = [f32p, C.c_float]
        _ref.ref_boundaries.argtypes = [f32p, C.c_int, C.c_int, C.c_int]
        _ref.ref_image.argtypes = [C.c_int, C.c_int, C.c_int]
        _ref.ref_case.argtypes = [np.ctypeslib.ndpointer(np.int32), f32p]
    return _ref


def params(nx=1, ny=1, nz=1, bc=(PERIODIC,) * 3, collision=BGK, halfway=0, omega=1.0,
           cs=0.0, dpdy=0.0, u_lid=0.0, dyn_y1=-1, y_wrap=-1, ghost_y=(0, 0, 0), v_in=None):
    """Oracle parameters.  v_in (float32 [nz*nx], the south inlet table) is
    kept alive on the returned object."""
    p = OrcParams()
    lib().orc_default_params(C.byref(p))
    p.nx, p.ny, p.nz = nx, ny, nz
    for a in range(3):
        p.bc[a] = bc[a]
    p.collision, p.halfway = collision, halfway
    p.omega, p.cs, p.dpdy, p.u_lid = omega, cs, dpdy, u_lid
    p.dyn_y1 = dyn_y1
    p.y_wrap = y_wrap
    for k in range(3):
        p.ghost_y[k] = ghost_y[k]
    if v_in is not None:
        p._v_in = np.ascontiguousarray(v_in, np.float32)
        p.v_in = p._v_in.ctypes.data
    return p


def plane_v_sum(f, y):
    """AverageV (MRTtest/DynInlet3.cu:2561-2600) restated: v = sum_i c_iy f_i
    at each node of the plane y in the reference's FP32 order
    (f2-f4+f5+f6-f7-f8+f11-f13+f16-f18), summed over all x and z in FP64.
    f: populations [19][nz][ny][nx]."""
    g = np.asarray(f, np.float32)[:, :, y, :]
    v = g[2] - g[4] + g[5] + g[6] - g[7] - g[8] + g[11] - g[13] + g[16] - g[18]
    return float(v.astype(np.float64).sum())


def inlet_control_update(dpdy, sum_v, nx, nz, umax, kp, dpdy0):
    """The host P-controller of main() (MRTtest/DynInlet3.cu:3159-3165), in
    FP32: Av_V = sum / ((nx - 2) nz), dpdy += (Av_V - UMAX)/UMAX * KP * |DPDY|.
    Returns (Av_V, new dpdy)."""
    f = np.float32
    av = f(sum_v) / f((nx - 2) * nz)
    if kp != 0:
        diff = (av - f(umax)) / f(umax)
        dpdy = f(f(dpdy) + diff * f(kp) * abs(f(dpdy0)))
    return float(av), float(dpdy)


def _ptr(a):
    return None if a is None else a.ctypes.data


def collide(p, f):
    f = np.ascontiguousarray(f).copy()
    suf = "f32" if f.dtype == np.float32 else "f64"
    getattr(lib(), "orc_collide_" + suf)(C.byref(p), f)
    return f


def moments(f):
    f = np.ascontiguousarray(f)
    m = np.zeros(19, f.dtype)
    suf = "f32" if f.dtype == np.float32 else "f64"
    getattr(lib(), "orc_moments_" + suf)(f, m)
    return m


def step(p, f, mask=None, u_in=None):
    f = np.ascontiguousarray(f)
    out = np.empty_like(f)
    force = np.zeros(3, np.float64)
    suf = "f32" if f.dtype == np.float32 else "f64"
    getattr(lib(), "orc_step_" + suf)(C.byref(p), _ptr(mask), _ptr(u_in), f, out, force.ctypes.data)
    return out, force


def run(p, f, nsteps, mask=None, u_in=None, forces=False):
    """Advance f (shape (19, nz, ny, nx)) nsteps; returns (f, forces[nsteps,3])."""
    fa = np.ascontiguousarray(f).copy()
    fb = np.empty_like(fa)
    fr = np.zeros((max(nsteps, 1), 3), np.float64) if forces else None
    suf = "f32" if fa.dtype == np.float32 else "f64"
    getattr(lib(), "orc_run_" + suf)(C.byref(p), _ptr(mask), _ptr(u_in), fa, fb, nsteps, _ptr(fr))
    return fa, fr


def macro(p, f):
    n = p.nx * p.ny * p.nz
    out = [np.empty(n, f.dtype) for _ in range(4)]
    suf = "f32" if f.dtype == np.float32 else "f64"
    getattr(lib(), "orc_macro_" + suf)(C.byref(p), np.ascontiguousarray(f), *[o.ctypes.data for o in out])
    return [o.reshape(p.nz, p.ny, p.nx) for o in out]


def feq_bgk(rho, u, v, w, dtype=np.float32):
    f = np.zeros(19, dtype)
    suf = "f32" if dtype == np.float32 else "f64"
    getattr(lib(), "orc_feq_bgk_" + suf)(rho, u, v, w, f)
    return f


def feq_mrt(rho, u, v, w, dtype=np.float32):
    f = np.zeros(19, dtype)
    suf = "f32" if dtype == np.float32 else "f64"
    getattr(lib(), "orc_feq_mrt_" + suf)(rho, u, v, w, f)
    return f


def moment_matrix():
    M = np.zeros((19, 19), np.float64)
    lib().orc_moment_matrix(M)
    return M


def inlet_profile(ny, nz, u0, amp, seed):
    out = np.empty(ny * nz, np.float32)
    lib().orc_inlet_profile(ny, nz, u0, amp, seed, out)
    return out


# ---- refinement ----
class OrcRefine(C.Structure):
    _fields_ = [("a", C.c_int * 3), ("nf", C.c_int * 3), ("omega_c", C.c_double), ("omega_f", C.c_double),
                ("cs", C.c_double), ("temporal", C.c_int), ("eq_kind", C.c_int)]


EQ_BGK_MEQ, EQ_BGK_PROJ, EQ_MRT_MEQ = 0, 1, 2


def _bind_refine():
    L = lib()
    if getattr(L, "_refine_bound", False):
        return L
    for suf, dt, cr in (("f32", np.float32, C.c_float), ("f64", np.float64, C.c_double)):
        fp = np.ctypeslib.ndpointer(dt, flags="C_CONTIGUOUS")
        getattr(L, "orc_phys_moments_" + suf).argtypes = [fp, fp]
        getattr(L, "orc_strain_rate_" + suf).argtypes = [fp, fp]
        getattr(L, "orc_invert_phys_les_cf_" + suf).argtypes = [fp, fp, cr, cr, C.c_int]
        getattr(L, "orc_invert_phys_les_fc_" + suf).argtypes = [fp, fp, cr, cr, C.c_int]
        getattr(L, "orc_order2_correction_" + suf).argtypes = [fp, fp, cr, cr, cr, C.c_int]
        getattr(L, "orc_extract_" + suf).argtypes = [C.POINTER(OrcParams), C.POINTER(OrcRefine), fp, fp]
        getattr(L, "orc_interp_cf_" + suf).argtypes = [C.POINTER(OrcRefine), fp, fp, cr, C.c_void_p, fp]
        getattr(L, "orc_interp_fc_" + suf).argtypes = [C.POINTER(OrcParams), C.POINTER(OrcRefine), fp, C.c_void_p,
                                                       fp]
        getattr(L, "orc_refine_step_" + suf).argtypes = [C.POINTER(OrcParams), C.POINTER(OrcParams),
                                                         C.POINTER(OrcRefine), C.c_void_p, C.c_void_p, C.c_void_p,
                                                         fp, fp, fp, fp, fp, fp, C.c_void_p]
    L._refine_bound = True
    return L


def _suf(dtype):
    return "f64" if np.dtype(dtype) == np.float64 else "f32"


def refine(a, nf, omega_c, omega_f=None, cs=0.0, temporal=1, eq_kind=0):
    r = OrcRefine()
    for k in range(3):
        r.a[k], r.nf[k] = a[k], nf[k]
    r.omega_c = omega_c
    r.omega_f = omega_f if omega_f is not None else float(np.float32(2) / (np.float32(1) + np.float32(2) * (np.float32(2) / np.float32(omega_c) - np.float32(1))))
    r.cs, r.temporal, r.eq_kind = cs, temporal, eq_kind
    return r


def footprint_size(r):
    return (r.nf[2] // 2 + 1) * (r.nf[1] // 2 + 1) * (r.nf[0] // 2 + 1)


def extract(pc, r, fc, dtype=None):
    """Footprint moments of a coarse lattice, in the lattice's precision
    (float32 unless fc or dtype is float64)."""
    dt = np.dtype(dtype or (np.float64 if np.asarray(fc).dtype == np.float64 else np.float32))
    M = np.zeros(9 * footprint_size(r), dt)
    getattr(_bind_refine(), "orc_extract_" + _suf(dt))(C.byref(pc), C.byref(r), np.ascontiguousarray(fc, dt), M)
    return M


def interp_cf(r, Mold, Mnew, alpha, ff, fmask=None):
    dt = np.asarray(Mold).dtype
    ff = np.ascontiguousarray(ff, dt).copy()
    getattr(_bind_refine(), "orc_interp_cf_" + _suf(dt))(C.byref(r), Mold, Mnew, alpha, _ptr(fmask), ff)
    return ff


def interp_fc(pc, r, ff, fc, cmask=None):
    dt = np.float64 if np.asarray(fc).dtype == np.float64 else np.float32
    fc = np.ascontiguousarray(fc, dt).copy()
    getattr(_bind_refine(), "orc_interp_fc_" + _suf(dt))(C.byref(pc), C.byref(r), np.ascontiguousarray(ff, dt),
                                                         _ptr(cmask), fc)
    return fc


def refine_run(pc, pf, r, ca, fa, nsteps, cmask=None, fmask=None, u_in=None, forces=False):
    """Advance coarse (ca) and fine (fa) lattices nsteps coarse steps, in
    float32, or in float64 when ca is float64 (the FP64 interface); with
    forces, also the fine level's first-sub-step momentum exchange of each
    coarse step, [nsteps, 3]."""
    L = _bind_refine()
    dt = np.float64 if np.asarray(ca).dtype == np.float64 else np.float32
    ca = np.ascontiguousarray(ca, dt).copy()
    fa = np.ascontiguousarray(fa, dt).copy()
    if u_in is not None:
        u_in = np.ascontiguousarray(u_in, dt)
    cb, fb = np.empty_like(ca), np.empty_like(fa)
    Mold = extract(pc, r, ca)
    Mnew = np.empty_like(Mold)
    F = np.zeros((max(nsteps, 1), 3), np.float64)
    step = getattr(L, "orc_refine_step_" + _suf(dt))
    for t in range(nsteps):
        step(C.byref(pc), C.byref(pf), C.byref(r), _ptr(cmask), _ptr(fmask), _ptr(u_in), ca, cb, fa, fb, Mold, Mnew,
             F[t].ctypes.data)
    return (ca, fa, F) if forces else (ca, fa)


def stats_update(p, f, stat, mode, n, w=1.0, mask=None):
    """One running-statistics update of stat (float32 [6, nz, ny, nx], in
    place) from a step's output populations f."""
    assert stat.dtype == np.float32 and stat.flags.c_contiguous
    f = np.ascontiguousarray(f)
    suf = "f32" if f.dtype == np.float32 else "f64"
    getattr(lib(), "orc_stats_update_" + suf)(C.byref(p), _ptr(mask), f, stat.reshape(-1), mode, n, w)


def xsym(f, side, y, z, ny, nz):
    """Symmetry-plane condition on one node's populations (side 'bot': x = 0,
    'top': x = nx-1)."""
    f = np.ascontiguousarray(f).copy()
    suf = "f32" if f.dtype == np.float32 else "f64"
    getattr(lib(), f"orc_xsym_{side}_" + suf)(f, y, z, ny, nz)
    return f
