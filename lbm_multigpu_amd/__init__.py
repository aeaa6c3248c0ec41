"""lbm_multigpu_amd -- Python binding of liblbm, the MI355X (gfx950) D3Q19
MRT-LES stream-collide and local-refinement library.

The library is the C ABI in include/lbm.h, built in-tree as
lbm_multigpu_amd/liblbm.so.  This module is a thin ctypes layer over it:
the hot path runs only in the HIP kernels; there is no CPU fallback, and
importing the solver without the built library raises.
"""
import ctypes as C
import os

import numpy as np

from .config import Config, read_inputs, write_inputs  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblbm.so")

# constants of include/lbm.h
MODEL_BGK, MODEL_MRT_CPU3D, MODEL_MRT_LES, MODEL_MRT_DHUMIERES = 0, 1, 2, 3
BC_PERIODIC, BC_CLAMP, BC_WALL = 0, 1, 2
STORE_F32, STORE_F16, STORE_BF16, STORE_F64 = 0, 1, 2, 3
GEOM_NONE, GEOM_CHANNEL, GEOM_CPU3D, GEOM_PRECURSOR = 0, 1, 2, 3
(NODE_FLUID, NODE_SOLID, NODE_SOLID_FORCE, NODE_INLET_WEST, NODE_OUTLET_EAST, NODE_LID_WEST_REG,
 NODE_SYM_XMIN, NODE_SYM_XMAX, NODE_OUTLET_NORTH, NODE_INLET_SOUTH, NODE_RECYCLE, NODE_VEL_FROM_BELOW,
 NODE_GHOST_W, NODE_GHOST_E, NODE_GHOST_S, NODE_GHOST_N, NODE_GHOST_DYN) = range(17)
INIT_BGK, INIT_MRT = 0, 1

# reference `im` code of each node code (MRTtest/DynInlet3.cu:1286-1390, :1711-1771,
# bgk_channel1/Re3000_Cyl2/DynInlet3.cu:1220-1239, CPU_3D/CPU_3D.cpp:1142).  The
# reference uses 53 twice: CPU_3D's regularised lid and MRTtest's ghost copy of
# row DYNY2; IM_TO_NODE maps 53 to the lid (CPU_3D's meaning).
NODE_TO_IM = {NODE_FLUID: 0, NODE_SOLID: 1, NODE_SOLID_FORCE: 10, NODE_INLET_WEST: 400,
              NODE_OUTLET_EAST: 300, NODE_LID_WEST_REG: 53, NODE_SYM_XMIN: 26, NODE_SYM_XMAX: 25,
              NODE_OUTLET_NORTH: 100, NODE_INLET_SOUTH: 200, NODE_RECYCLE: 80, NODE_VEL_FROM_BELOW: 70,
              NODE_GHOST_W: 50, NODE_GHOST_E: 51, NODE_GHOST_S: 52, NODE_GHOST_N: 53, NODE_GHOST_DYN: 54}
IM_TO_NODE = {v: k for k, v in NODE_TO_IM.items() if k != NODE_GHOST_N}


class LBMError(RuntimeError):
    """A liblbm call returned a negative LBM_E* code."""

    def __init__(self, code, msg):
        super().__init__(f"liblbm error {code}: {msg}")
        self.code = code


class lbm_config(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("bc", C.c_int * 3),
                ("model", C.c_int), ("halfway", C.c_int), ("storage", C.c_int),
                ("omega", C.c_float), ("cs", C.c_float), ("dpdy", C.c_float), ("umax", C.c_float),
                ("inlet_amp", C.c_float), ("seed", C.c_uint64), ("s_dhum", C.c_float * 19),
                ("geometry", C.c_int), ("obst_x", C.c_float), ("obst_y", C.c_float),
                ("obst_z", C.c_float), ("obst_r", C.c_float), ("startf", C.c_int),
                ("force_steps", C.c_int), ("refinement", C.c_int), ("lr_nx", C.c_int),
                ("lr_ny", C.c_int), ("lr_nz", C.c_int), ("lr_x0", C.c_float), ("lr_y0", C.c_float),
                ("lr_z0", C.c_float), ("temporal_interp", C.c_int), ("interface_eq", C.c_int),
                ("dyn_y1", C.c_int), ("kp", C.c_float), ("ctrl_every", C.c_int), ("ctrl_start", C.c_int),
                ("dyn_mode", C.c_int), ("dyn_y2", C.c_int), ("obst2_x", C.c_float), ("obst2_y", C.c_float),
                ("obst2_z", C.c_float), ("obst2_r", C.c_float), ("trip_until", C.c_int), ("trip_z", C.c_int)]


class lbm_sizes(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("nx_pad", C.c_int),
                ("cells", C.c_size_t), ("device_bytes", C.c_size_t), ("force_steps", C.c_int)]


EXPORTS = [
    "lbm_config_default", "lbm_create", "lbm_set_geometry", "lbm_get_mask",
    "lbm_init_equilibrium", "lbm_init_taylor_green", "lbm_step", "lbm_step_timed",
    "lbm_get_moments", "lbm_get_populations", "lbm_get_populations_f64",
    "lbm_set_populations", "lbm_set_populations_f64", "lbm_get_forces",
    "lbm_get_mass_momentum", "lbm_write_tecplot", "lbm_query_sizes", "lbm_get_time",
    "lbm_last_error", "lbm_destroy",
    "lbm_mlp_create", "lbm_mlp_destroy", "lbm_mlp_last_error", "lbm_mlp_num_params",
    "lbm_mlp_set_params", "lbm_mlp_get_params", "lbm_mlp_set_norm", "lbm_mlp_set_data",
    "lbm_mlp_gradient", "lbm_mlp_train", "lbm_mlp_forward", "lbm_set_closure",
    "lbm_closure_samples", "lbm_set_closure_target", "lbm_stats_enable", "lbm_stats_get", "lbm_set_closure_interval",
    "lbm_get_inlet_control", "lbm_plane_mean_v", "lbm_get_dpdy", "lbm_mass_momentum_noalloc",
    "lbm_get_closure_rates", "lbm_set_closure_base",
]

# closure network: 7 -> 64 -> 64 -> 64 -> 1, flat FP32 parameter vector
MLP_IN, MLP_H = 7, 64
MLP_NPARAM = MLP_IN * MLP_H + MLP_H + 2 * (MLP_H * MLP_H + MLP_H) + MLP_H + 1

_lib = None


def load_library(path=LIB_PATH):
    """Load liblbm.so (raises OSError if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise OSError(f"liblbm.so not built at {path}: run __graft_entry__.build()")
    lib = C.CDLL(path)
    P = C.c_void_p
    f32 = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
    f64 = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
    u8 = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
    lib.lbm_config_default.argtypes = [C.POINTER(lbm_config)]
    lib.lbm_config_default.restype = None
    lib.lbm_create.argtypes = [C.POINTER(lbm_config), C.POINTER(P)]
    lib.lbm_set_geometry.argtypes = [P, P]
    lib.lbm_get_mask.argtypes = [P, C.c_int, u8]
    lib.lbm_init_equilibrium.argtypes = [P, C.c_int, P, P, P, P]
    lib.lbm_init_taylor_green.argtypes = [P, C.c_float]
    lib.lbm_step.argtypes = [P, C.c_int]
    lib.lbm_step_timed.argtypes = [P, C.c_int, C.POINTER(C.c_float)]
    lib.lbm_get_moments.argtypes = [P, C.c_int, P, P, P, P]
    lib.lbm_get_populations.argtypes = [P, C.c_int, f32]
    lib.lbm_get_populations_f64.argtypes = [P, C.c_int, f64]
    lib.lbm_set_populations.argtypes = [P, C.c_int, f32]
    lib.lbm_set_populations_f64.argtypes = [P, C.c_int, f64]
    lib.lbm_get_forces.argtypes = [P, P, C.c_int, C.POINTER(C.c_int)]
    lib.lbm_get_mass_momentum.argtypes = [P, C.c_int, f64]
    lib.lbm_mass_momentum_noalloc.argtypes = [P, C.c_int, P, P]
    lib.lbm_write_tecplot.argtypes = [P, C.c_char_p]
    lib.lbm_query_sizes.argtypes = [P, C.c_int, C.POINTER(lbm_sizes)]
    lib.lbm_get_time.argtypes = [P, C.POINTER(C.c_longlong)]
    lib.lbm_last_error.argtypes = [P]
    lib.lbm_last_error.restype = C.c_char_p
    lib.lbm_destroy.argtypes = [P]
    lib.lbm_destroy.restype = None
    lib.lbm_hash_u32.argtypes = [C.c_uint64, C.c_uint64]
    lib.lbm_hash_u32.restype = C.c_uint32
    lib.lbm_inlet_profile.argtypes = [C.c_int, C.c_int, C.c_float, C.c_float, C.c_uint64, f32]
    lib.lbm_inlet_profile.restype = None
    lib.lbm_stats_enable.argtypes = [P, C.c_longlong, C.c_longlong]
    lib.lbm_get_inlet_control.argtypes = [P, P, C.c_int, C.POINTER(C.c_int)]
    lib.lbm_plane_mean_v.argtypes = [P, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    lib.lbm_get_dpdy.argtypes = [P, C.POINTER(C.c_float)]
    lib.lbm_stats_get.argtypes = [P, C.c_int, P, P]
    lib.lbm_mlp_create.argtypes = [C.c_uint64, C.POINTER(P)]
    lib.lbm_mlp_destroy.argtypes = [P]
    lib.lbm_mlp_destroy.restype = None
    lib.lbm_mlp_last_error.argtypes = [P]
    lib.lbm_mlp_last_error.restype = C.c_char_p
    lib.lbm_mlp_num_params.argtypes = []
    lib.lbm_mlp_set_params.argtypes = [P, f32]
    lib.lbm_mlp_get_params.argtypes = [P, f32]
    lib.lbm_mlp_set_norm.argtypes = [P, f32, f32, f32, f32, C.c_float]
    lib.lbm_mlp_set_data.argtypes = [P, f32, f32, C.c_longlong]
    lib.lbm_mlp_gradient.argtypes = [P, C.c_longlong, C.c_int, P, C.POINTER(C.c_double)]
    lib.lbm_mlp_train.argtypes = [P, C.c_int, C.c_int, C.c_float, P]
    lib.lbm_mlp_forward.argtypes = [P, f32, C.c_longlong, f32]
    lib.lbm_set_closure.argtypes = [P, P, C.c_float]
    lib.lbm_set_closure_interval.argtypes = [P, C.c_int]
    lib.lbm_closure_samples.argtypes = [P, C.c_uint64, C.c_float, f32, f32, C.c_int, C.POINTER(C.c_int)]
    lib.lbm_set_closure_target.argtypes = [P, C.c_int]
    lib.lbm_get_closure_rates.argtypes = [P, P]
    lib.lbm_set_closure_base.argtypes = [P, C.c_int]
    _lib = lib
    return lib


def default_config():
    cfg = lbm_config()
    load_library().lbm_config_default(C.byref(cfg))
    return cfg


def inlet_profile(ny, nz, u0, amp, seed):
    """The library's seeded inlet table u_in[z*ny + y] (host code, no GPU)."""
    out = np.empty(ny * nz, np.float32)
    load_library().lbm_inlet_profile(ny, nz, u0, amp, seed, out)
    return out


class Solver:
    """One liblbm context.  Methods mirror the C ABI one to one."""

    def __init__(self, cfg):
        stats = None
        if isinstance(cfg, Config):
            if cfg.velav:
                stats = (cfg.start_velav, max(cfg.start_velfluc, cfg.start_velav))
            cfg = cfg.to_c()
        self.lib = load_library()
        self.cfg = cfg
        h = C.c_void_p()
        rc = self.lib.lbm_create(C.byref(cfg), C.byref(h))
        if rc != 0:
            raise LBMError(rc, self.lib.lbm_last_error(None).decode())
        self.h = h
        if stats:
            self.stats_enable(*stats)

    def _check(self, rc):
        if rc != 0:
            raise LBMError(rc, self.lib.lbm_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.lbm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def sizes(self, level=0):
        s = lbm_sizes()
        self._check(self.lib.lbm_query_sizes(self.h, level, C.byref(s)))
        return s

    def shape(self, level=0):
        s = self.sizes(level)
        return (s.nz, s.ny, s.nx)

    def set_geometry(self, mask=None):
        if mask is None:
            self._check(self.lib.lbm_set_geometry(self.h, None))
        else:
            m = np.ascontiguousarray(mask, dtype=np.uint8)
            self._check(self.lib.lbm_set_geometry(self.h, m.ctypes.data))

    def get_mask(self, level=0):
        out = np.empty(self.shape(level), np.uint8)
        self._check(self.lib.lbm_get_mask(self.h, level, out.reshape(-1)))
        return out

    def init_equilibrium(self, kind=INIT_BGK, rho=None, u=None, v=None, w=None):
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float32) for a in (rho, u, v, w)]
        ptrs = [None if a is None else a.ctypes.data for a in arrs]
        self._check(self.lib.lbm_init_equilibrium(self.h, kind, *ptrs))

    def init_taylor_green(self, u0):
        self._check(self.lib.lbm_init_taylor_green(self.h, u0))

    def step(self, n=1):
        self._check(self.lib.lbm_step(self.h, n))

    def step_timed(self, n):
        ms = C.c_float()
        self._check(self.lib.lbm_step_timed(self.h, n, C.byref(ms)))
        return ms.value

    def moments(self, level=0):
        shp = self.shape(level)
        out = [np.empty(shp, np.float64) for _ in range(4)]
        self._check(self.lib.lbm_get_moments(self.h, level, *[o.ctypes.data for o in out]))
        return out

    def populations(self, level=0, f64=False):
        shp = (19,) + self.shape(level)
        out = np.empty(shp, np.float64 if f64 else np.float32)
        fn = self.lib.lbm_get_populations_f64 if f64 else self.lib.lbm_get_populations
        self._check(fn(self.h, level, out.reshape(-1)))
        return out

    def set_populations(self, f, level=0):
        f = np.ascontiguousarray(f)
        if f.dtype == np.float64:
            self._check(self.lib.lbm_set_populations_f64(self.h, level, f.reshape(-1)))
        else:
            self._check(self.lib.lbm_set_populations(self.h, level, f.astype(np.float32).reshape(-1)))

    def forces(self):
        n = C.c_int()
        self._check(self.lib.lbm_get_forces(self.h, None, 0, C.byref(n)))
        out = np.zeros((max(n.value, 0), 3), np.float64)
        if n.value:
            self._check(self.lib.lbm_get_forces(self.h, out.ctypes.data, n.value, C.byref(n)))
        return out

    def inlet_control(self):
        """Dynamic-inlet control history: rows (t, Av_V, dpdy after the update)."""
        n = C.c_int()
        self._check(self.lib.lbm_get_inlet_control(self.h, None, 0, C.byref(n)))
        out = np.zeros((max(n.value, 0), 3), np.float64)
        if n.value:
            self._check(self.lib.lbm_get_inlet_control(self.h, out.ctypes.data, n.value, C.byref(n)))
        return out

    def plane_mean_v(self, y):
        """AverageV on the coarse lattice: (FP64 sum of v over plane y, the
        reference's mean sum / ((nx - 2) nz))."""
        sm, mean = C.c_double(), C.c_double()
        self._check(self.lib.lbm_plane_mean_v(self.h, int(y), C.byref(sm), C.byref(mean)))
        return sm.value, mean.value

    def dpdy(self):
        v = C.c_float()
        self._check(self.lib.lbm_get_dpdy(self.h, C.byref(v)))
        return v.value

    def mass_momentum(self, level=0):
        out = np.zeros(4, np.float64)
        self._check(self.lib.lbm_get_mass_momentum(self.h, level, out))
        return out

    def write_tecplot(self, casename):
        self._check(self.lib.lbm_write_tecplot(self.h, casename.encode()))

    def time(self):
        t = C.c_longlong()
        self._check(self.lib.lbm_get_time(self.h, C.byref(t)))
        return t.value

    def stats_enable(self, start_av, start_fluc):
        """Running velocity mean from step start_av, mean squared deviation
        from step start_fluc (the reference's VELAV statistics)."""
        self._check(self.lib.lbm_stats_enable(self.h, start_av, start_fluc))

    def stats(self, level=0):
        """(mean[3, nz, ny, nx], dev[3, nz, ny, nx]) of u, v, w."""
        shp = (3,) + self.shape(level)
        mean = np.empty(shp, np.float32)
        dev = np.empty(shp, np.float32)
        self._check(self.lib.lbm_stats_get(self.h, level, mean.ctypes.data, dev.ctypes.data))
        return mean, dev

    def set_closure(self, mlp, nu_cap=0.05, every=1, smagorinsky_base=False):
        """Attach a trained Mlp (None detaches): the collide then takes its
        eddy viscosity from the network, Smagorinsky outside its range; the
        network is evaluated every `every` steps.  With smagorinsky_base the
        network's nu_t adds to the Smagorinsky relaxation time of the
        configured CS instead of to tau0."""
        self._closure = mlp  # keep the network alive while attached
        self._check(self.lib.lbm_set_closure(self.h, None if mlp is None else mlp.h, nu_cap))
        if mlp is not None:
            self._check(self.lib.lbm_set_closure_interval(self.h, every))
            self._check(self.lib.lbm_set_closure_base(self.h, int(bool(smagorinsky_base))))

    def closure_rates(self):
        """Per-node collision rate of the last closure step [nz, ny, nx]:
        > 0 the network's, < 0 minus the Smagorinsky fallback's."""
        out = np.empty(self.shape(0), np.float32)
        self._check(self.lib.lbm_get_closure_rates(self.h, out.ctypes.data))
        return out

    def set_closure_target(self, with_model_stress):
        """closure_samples' target: the subfilter stress alone (False) or with
        this run's own Smagorinsky stress, filtered (True)."""
        self._check(self.lib.lbm_set_closure_target(self.h, int(bool(with_model_stress))))

    def closure_samples(self, seed, keep, max_samples):
        """Features X[n, 7] and eddy-viscosity targets y[n] from this run,
        box-filtered by 2, at a hashed fraction `keep` of the filtered cells."""
        X = np.empty((max_samples, MLP_IN), np.float32)
        y = np.empty(max_samples, np.float32)
        n = C.c_int()
        rc = self.lib.lbm_closure_samples(self.h, seed, keep, X.reshape(-1), y, max_samples, C.byref(n))
        if rc == -22 and n.value > max_samples:  # more qualify: fetch them all with a buffer of that size
            return self.closure_samples(seed, keep, n.value)
        self._check(rc)
        return X[: n.value].copy(), y[: n.value].copy()


class Mlp:
    """The closure network on the GPU (include/lbm.h lbm_mlp_*)."""

    def __init__(self, seed=0):
        self.lib = load_library()
        h = C.c_void_p()
        rc = self.lib.lbm_mlp_create(seed, C.byref(h))
        if rc != 0:
            raise LBMError(rc, self.lib.lbm_mlp_last_error(None).decode())
        self.h = h

    def _check(self, rc):
        if rc != 0:
            raise LBMError(rc, self.lib.lbm_mlp_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.lbm_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self):
        out = np.empty(MLP_NPARAM, np.float32)
        self._check(self.lib.lbm_mlp_get_params(self.h, out))
        return out

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32).reshape(-1)
        if p.size != MLP_NPARAM:
            raise ValueError(f"expected {MLP_NPARAM} parameters, got {p.size}")
        self._check(self.lib.lbm_mlp_set_params(self.h, p))

    def set_norm(self, mu, sd, lo, hi, nu_scale):
        a = [np.ascontiguousarray(v, dtype=np.float32).reshape(MLP_IN) for v in (mu, sd, lo, hi)]
        self._check(self.lib.lbm_mlp_set_norm(self.h, *a, nu_scale))

    def set_data(self, X, y):
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32).reshape(-1)
        if X.shape != (y.size, MLP_IN):
            raise ValueError("X must be [n, 7] and y [n]")
        self._check(self.lib.lbm_mlp_set_data(self.h, X.reshape(-1), y, y.size))

    def gradient(self, offset, count):
        g = np.empty(MLP_NPARAM, np.float32)
        loss = C.c_double()
        self._check(self.lib.lbm_mlp_gradient(self.h, offset, count, g.ctypes.data, C.byref(loss)))
        return g, loss.value

    def train(self, steps, batch, lr):
        hist = np.empty(steps, np.float64)
        self._check(self.lib.lbm_mlp_train(self.h, steps, batch, lr, hist.ctypes.data if steps else None))
        return hist

    def forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        out = np.empty(X.shape[0], np.float32)
        self._check(self.lib.lbm_mlp_forward(self.h, X.reshape(-1), X.shape[0], out))
        return out
