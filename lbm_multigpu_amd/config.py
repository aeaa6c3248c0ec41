"""Run configuration and the reference's `.inputs` format.

The reference configures each program with compile-time #defines
(MRTtest/DynInlet3.cu:10-60) and echoes them at start-up to CASENAME.inputs
with WriteInputs (MRTtest/DynInlet3.cu:2805-2833); no reference program reads
that file back.  This module reads and writes the same keys, with the
separators the reference produces ("Key: \\tv", "Key \\t: v", "Key \\tv"), so
a run record of the reference sets up the same case here.  Keys that the
reference does not have (boundary modes, storage, seed ...) follow the
reference keys under an "# lbm-mi355x" line.
"""
import dataclasses
import math
import re

# keys of WriteInputs in the reference's order (MRTtest/DynInlet3.cu:2807-2831;
# Test/im80.inputs adds "KP:")
REFERENCE_KEYS = [
    "Base domain size", "Base blocksize:", "Obst1 location:", "Obst1 radius:", "Obst2 location:",
    "Obst2 radius:", "RE:", "UMAX:", "omega", "DPDY", "TMAX:", "STARTF:", "START_VELAV:",
    "START_VELFLUC:", "REFINEMENT:", "MODEL:", "Smagorinsky LES:", "CS:", "LR domain size",
    "LR factor", "LR location", "LR blocksize:", "omega in LR", "GPUs per node", "KP:",
]
EXTENSION_KEYS = [
    "MRT rates:", "Boundary x:", "Boundary y:", "Boundary z:", "Halfway bounce-back:", "Storage:",
    "Geometry:", "Inlet perturbation:", "Seed:", "Temporal interpolation:", "Force steps:",
    "Statistics:", "DYNY1:", "Inlet control:", "DYNY2:", "Dynamic inlet mode:", "Trip until:", "Trip z:",
]

_BC = {"periodic": 0, "clamp": 1, "wall": 2}
_STORE = {"f32": 0, "f16": 1, "bf16": 2, "f64": 3}
_GEOM = {"none": 0, "channel": 1, "cpu3d": 2, "precursor": 3}
_RATES = {"lrmgpu": 2, "dhumieres": 3, "cpu3d": 1}


@dataclasses.dataclass
class Config:
    nx: int = 64
    ny: int = 64
    nz: int = 64
    bc: tuple = ("periodic", "periodic", "periodic")
    model: str = "BGK"          # "BGK" or "MRT" (MODEL)
    mrt_rates: str = "lrmgpu"   # lrmgpu (DynInlet3), dhumieres (:774-783), cpu3d (CPU_3D)
    les: bool = False           # "Smagorinsky LES"
    cs: float = 0.0             # "CS" (the reference's CS = Cs^2)
    omega: float = 1.0
    re: float = 0.0             # "RE" (omega derives from it when omega is absent)
    umax: float = 0.0           # "UMAX"
    dpdy: float = 0.0           # "DPDY"
    tmax: int = 0               # "TMAX"
    startf: int = 0             # "STARTF"
    start_velav: int = 0
    start_velfluc: int = 0
    velav: bool = False         # VELAV: running velocity statistics (MRTtest/DynInlet3.cu:58)
    halfway: bool = False
    storage: str = "f32"
    geometry: str = "none"
    obst1: tuple = (0.0, 0.0, 0.0)
    obst1_r: float = 0.0
    obst2: tuple = (0.0, 0.0, 0.0)
    obst2_r: float = 0.0
    inlet_amp: float = 0.0
    seed: int = 0
    force_steps: int = 0
    refinement: bool = False
    lr_size: tuple = (0, 0, 0)
    lr_factor: float = 0.5
    lr_loc: tuple = (0.0, 0.0, 0.0)
    omega_lr: float = 0.0
    temporal_interp: bool = True
    interface_eq: int = 0       # 0: the collide's own equilibrium; 1: mrt_meq as coded
    gpus: int = 1
    # dynamic-inlet precursor (MRTtest/DynInlet3.cu:1744, :2561-2600, :3147-3170)
    dyn_y1: int = -1            # DYNY1: rows past it get no body force; -1: none cut
    kp: float = 0.0             # "KP": P-control gain on dpdy
    ctrl_every: int = 100       # control interval (steps); 0: off
    ctrl_start: int = 10000     # control from steps t > ctrl_start
    # bgk_channel1/Re3000_Cyl2 precursor: 1 = rows 0..DYNY1 periodic in y, force on every row
    dyn_mode: int = 0
    dyn_y2: int = 1             # DYNY2: source row of MRTtest's im 53 ghost copy
    trip_until: int = 0         # the trip obstacle (Obst2) exists for steps t < trip_until
    trip_z: int = -1            # z plane of the trip's offset block; -1: none

    def model_code(self):
        if self.model.upper() == "BGK":
            return 0
        return _RATES[self.mrt_rates]

    def to_c(self):
        """The lbm_config struct of include/lbm.h."""
        from . import default_config

        c = default_config()
        c.nx, c.ny, c.nz = self.nx, self.ny, self.nz
        for a in range(3):
            c.bc[a] = _BC[self.bc[a]]
        c.model = self.model_code()
        c.halfway = int(self.halfway)
        c.storage = _STORE[self.storage]
        c.omega = self.omega
        c.cs = self.cs if self.les else 0.0
        c.dpdy = self.dpdy
        c.umax = self.umax
        c.inlet_amp = self.inlet_amp
        c.seed = self.seed
        c.geometry = _GEOM[self.geometry]
        c.obst_x, c.obst_y, c.obst_z = self.obst1
        c.obst_r = self.obst1_r
        c.startf = self.startf
        c.force_steps = self.force_steps
        c.refinement = int(self.refinement)
        c.lr_nx, c.lr_ny, c.lr_nz = self.lr_size
        c.lr_x0, c.lr_y0, c.lr_z0 = self.lr_loc
        c.temporal_interp = int(self.temporal_interp)
        c.interface_eq = int(self.interface_eq)
        c.dyn_y1, c.kp = int(self.dyn_y1), float(self.kp)
        c.ctrl_every, c.ctrl_start = int(self.ctrl_every), int(self.ctrl_start)
        c.dyn_mode, c.dyn_y2 = int(self.dyn_mode), int(self.dyn_y2)
        c.obst2_x, c.obst2_y, c.obst2_z = self.obst2
        c.obst2_r = self.obst2_r
        c.trip_until, c.trip_z = int(self.trip_until), int(self.trip_z)
        return c


def omega_from_re(umax, obst_r, re):
    """main(): omega = 1/(3 UMAX (2 OBSTR1)/RE + 0.5) (MRTtest/DynInlet3.cu:2867-2869)."""
    return 1.0 / (3.0 * (umax * 2.0 * obst_r / re) + 0.5)


def omega_lr(omega, factor=0.5):
    """main(): omegaLR = 2/(1 + 2(2/omega - 1)) per halving (MRTtest/DynInlet3.cu:2870-2877)."""
    o = omega
    n = int(round(math.log2(1.0 / factor)))
    for _ in range(n):
        o = 2.0 / (1.0 + 2.0 * (2.0 / o - 1.0))
    return o


def _num(s):
    s = s.strip()
    try:
        return int(s)
    except ValueError:
        return float(s)


def _triple(s, sep="x"):
    s = s.strip().strip("()")
    parts = re.split(r"[x,]" if sep == "x" else ",", s)
    return tuple(_num(p) for p in parts if p.strip() != "")


def _split(line, keys):
    """Match the longest known key at the start of a line; return (key, value)."""
    for k in sorted(keys, key=len, reverse=True):
        if line.startswith(k):
            rest = line[len(k):]
            if k == "omega" and rest.strip().startswith("in LR"):
                continue
            return k, rest.strip().lstrip(":").strip()
    return None, None


def read_inputs(path_or_text):
    """Parse a `.inputs` run record into a Config (tolerant to the reference's
    inconsistent separators; unknown lines are ignored)."""
    text = path_or_text
    if "\n" not in path_or_text and not path_or_text.startswith("Base"):
        with open(path_or_text) as fh:
            text = fh.read()
    cfg = Config()
    seen = {}
    for raw in text.splitlines():
        line = raw.rstrip()
        if not line or line.startswith("#"):
            continue
        k, v = _split(line, REFERENCE_KEYS + EXTENSION_KEYS)
        if k is None or v == "":
            continue  # unknown key, or a truncated line: keep the default
        seen[k] = v
    if "Base domain size" in seen:
        cfg.nx, cfg.ny, cfg.nz = _triple(seen["Base domain size"])
    if "Obst1 location:" in seen:
        cfg.obst1 = tuple(float(a) for a in _triple(seen["Obst1 location:"], ","))
    if "Obst1 radius:" in seen:
        cfg.obst1_r = float(seen["Obst1 radius:"])
    if "Obst2 location:" in seen:
        cfg.obst2 = tuple(float(a) for a in _triple(seen["Obst2 location:"], ","))
    if "Obst2 radius:" in seen:
        cfg.obst2_r = float(seen["Obst2 radius:"])
    for key, attr, typ in (("RE:", "re", float), ("UMAX:", "umax", float), ("DPDY", "dpdy", float),
                           ("TMAX:", "tmax", int), ("STARTF:", "startf", int),
                           ("START_VELAV:", "start_velav", int), ("START_VELFLUC:", "start_velfluc", int),
                           ("CS:", "cs", float), ("omega in LR", "omega_lr", float),
                           ("LR factor", "lr_factor", float), ("GPUs per node", "gpus", int),
                           ("KP:", "kp", float), ("DYNY1:", "dyn_y1", int), ("DYNY2:", "dyn_y2", int),
                           ("Dynamic inlet mode:", "dyn_mode", int), ("Trip until:", "trip_until", int),
                           ("Trip z:", "trip_z", int)):
        if key in seen:
            setattr(cfg, attr, typ(float(seen[key])) if typ is int else typ(seen[key]))
    if "REFINEMENT:" in seen:
        cfg.refinement = bool(int(float(seen["REFINEMENT:"])))
    if "MODEL:" in seen:
        cfg.model = seen["MODEL:"].strip()
    if "Smagorinsky LES:" in seen:
        cfg.les = bool(int(float(seen["Smagorinsky LES:"])))
    if "LR domain size" in seen:
        cfg.lr_size = tuple(int(a) for a in _triple(seen["LR domain size"]))
    if "LR location" in seen:
        cfg.lr_loc = tuple(float(a) for a in _triple(seen["LR location"]))
    if "omega" in seen:
        cfg.omega = float(seen["omega"])
    elif cfg.re > 0:
        cfg.omega = omega_from_re(cfg.umax, cfg.obst1_r, cfg.re)
    ext = {"MRT rates:": ("mrt_rates", str), "Halfway bounce-back:": ("halfway", lambda s: bool(int(s))),
           "Storage:": ("storage", str), "Geometry:": ("geometry", str),
           "Inlet perturbation:": ("inlet_amp", float), "Seed:": ("seed", int),
           "Temporal interpolation:": ("temporal_interp", lambda s: bool(int(s))),
           "Force steps:": ("force_steps", int), "Statistics:": ("velav", lambda s: bool(int(s)))}
    for k, (attr, typ) in ext.items():
        if k in seen:
            setattr(cfg, attr, typ(seen[k].strip()))
    if "START_VELAV:" in seen and "Statistics:" not in seen:
        cfg.velav = True  # the reference programs that print START_VELAV build with VELAV 1
    bc = list(cfg.bc)
    for a, k in enumerate(("Boundary x:", "Boundary y:", "Boundary z:")):
        if k in seen:
            bc[a] = seen[k].strip()
    cfg.bc = tuple(bc)
    if "Inlet control:" in seen:  # "<every>x<start>"
        every, start = (int(float(a)) for a in seen["Inlet control:"].split("x"))
        cfg.ctrl_every, cfg.ctrl_start = every, start
    return cfg


def _g(v):
    """Format like C++ ostream's default (6 significant digits, %g)."""
    if isinstance(v, int):
        return str(v)
    return "%g" % v


def write_inputs(cfg, path=None):
    """Emit the reference's WriteInputs echo, then the extension keys."""
    o1 = cfg.obst1
    o2 = cfg.obst2
    lines = [
        f"Base domain size \t{cfg.nx}x{cfg.ny}x{cfg.nz}",
        "Base blocksize: \t256x1x1",
        f"Obst1 location: \t({_g(o1[0])},{_g(o1[1])},{_g(o1[2])})",
        f"Obst1 radius: \t{_g(cfg.obst1_r)}",
        f"Obst2 location: \t({_g(o2[0])},{_g(o2[1])},{_g(o2[2])})",
        f"Obst2 radius: \t{_g(cfg.obst2_r)}",
        f"RE: \t{_g(cfg.re)}",
        f"UMAX: \t{_g(cfg.umax)}",
        f"omega \t: {_g(cfg.omega)}",
        f"DPDY \t: {_g(cfg.dpdy)}",
        f"TMAX: \t{cfg.tmax}",
        f"STARTF: \t{cfg.startf}",
        f"START_VELAV: \t{cfg.start_velav}",
        f"START_VELFLUC: \t{cfg.start_velfluc}",
        f"REFINEMENT: \t{int(cfg.refinement)}",
        f"MODEL: \t{cfg.model}",
        f"Smagorinsky LES: \t{int(cfg.les)}",
        f"CS: \t{_g(cfg.cs)}",
        f"LR domain size \t{cfg.lr_size[0]}x{cfg.lr_size[1]}x{cfg.lr_size[2]}",
        f"LR factor \t{_g(cfg.lr_factor)}",
        f"LR location \t{_g(cfg.lr_loc[0])}x{_g(cfg.lr_loc[1])}x{_g(cfg.lr_loc[2])}",
        "LR blocksize: \t256x1x1",
        f"omega in LR \t: {_g(cfg.omega_lr or omega_lr(cfg.omega, cfg.lr_factor))}",
        f"GPUs per node \t: {cfg.gpus}",
        "# lbm-mi355x",
        f"MRT rates: \t{cfg.mrt_rates}",
        f"Boundary x: \t{cfg.bc[0]}",
        f"Boundary y: \t{cfg.bc[1]}",
        f"Boundary z: \t{cfg.bc[2]}",
        f"Halfway bounce-back: \t{int(cfg.halfway)}",
        f"Storage: \t{cfg.storage}",
        f"Geometry: \t{cfg.geometry}",
        f"Inlet perturbation: \t{_g(cfg.inlet_amp)}",
        f"Seed: \t{cfg.seed}",
        f"Temporal interpolation: \t{int(cfg.temporal_interp)}",
        f"Force steps: \t{cfg.force_steps}",
        f"Statistics: \t{int(cfg.velav)}",
        f"DYNY1: \t{cfg.dyn_y1}",
        f"KP: \t{_g(cfg.kp)}",
        f"Inlet control: \t{cfg.ctrl_every}x{cfg.ctrl_start}",
        f"DYNY2: \t{cfg.dyn_y2}",
        f"Dynamic inlet mode: \t{cfg.dyn_mode}",
        f"Trip until: \t{cfg.trip_until}",
        f"Trip z: \t{cfg.trip_z}",
    ]
    text = "\n".join(lines) + "\n"
    if path:
        with open(path, "w") as fh:
            fh.write(text)
    return text


# ---------------------------------------------------------------- workloads

def taylor_green(n=64, tau=0.6):
    """64^3 fully periodic BGK, tau=0.6 (Taylor-Green initial state)."""
    return Config(nx=n, ny=n, nz=n, model="BGK", omega=1.0 / tau, tmax=1000)


def channel(nx=512, ny=128, nz=128, side=16, x_center=128, tau=0.52, cs_smag=0.1, u0=0.05,
            amp=1e-3, seed=0, storage="f32", rates="dhumieres", force_steps=0):
    """Channel with a square cylinder (side `side` cells at x=x_center,
    centred in y and z), halfway walls at the y/z bounds, West_Extrap
    velocity inlet with a seeded perturbation and an East_Extrap outlet."""
    half = side / 2.0
    return Config(nx=nx, ny=ny, nz=nz, bc=("clamp", "wall", "wall"), model="MRT", mrt_rates=rates,
                  les=cs_smag > 0, cs=cs_smag * cs_smag, omega=1.0 / tau, umax=u0, halfway=True,
                  storage=storage, geometry="channel",
                  obst1=(x_center - 0.5, (ny - 1) / 2.0, (nz - 1) / 2.0), obst1_r=half,
                  inlet_amp=amp, seed=seed, force_steps=force_steps, tmax=5000)


def cpu3d_case():
    """The reference CPU program's own case (CPU_3D/CPU_3D.cpp:17-31):
    251x43x43, circular cylinder R=5 at (50.5, 20.5), moving regularised
    x=0 plane (v=UMAX), node walls elsewhere, clamped streaming, MRT with
    m13/m14 at rate 1, RE=50, UMAX=0.04, 100 steps."""
    import numpy as np

    f = np.float32
    umax, re, r = 0.04, 50.0, 5.0
    # FP32 as main() computes it (CPU_3D/CPU_3D.cpp:1221, :1299)
    om = float(f(1) / (f(3) * (f(umax) * (f(r) * f(2)) / f(re)) + f(0.5)))
    return Config(nx=251, ny=43, nz=43, bc=("clamp", "clamp", "clamp"), model="MRT",
                  mrt_rates="cpu3d", omega=om, umax=umax, re=re, geometry="cpu3d",
                  obst1=(50.5, 20.5, 20.5), obst1_r=r, tmax=100)


def patch_offset(seed):
    """Seeded integer offset in [-4, 4]^3 (counter-based hash, as the inlet)."""
    from . import load_library

    lib = load_library()
    return tuple(int(lib.lbm_hash_u32(seed, 1_000_000 + a) % 9) - 4 for a in range(3))


def refined_channel(nx=256, ny=128, nz=128, side=16, x_center=64, patch=(128, 96, 96), tau=0.52, cs_smag=0.1,
                    u0=0.05, amp=1e-3, seed=0, rates="dhumieres", temporal=True, storage="f32"):
    """The channel with a 2:1 fine patch of `patch` fine nodes around the
    cylinder; the patch's coarse origin is centred on (x_center + side,
    ny/2, nz/2) and shifted by a seeded offset within +-4 coarse cells."""
    c = channel(nx=nx, ny=ny, nz=nz, side=side, x_center=x_center, tau=tau, cs_smag=cs_smag, u0=u0, amp=amp,
                seed=seed, storage=storage, rates=rates)
    off = patch_offset(seed)
    half = [p // 4 for p in patch]            # half the patch span in coarse cells
    centre = (x_center + side, ny // 2, nz // 2)
    a = [centre[k] - half[k] + off[k] for k in range(3)]
    c.refinement = True
    c.lr_size = tuple(patch)
    c.lr_loc = tuple(float(a[k]) + 0.25 for k in range(3))
    c.temporal_interp = temporal
    c.tmax = 2000
    return c


def precursor(nx=42, ny=96, nz=8, dyn_y1=24, side=12, obst_y=60, trip=(20, 10, 4), trip_until=200,
              tau=0.55, cs_smag=0.1, u0=0.06, dpdy=-5e-7, kp=0.3, ctrl=(100, 1000), storage="f32"):
    """A y-streamwise channel with a periodic inflow precursor in the form of
    bgk_channel1/Re3000_Cyl2/DynInlet3.cu (its 224x800x4 case scaled down):
    node walls at x = 0 and x = nx-1, rows 0..DYNY1 periodic in y and driven
    by the P-controlled body force, the recycling plane y = DYNY1+1 (im 80)
    feeding the channel above, a force-summed square cylinder of side `side`
    (Obst1) at y = obst_y, a trip obstacle (Obst2, half-width trip[2] at
    (trip[0], trip[1])) for t < trip_until, and the rho = 1 north outlet."""
    half = side / 2.0
    return Config(nx=nx, ny=ny, nz=nz, bc=("periodic", "periodic", "periodic"), model="MRT", mrt_rates="lrmgpu",
                  les=cs_smag > 0, cs=cs_smag * cs_smag, omega=1.0 / tau, umax=u0, dpdy=dpdy,
                  geometry="precursor", obst1=(nx / 2.0 - 0.5, obst_y + 0.5, nz / 2.0), obst1_r=half,
                  obst2=(trip[0] + 0.5, trip[1] + 0.5, nz / 2.0), obst2_r=float(trip[2]), trip_until=trip_until,
                  trip_z=0, dyn_y1=dyn_y1, dyn_mode=1, kp=kp, ctrl_every=ctrl[0], ctrl_start=ctrl[1],
                  storage=storage, tmax=2000)
