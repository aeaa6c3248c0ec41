"""Storage-precision study: the channel at 1024x256x256 with populations in
FP16 and BF16 (shifted deviations f_i - w_i, FP32 collide), FP32, and FP64.

Every run starts from the same seeded state; after `steps` steps the velocity
and density fields are compared with the FP64 run (L2 and Linf), and each
run's global mass drift is taken from FP64 sums over the nodes
(lbm_get_mass_momentum).  Throughput of each storage is measured with HIP
events on the same runs.

Run: python -m lbm_multigpu_amd.precision [--scale S] [--steps N] [--out FILE]
"""
import argparse
import json
import time

import numpy as np

from . import Solver, INIT_BGK
from . import config as cfgmod


def workload(scale=1.0, storage="f32"):
    """The 512x128x128 channel scaled by 2*scale in each direction (cylinder
    side, position and inlet recipe scaled alike; tau 0.52, U 0.05)."""
    s = 2.0 * scale
    return cfgmod.channel(nx=int(512 * s), ny=int(128 * s), nz=int(128 * s), side=int(16 * s),
                          x_center=int(128 * s), storage=storage)


def run_one(cfg, steps):
    with Solver(cfg) as s:
        s.set_geometry()
        s.init_equilibrium(INIT_BGK)
        m0 = s.mass_momentum()
        ms = s.step_timed(steps)
        m1 = s.mass_momentum()
        rho, u, v, w = s.moments()
    return {"rho": rho, "u": np.stack([u, v, w]), "mass0": m0[0], "mass1": m1[0], "ms": ms}


def errors(r, ref, solid):
    fl = ~solid
    du = (r["u"] - ref["u"])[:, fl]
    uref = ref["u"][:, fl]
    drho = (r["rho"] - ref["rho"])[fl]
    return {"u_l2": float(np.sqrt((du ** 2).sum() / (uref ** 2).sum())),
            "u_linf": float(np.abs(du).max() / np.sqrt((uref ** 2).sum(0)).max()),
            "rho_l2": float(np.sqrt((drho ** 2).sum() / ((ref["rho"][fl] - 1.0) ** 2).sum())),
            "rho_linf_abs": float(np.abs(drho).max())}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--scale", type=float, default=1.0, help="1: 1024x256x256")
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    t0 = time.time()
    res = {"steps": a.steps}
    ref = None
    solid = None
    for st in ("f64", "f32", "f16", "bf16"):
        cfg = workload(a.scale, st)
        r = run_one(cfg, a.steps)
        if solid is None:
            with Solver(cfg) as s:
                s.set_geometry()
                solid = np.isin(s.get_mask(), (1, 2))
        cells = cfg.nx * cfg.ny * cfg.nz
        out = {"grid": [cfg.nx, cfg.ny, cfg.nz], "glups": cells * a.steps / (r["ms"] * 1e-3) / 1e9,
               "ms_per_step": r["ms"] / a.steps, "mass_drift": (r["mass1"] - r["mass0"]) / r["mass0"]}
        if st == "f64":
            ref = r
        else:
            out.update(errors(r, ref, solid))
        res[st] = out
        print(st, json.dumps(out), flush=True)
        del r
    res["wall_s"] = time.time() - t0
    txt = json.dumps(res, indent=1)
    if a.out:
        with open(a.out, "w") as fh:
            fh.write(txt + "\n")
    print(txt)


if __name__ == "__main__":
    main()
