"""Learned eddy-viscosity closure: data generation, training and evaluation.

The closure network (include/lbm.h, lbm_mlp_*) replaces the constant-Cs
Smagorinsky rate of the reference's MRT-LES collide
(MRTtest/DynInlet3.cu:628-735).  This module drives it end to end on the
solver's own data:

1. a fine channel run (2x the target resolution) is advanced past its
   start-up; every `every` steps the library box-filters it by 2 and returns
   the closure inputs and the eddy viscosity nu_t = max(0, -tau_dev:S /
   (2 S:S)) of a hashed fraction of the filtered cells (lbm_closure_samples);
2. the network is trained on the GPU by the library's own forward, backward
   and Adam kernels (lbm_mlp_train), with the inputs standardised and their
   training range recorded (outside it the collide falls back to
   Smagorinsky);
3. coarse runs with the learned closure (evaluated every step, and every k
   steps with the rates held in between) and with Smagorinsky, and Smagorinsky
   runs at intermediate resolutions, collect the time-averaged velocity with
   the fused running statistics; each is compared with the fine run's mean
   at the coarse cell centres, and the GPU time of each run is measured with
   HIP events.
4. optionally (calibrate) the network's output scale and the Smagorinsky
   constant are chosen a posteriori: short coarse runs over a grid of each,
   scored against the fine mean over the training window (the time of the
   training snapshots); every model is then compared over the held-out test
   window (the time of the last 20 % of the snapshots), which neither the
   training nor the calibration saw.

Every resolution uses the same physical set-up: channel length 4 H, square
cylinder of side H/8 at x = H, halfway walls, seeded inlet perturbation,
U = 0.05 (acoustic scaling) and the fine run's viscosity, so
tau_r = 1/2 + 3 nu_fine r for a grid r times the fine one.

Run: python -m lbm_multigpu_amd.closure [--scale S] [--out FILE]
"""
import argparse
import json
import math
import time

import numpy as np

from . import Mlp, Solver
from . import config as cfgmod

FEATURES = ("log|S|", "3sqrt6 det(S^)", "S^xy", "S^xz", "S^yz", "log d_wall", "|u|")


def channel_at(r, base=(512, 128, 128), tau_fine=0.52, cs_smag=0.1, seed=0):
    """The channel case at r times the base grid (r = 1: the fine run)."""
    nx, ny, nz = (int(round(b * r)) for b in base)
    nu = (tau_fine - 0.5) / 3.0 * r
    return cfgmod.channel(nx=nx, ny=ny, nz=nz, side=int(round(16 * r * base[1] / 128)),
                          x_center=int(round(128 * r * base[1] / 128)), tau=0.5 + 3.0 * nu,
                          cs_smag=cs_smag, seed=seed)


def generate_samples(cfg, spinup, snapshots, every, keep=0.02, seed=1, max_per_snapshot=200000,
                     stats_from=None, split=None, with_model_stress=False, log=print):
    """Advance the fine run and collect box-filtered training samples.

    Returns (X, y, snap_id, solver-state dict with the fine mean velocity
    when stats_from is given -- also the mean up to snapshot `split` as
    "mean_train" -- GPU ms)."""
    s = Solver(cfg)
    s.set_geometry()
    s.init_equilibrium()
    s.set_closure_target(with_model_stress)
    if stats_from is not None:
        s.stats_enable(stats_from, 1 << 62)
    ms = s.step_timed(spinup)
    Xs, ys, ids = [], [], []
    for k in range(snapshots):
        ms += s.step_timed(every)
        X, y = s.closure_samples(seed + k, keep, max_per_snapshot)
        Xs.append(X)
        ys.append(y)
        ids.append(np.full(len(y), k, np.int32))
        if stats_from is not None and split is not None and k + 1 == split:
            out_train = s.stats()[0]
        if (k + 1) % 20 == 0:
            log(f"  fine snapshot {k + 1}/{snapshots}: {len(y)} samples, t={s.time()}")
    out = {"mask": s.get_mask()}
    if stats_from is not None:
        out["mean"] = s.stats()[0]
        if split is not None:
            out["mean_train"] = out_train
    s.close()
    return np.concatenate(Xs), np.concatenate(ys), np.concatenate(ids), out, ms


def smagorinsky_nu(X, c):
    """A-priori Smagorinsky viscosity c sqrt(2 S:S) from the inputs
    (|S_dev| = exp(x0); the solver's CS = Cs^2 plays the role of c)."""
    return c * math.sqrt(2.0) * np.exp(X[:, 0].astype(np.float64))


def fit(X, y, seed=0, steps=3000, batch=4096, lr=2e-3, lr_final=2e-4, log=print):
    """Train a network on (X, y); returns (mlp, norm dict, loss history)."""
    rng = np.random.default_rng(seed)
    perm = rng.permutation(len(y))
    X, y = X[perm], y[perm]
    mu = X.mean(0).astype(np.float32)
    sd = (X.std(0) + 1e-6).astype(np.float32)
    lo, hi = X.min(0).astype(np.float32), X.max(0).astype(np.float32)
    nu_scale = float(max(y.mean(), 1e-12))
    m = Mlp(seed)
    m.set_norm(mu, sd, lo, hi, nu_scale)
    m.set_data(X, y)
    hist = []
    chunks = 10
    for c in range(chunks):
        # geometric learning-rate decay in ten stages
        lr_c = lr * (lr_final / lr) ** (c / max(chunks - 1, 1))
        h = m.train(steps // chunks, min(batch, len(y)), lr_c)
        hist.append(h)
        log(f"  train stage {c + 1}/{chunks}: lr {lr_c:.2e} loss {h[-50:].mean():.4f}")
    norm = {"mu": mu.tolist(), "sd": sd.tolist(), "lo": lo.tolist(), "hi": hi.tolist(), "nu_scale": nu_scale}
    return m, norm, np.concatenate(hist)


def mean_velocity(cfg, spinup, avg_steps, mlp=None, refresh=1, split=None, smag_base=False):
    """Time-averaged velocity of a run after its start-up; returns (means,
    mask, GPU ms of the whole run, ms per step), where means holds the mean
    u[3, nz, ny, nx] over the whole averaging window ("all") and, with
    split, over its first `split` steps ("train") and the rest ("test").
    With a network, it is evaluated every `refresh` steps (on top of the
    Smagorinsky relaxation time with smag_base)."""
    s = Solver(cfg)
    s.set_geometry()
    s.init_equilibrium()
    if mlp is not None:
        s.set_closure(mlp, every=refresh, smagorinsky_base=smag_base)
    s.stats_enable(spinup, 1 << 62)
    ms = s.step_timed(spinup)
    means = {}
    if split:
        ms2 = s.step_timed(split)
        means["train"] = s.stats()[0].astype(np.float64)
        ms2 += s.step_timed(avg_steps - split)
    else:
        ms2 = s.step_timed(avg_steps)
    means["all"] = s.stats()[0].astype(np.float64)
    if split:
        means["test"] = (means["all"] * avg_steps - means["train"] * split) / (avg_steps - split)
    mask = s.get_mask()
    s.close()
    return means, mask, ms + ms2, ms2 / avg_steps


def at_coarse_centres(field, r_src, r_dst):
    """Trilinear samples of a node field of the r_src grid at the node
    positions of the r_dst grid (node i of grid r sits at (i + 1/2)/r in
    fine-grid units); r_src = 2 r_dst gives the 2x box filter."""
    from scipy.ndimage import map_coordinates

    nz, ny, nx = (int(round(n * r_dst / r_src)) for n in field.shape)
    zz, yy, xx = np.meshgrid(*((np.arange(n) + 0.5) / r_dst * r_src - 0.5 for n in (nz, ny, nx)), indexing="ij")
    return map_coordinates(field.astype(np.float64), [zz, yy, xx], order=1, mode="nearest")


def eval_region(mask_c):
    """Coarse cells used for the error: fluid cells at least two cells from
    any solid node (the obstacle's staircase differs between grids)."""
    from scipy.ndimage import binary_dilation

    solid = np.isin(mask_c, (1, 2))
    return ~binary_dilation(solid, iterations=2)


def l2_errors(u, u_ref, region, u0):
    d = (u - u_ref)[region]
    return {"rel_l2": float(np.linalg.norm(d) / np.linalg.norm(u_ref[region])),
            "wake_l2": float(np.linalg.norm(d) / np.linalg.norm((u_ref - u0)[region]))}


ALPHAS = (0.25, 0.5, 1.0, 2.0, 4.0, 8.0)     # network output scales tried a posteriori
CS_GRID = (0.03, 0.05, 0.07, 0.1, 0.14, 0.2)  # Smagorinsky constants tried a posteriori
HYBRID_ALPHAS = (0.5, 1.0, 2.0, 4.0)          # network scales on top of Smagorinsky


def run(scale=1.0, spinup_fine=16000, snapshots=400, every=100, keep=0.02, train_steps=3000,
        resolutions=(0.5625, 0.625, 0.6875, 0.75), seed=0, refresh=(1, 10), tau_fine=0.52, calibrate=False,
        alphas=ALPHAS, cs_grid=CS_GRID, target="resolved", log=print):
    """The full experiment at `scale` times the 512x128x128 fine grid.

    Coarse runs (half the fine resolution): the learned closure at each
    network interval in `refresh`, Smagorinsky Cs = 0.1 (the solver's
    constant), Smagorinsky with the a-priori best-fit constant, and no
    sub-grid model; then Smagorinsky Cs = 0.1 at each of `resolutions`
    (fractions of the fine grid), the curve against which the learned run's
    error is priced in GPU time.  Every run's error is reported over the
    whole averaging window ("rel_l2", "wake_l2"), the training window
    ("train") and the held-out test window ("test").

    With calibrate, the network's output scale alpha (nu_t -> alpha nu_t)
    and the Smagorinsky constant are each chosen from a grid by the wake
    error over the training window of coarse runs (the network evaluated
    every max(refresh) steps there), and the runs "coarse_learned_calibrated"
    (every step) and "coarse_smagorinsky_calibrated" are added; so is the
    learned correction to Smagorinsky (the network's nu_t on top of the
    Smagorinsky relaxation time, "coarse_hybrid_calibrated"), its constant
    and scale chosen together on the same grid.

    target: "resolved" trains on the filtered field's subfilter stress
    alone; "consistent" adds the fine run's own Smagorinsky stress, filtered
    (lbm_set_closure_target), which the filtered fine equations carry too."""
    t0 = time.time()
    base = tuple(int(round(b * scale)) for b in (512, 128, 128))
    fine = channel_at(1.0, base, tau_fine=tau_fine, seed=seed)
    u0 = fine.umax
    fine_avg = snapshots * every
    # hold out the last 20% of the snapshots (and of the averaging window)
    split = int(snapshots * 0.8)
    log(f"fine run {base} tau {1 / fine.omega:.4f}: spin-up {spinup_fine}, {snapshots} snapshots every {every}")
    X, y, ids, fout, ms_fine = generate_samples(fine, spinup_fine, snapshots, every, keep, seed + 1,
                                                stats_from=spinup_fine, split=split,
                                                with_model_stress=(target == "consistent"), log=log)
    log(f"  {len(y)} samples, nu_t mean {y.mean():.3e}, zero fraction {np.mean(y == 0):.3f}")
    tr, te = ids < split, ids >= split
    mlp, norm, hist = fit(X[tr], y[tr], seed, train_steps, log=log)
    pred = mlp.forward(X[te])
    c_fit = float(np.dot(smagorinsky_nu(X[tr], 1.0), y[tr]) / np.dot(smagorinsky_nu(X[tr], 1.0),
                                                                      smagorinsky_nu(X[tr], 1.0)))
    var = float(np.mean((y[te] - y[te].mean()) ** 2))
    apriori = {"mse_mlp": float(np.mean((pred - y[te]) ** 2)),
               "mse_smag_bestfit": float(np.mean((smagorinsky_nu(X[te], c_fit) - y[te]) ** 2)),
               "mse_smag_cs0.1": float(np.mean((smagorinsky_nu(X[te], 0.01) - y[te]) ** 2)),
               "var_target": var, "c_bestfit": c_fit,
               "corr_mlp": float(np.corrcoef(pred, y[te])[0, 1]),
               "corr_smag": float(np.corrcoef(smagorinsky_nu(X[te], 1.0), y[te])[0, 1])}
    log(f"  a-priori test MSE: mlp {apriori['mse_mlp']:.3e}, best-fit Smagorinsky {apriori['mse_smag_bestfit']:.3e}")
    f_all = fout["mean"][0].astype(np.float64)
    f_train = fout["mean_train"][0].astype(np.float64)
    f_test = (f_all * snapshots - f_train * split) / (snapshots - split)
    u_ref = {w: at_coarse_centres(f, 1.0, 0.5) for w, f in (("all", f_all), ("train", f_train), ("test", f_test))}
    # the same physical time as the fine run: half the steps
    sp_c, av_c = spinup_fine // 2, fine_avg // 2
    cs_fit = math.sqrt(max(c_fit, 0.0))
    runs = {}
    region = None

    def coarse_run(name, r, k, cs, alpha=1.0, smag_base=False):
        nonlocal region
        cfg = channel_at(r, base, tau_fine=tau_fine, cs_smag=cs, seed=seed)
        sp, av = int(round(spinup_fine * r)), int(round(fine_avg * r))
        if k:
            mlp.set_norm(norm["mu"], norm["sd"], norm["lo"], norm["hi"], alpha * norm["nu_scale"])
        means, mask, ms, ms_step = mean_velocity(cfg, sp, av, mlp if k else None, max(k, 1),
                                                 split=int(round(av * split / snapshots)), smag_base=smag_base)
        if k:
            mlp.set_norm(norm["mu"], norm["sd"], norm["lo"], norm["hi"], norm["nu_scale"])
        if r == 0.5 and region is None:
            region = eval_region(mask)
        u_c = {w: (at_coarse_centres(m[0], r, 0.5) if r != 0.5 else m[0]) for w, m in means.items()}
        rr = {"grid": [cfg.nx, cfg.ny, cfg.nz], "tau": 1 / cfg.omega, "steps": sp + av, "cs": cs,
              "gpu_ms": ms, "ms_per_step": ms_step}
        if k:
            rr["alpha"] = alpha
            rr["base"] = "smagorinsky" if smag_base else "tau0"
        rr.update(l2_errors(u_c["all"], u_ref["all"], region, u0))
        for w in ("train", "test"):
            rr[w] = l2_errors(u_c[w], u_ref[w], region, u0)
        runs[name] = rr
        log(f"  {name}: grid {cfg.nx}x{cfg.ny}x{cfg.nz}, {sp + av} steps, {ms / 1e3:.2f} s GPU, "
            f"wake L2 train {rr['train']['wake_l2']:.4f} test {rr['test']['wake_l2']:.4f}")
        return rr

    cases = [("coarse_learned" if k == 1 else f"coarse_learned_refresh{k}", 0.5, k, 0.1) for k in refresh]
    cases += [("coarse_smagorinsky", 0.5, 0, 0.1), ("coarse_smagorinsky_bestfit", 0.5, 0, cs_fit),
              ("coarse_no_model", 0.5, 0, 0.0)]
    cases += [(f"smagorinsky_r{r}", r, 0, 0.1) for r in resolutions]
    for name, r, k, cs in cases:
        coarse_run(name, r, k, cs)
    assert sp_c + av_c == runs["coarse_smagorinsky"]["steps"]
    calib = None
    if calibrate:
        kc = max(refresh)
        scores_a = {}
        for a in alphas:
            rr = coarse_run(f"calib_alpha{a}", 0.5, kc, 0.1, alpha=a)
            scores_a[a] = rr["train"]["wake_l2"]
        scores_c = {}
        for c in cs_grid:
            rr = runs.get("coarse_smagorinsky") if c == 0.1 else coarse_run(f"calib_cs{c}", 0.5, 0, c)
            scores_c[c] = rr["train"]["wake_l2"]
        scores_h = {}
        for c in cs_grid:
            for a in HYBRID_ALPHAS:
                rr = coarse_run(f"calib_hybrid_cs{c}_alpha{a}", 0.5, kc, c, alpha=a, smag_base=True)
                scores_h[(c, a)] = rr["train"]["wake_l2"]
        finite = lambda d: {k: v for k, v in d.items() if math.isfinite(v)}
        a_best = min(finite(scores_a), key=scores_a.get)
        c_best = min(finite(scores_c), key=scores_c.get)
        h_best = min(finite(scores_h), key=scores_h.get)
        calib = {"alpha": a_best, "cs": c_best, "hybrid_cs": h_best[0], "hybrid_alpha": h_best[1],
                 "train_wake_l2_by_alpha": scores_a, "train_wake_l2_by_cs": scores_c,
                 "train_wake_l2_hybrid": {f"{c},{a}": v for (c, a), v in scores_h.items()}}
        log(f"  calibrated: alpha {a_best}, Cs {c_best}, hybrid Cs {h_best[0]} alpha {h_best[1]}")
        coarse_run("coarse_learned_calibrated", 0.5, 1, 0.1, alpha=a_best)
        coarse_run("coarse_smagorinsky_calibrated", 0.5, 0, c_best)
        coarse_run("coarse_hybrid_calibrated", 0.5, 1, h_best[0], alpha=h_best[1], smag_base=True)
    for name, rr in runs.items():
        log(f"  {name}: rel L2 {rr['rel_l2']:.4f}, wake L2 {rr['wake_l2']:.4f}, test wake L2 {rr['test']['wake_l2']:.4f}")
    res = {"fine": {"grid": list(base), "tau": 1 / fine.omega, "steps": spinup_fine + fine_avg, "gpu_ms": ms_fine,
                    "samples": int(len(y)), "target": target},
           "apriori": apriori, "training": {"steps": train_steps, "loss_first": float(hist[:50].mean()),
                                            "loss_last": float(hist[-50:].mean())},
           "norm": norm, "runs": runs, "calibration": calib, "wall_s": time.time() - t0,
           "eval_cells": int(region.sum()),
           "windows": {"train_coarse_steps": [sp_c, sp_c + av_c * split // snapshots],
                       "test_coarse_steps": [sp_c + av_c * split // snapshots, sp_c + av_c]}}
    pts = sorted((v["gpu_ms"], v["wake_l2"]) for k, v in runs.items()
                 if k.startswith("smagorinsky_r") or k == "coarse_smagorinsky")
    err_learned = runs[cases[0][0]]["wake_l2"]
    tms = [p[0] for p in pts]
    errs = [p[1] for p in pts]
    res["smagorinsky_ms_at_learned_error"], res["smagorinsky_cheapest_meets_learned_error"] = \
        time_for_error(tms, errs, err_learned)
    res["smagorinsky_curve"] = {"gpu_ms": tms, "wake_l2": errs}
    return res


def time_for_error(tms, errs, target):
    """GPU time a resolution curve (points sorted by GPU time) needs to
    reach an error of target: the cheapest point's time when that point
    already meets it (second value True), else the log-log interpolation at
    the cheapest crossing from above, else None."""
    if not tms or not all(math.isfinite(e) for e in errs):
        return None, False
    if errs[0] <= target:
        return tms[0], True
    for (t1, e1), (t2, e2) in zip(zip(tms, errs), zip(tms[1:], errs[1:])):
        if e1 > target >= e2:
            a = (math.log(target) - math.log(e1)) / (math.log(e2) - math.log(e1))
            return math.exp(math.log(t1) + a * (math.log(t2) - math.log(t1))), False
    return None, False


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--scale", type=float, default=1.0, help="grid scale (1: the 512x128x128 fine run)")
    ap.add_argument("--spinup", type=int, default=16000, help="fine-run steps before sampling")
    ap.add_argument("--snapshots", type=int, default=400)
    ap.add_argument("--every", type=int, default=100)
    ap.add_argument("--keep", type=float, default=0.02)
    ap.add_argument("--train-steps", type=int, default=3000)
    ap.add_argument("--refresh", type=int, nargs="+", default=[1, 10],
                    help="network evaluation intervals of the learned coarse runs")
    ap.add_argument("--tau", type=float, default=0.52, help="relaxation time of the fine run")
    ap.add_argument("--calibrate", action="store_true",
                    help="choose the network output scale and the Smagorinsky constant a posteriori")
    ap.add_argument("--target", choices=["resolved", "consistent"], default="resolved",
                    help="training target: the subfilter stress alone, or with the fine run's own model stress")
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    res = run(a.scale, a.spinup, a.snapshots, a.every, a.keep, a.train_steps, refresh=tuple(a.refresh),
              tau_fine=a.tau, calibrate=a.calibrate, target=a.target, log=lambda *m: print(*m, flush=True))
    txt = json.dumps(res, indent=1)
    if a.out:
        with open(a.out, "w") as fh:
            fh.write(txt + "\n")
    print(txt)


if __name__ == "__main__":
    main()
