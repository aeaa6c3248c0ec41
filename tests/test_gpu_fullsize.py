"""The bench's channel at its full size (512x128x128, square cylinder side 16,
MRT + Smagorinsky), checked through properties that do not depend on an
oracle run of that size:

* mirror symmetry: without the inlet perturbation the case is symmetric under
  y -> ny-1-y and under z -> nz-1-z (rod centred at (ny-1)/2, spanning z;
  halfway walls on both faces), so after 200 steps u and w are even and v
  odd in y (u, v even and w odd in z) to FP32 rounding: 1e-5 of max|u|;
* 16-bit storage: FP16 populations with the FP32 collide against the FP64
  run of the same case, relative velocity error <= 1e-3 (BF16: 1e-2), the
  bars the storage format is held to;
* conservation: with the seeded inlet the FP64 mass sums of the FP32 and
  FP64 runs differ by less than 1e-6 relative after 200 steps (the open
  inlet and outlet change the mass of both runs alike; measured 4e-8).
"""
import sys

import numpy as np
import pytest

import lbm_multigpu_amd as L
from lbm_multigpu_amd import config as cfgmod

pytestmark = pytest.mark.gpu
STEPS = 200


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _run(storage, amp):
    c = cfgmod.channel(storage=storage, amp=amp).to_c()
    with L.Solver(c) as s:
        s.set_geometry(None)
        mask = s.get_mask()
        s.init_equilibrium(L.INIT_BGK)
        s.step(STEPS)
        rho, u, v, w = s.moments()
        mass = s.mass_momentum()[0]
    return mask, rho, u, v, w, mass


def _rel_vel(got, ref):
    scale = np.sqrt(sum(r.astype(np.float64) ** 2 for r in ref)).max()
    return max(np.abs(g - r).max() for g, r in zip(got, ref)) / scale


def test_full_size_channel_is_mirror_symmetric():
    say("full-size channel, symmetry")
    mask, rho, u, v, w, _ = _run("f32", 0.0)
    assert np.array_equal(mask, mask[:, ::-1, :]) and np.array_equal(mask, mask[::-1, :, :])
    fl = mask == L.NODE_FLUID
    scale = np.sqrt(u.astype(np.float64) ** 2 + v ** 2 + w ** 2)[fl].max()
    # y mirror (axis 1 of [z][y][x]): u, w even, v odd
    ey = max(np.abs(u - u[:, ::-1, :])[fl].max(), np.abs(v + v[:, ::-1, :])[fl].max(),
             np.abs(w - w[:, ::-1, :])[fl].max()) / scale
    # z mirror (axis 0): u, v even, w odd
    ez = max(np.abs(u - u[::-1])[fl].max(), np.abs(v - v[::-1])[fl].max(), np.abs(w + w[::-1])[fl].max()) / scale
    say(f"  max|u| {scale:.4f}, asymmetry y {ey:.2e}, z {ez:.2e}")
    assert np.isfinite(scale) and scale > 0.04
    assert ey < 1e-5 and ez < 1e-5


@pytest.mark.parametrize("storage,tol", [("f16", 1e-3), ("bf16", 1e-2)])
def test_full_size_16bit_storage_vs_fp64(storage, tol):
    say("full-size channel,", storage, "vs f64")
    _, _, u6, v6, w6, _ = _run("f64", 1e-3)
    _, _, u, v, w, _ = _run(storage, 1e-3)
    err = _rel_vel((u, v, w), (u6, v6, w6))
    say(f"  {storage}: velocity error vs FP64 {err:.3e} (bar {tol})")
    assert err < tol


def test_full_size_fp32_mass_tracks_fp64():
    say("full-size channel, mass")
    *_, m32 = _run("f32", 1e-3)
    *_, m64 = _run("f64", 1e-3)
    say(f"  mass FP32 {m32:.9e}, FP64 {m64:.9e}")
    assert abs(m32 - m64) < 1e-6 * abs(m64)


def test_config4_1024_storage_formats_against_fp64():
    """The channel scaled to 1024x256x256 (cylinder side 32), 2000 steps from
    the same seeded state in FP64, FP32, FP16 and BF16 storage: velocity
    error against the FP64 run, L2 over the fluid and Linf relative to
    max|u|.  Bars: FP16 <= 1e-3 in both (the storage target), FP32 <= 1e-5,
    BF16 (8 significant bits, reported beside) L2 <= 2e-3 and Linf <= 3e-2;
    every format's FP64 mass sum within 1e-6 (FP32, FP16) or 1e-4 (BF16) of
    the FP64 run's."""
    from lbm_multigpu_amd import precision as P

    ref, solid = None, None
    bars = {"f32": (1e-5, 1e-5, 1e-6), "f16": (1e-3, 1e-3, 1e-6), "bf16": (2e-3, 3e-2, 1e-4)}
    for st in ("f64", "f32", "f16", "bf16"):
        say("config 4 storage", st)
        cfg = P.workload(1.0, st)
        assert (cfg.nx, cfg.ny, cfg.nz) == (1024, 256, 256)
        r = P.run_one(cfg, 2000)
        if st == "f64":
            ref = r
            with L.Solver(cfg) as s:
                s.set_geometry()
                solid = np.isin(s.get_mask(), (L.NODE_SOLID, L.NODE_SOLID_FORCE))
            continue
        e = P.errors(r, ref, solid)
        dm = abs((r["mass1"] - ref["mass1"]) / ref["mass1"])
        say(f"  {st}: u L2 {e['u_l2']:.2e}, Linf {e['u_linf']:.2e}, mass vs FP64 {dm:.1e}, "
            f"{cfg.nx * cfg.ny * cfg.nz * 2000 / (r['ms'] * 1e-3) / 1e9:.1f} GLUPS")
        l2, linf, mass = bars[st]
        assert e["u_l2"] <= l2 and e["u_linf"] <= linf and dm <= mass
        del r


def _levels_run(cfg, steps, levels):
    with L.Solver(cfg.to_c()) as s:
        s.set_geometry(None)
        s.init_equilibrium(L.INIT_BGK)
        s.step(steps)
        out = {"mass": s.mass_momentum()[0]}
        for lev in levels:
            rho, u, v, w = s.moments(lev)
            out[lev] = (np.stack([u, v, w]), s.get_mask(lev))
    return out


def _u_errors(got, ref):
    u, mask = got
    ur = ref[0]
    fl = ~np.isin(mask, (L.NODE_SOLID, L.NODE_SOLID_FORCE))
    du = (u - ur)[:, fl]
    return (float(np.sqrt((du ** 2).sum() / (ur[:, fl] ** 2).sum())),
            float(np.abs(du).max() / np.sqrt((ur[:, fl] ** 2).sum(0)).max()))


# velocity L2, Linf (relative to max|u|) and FP64 mass-sum bars against the FP64 run
FULL_BARS = {"f32": (1e-5, 1e-5, 1e-6), "f16": (1e-3, 1e-3, 1e-5), "bf16": (2e-2, 1e-1, 1e-4)}


@pytest.mark.parametrize("workload", ["channel", "refined"])
def test_full_length_storage_formats_against_fp64(workload):
    """The bench channel (512x128x128) over its full 5000 steps, and the
    refined channel (256x128x128 coarse, 128x96x96 fine patch) over its full
    2000 coarse steps, in FP32, FP16 and BF16 storage against the FP64 run
    of the same case (FP64 end to end, the refined interface included):
    FP32 within 1e-5 and FP16 within 1e-3 in both L2 and Linf on every level
    (measured: FP32 9e-7 / 2.6e-6, FP16 1.7e-4 / 8.7e-4 at most); BF16
    reported beside (8.1e-3 / 5.3e-2)."""
    steps, levels = (5000, (0,)) if workload == "channel" else (2000, (0, 1))
    mk = (lambda st: cfgmod.channel(storage=st)) if workload == "channel" else \
        (lambda st: cfgmod.refined_channel(storage=st))
    say(f"full-length {workload}: {steps} steps")
    ref = _levels_run(mk("f64"), steps, levels)
    for st in ("f32", "f16", "bf16"):
        r = _levels_run(mk(st), steps, levels)
        dm = abs(r["mass"] / ref["mass"] - 1)
        l2b, linfb, mb = FULL_BARS[st]
        for lev in levels:
            l2, linf = _u_errors(r[lev], ref[lev])
            say(f"  {st} level {lev}: u L2 {l2:.2e}, Linf {linf:.2e}, mass {dm:.1e}")
            assert l2 <= l2b and linf <= linfb
        assert dm <= mb


def _stated_channel_mask(nx, ny, nz, side, x_center):
    """The channel's node codes built from its statement alone, [z][y][x]:
    a square rod of side `side` cells spanning z, its x cells
    x_center - side/2 .. x_center + side/2 - 1, centred in y; the velocity
    inlet plane x = 0 and the outlet plane x = nx-1; halfway walls on the y
    and z faces (no wall nodes), fluid everywhere else."""
    m = np.full((nz, ny, nx), L.NODE_FLUID, np.uint8)
    h = side // 2
    y0 = ny // 2 - h  # centred: ny/2 - h .. ny/2 + h - 1 is symmetric about (ny-1)/2
    m[:, y0:y0 + side, x_center - h:x_center + h] = L.NODE_SOLID_FORCE
    m[:, :, 0] = L.NODE_INLET_WEST
    m[:, :, nx - 1] = L.NODE_OUTLET_EAST
    return m


@pytest.mark.parametrize("dims", [(512, 128, 128, 16, 128), (256, 64, 64, 8, 64)])
def test_channel_mask_is_the_stated_geometry(dims):
    """The geometry classifier (lbm_set_geometry with no mask) on the bench's
    channel and on the learned-closure workload's coarse channel, node for
    node against the stated geometry."""
    nx, ny, nz, side, xc = dims
    say("channel mask", dims)
    c = cfgmod.channel(nx=nx, ny=ny, nz=nz, side=side, x_center=xc).to_c()
    with L.Solver(c) as s:
        s.set_geometry(None)
        got = s.get_mask()
    exp = _stated_channel_mask(nx, ny, nz, side, xc)
    bad = np.argwhere(got != exp)
    assert bad.size == 0, f"{len(bad)} nodes differ, first [z,y,x] {bad[0]}: {got[tuple(bad[0])]} vs {exp[tuple(bad[0])]}"
    assert (got == L.NODE_SOLID_FORCE).sum() == side * side * nz
