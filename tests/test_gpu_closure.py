"""Learned eddy-viscosity closure on the GPU against the NumPy reference
(tests/mlp_ref.py, itself pinned by finite differences and closed forms in
tests/test_mlp_ref.py) and the oracle's collide.

Tolerances, from the BF16 arithmetic the kernels use (8-bit mantissa):
* forward: against the same BF16 rounding points in NumPy, 5e-3 relative
  (L2 over the batch); against plain FP64, 3e-2;
* gradient: against the BF16 NumPy backward, 2e-2 relative L2 per layer;
  against FP64, 8e-2 per layer and cosine similarity >= 0.995 overall;
* fused collide: one step against the oracle's collide at the per-node rate
  the NumPy network gives, 1e-5 relative to max|f| (a BF16 rounding flip in
  the network moves nu_t by ~1e-3 relative); the closure's own effect on the
  same step is checked to be far above that bar.
"""
import math
import sys

import numpy as np
import pytest

import lbm_multigpu_amd as L
import mlp_ref as M
import oracle_lib as O
from lbm_multigpu_amd import config as cfgmod
from test_gpu_parity import orc_params, random_state

pytestmark = pytest.mark.gpu
INF = np.float32(3e38)
# lattice velocities in the solver's order (MRTtest/DynInlet3.cu:603-605)
CV = np.array([(0, 0, 0), (1, 0, 0), (0, 1, 0), (-1, 0, 0), (0, -1, 0), (1, 1, 0), (-1, 1, 0), (-1, -1, 0),
               (1, -1, 0), (0, 0, 1), (1, 0, 1), (0, 1, 1), (-1, 0, 1), (0, -1, 1), (0, 0, -1), (1, 0, -1),
               (0, 1, -1), (-1, 0, -1), (0, -1, -1)], np.float64)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def synthetic_inputs(seed, n):
    """Raw closure inputs spread like a wake flow's."""
    rng = np.random.default_rng(seed)
    X = np.empty((n, 7), np.float32)
    X[:, 0] = rng.normal(-6.0, 1.0, n)
    X[:, 1] = rng.uniform(-1, 1, n)
    X[:, 2:5] = rng.uniform(-0.7, 0.7, (n, 3))
    X[:, 5] = np.log(rng.uniform(0.5, 30.0, n))
    X[:, 6] = rng.uniform(0.0, 0.08, n)
    return X


def norm_of(X):
    mu = X.mean(0).astype(np.float32)
    sd = X.std(0).astype(np.float32) + np.float32(1e-6)
    return mu, sd


def rl2(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


def test_params_roundtrip_and_device_init():
    m = L.Mlp(seed=9)
    p0 = m.params
    W1 = M.unpack(p0)["W1"]
    a = math.sqrt(6.0 / (7 + 64))
    assert np.abs(W1).max() <= a and W1.std() > 0.45 * a     # uniform on [-a, a]: std a/sqrt(3)
    assert np.all(M.unpack(p0)["b2"] == 0)
    assert np.array_equal(L.Mlp(seed=9).params, p0)          # seeded
    assert not np.array_equal(L.Mlp(seed=10).params, p0)
    p = M.glorot(3)
    m.params = p
    assert np.array_equal(m.params, p)


@pytest.mark.parametrize("n", [1, 1000])
def test_forward_matches_reference(n):
    say("mlp forward n", n)
    X = synthetic_inputs(1, max(n, 64))[:n]
    mu, sd = norm_of(synthetic_inputs(1, 4096))
    p = M.glorot(7)
    nu_scale = 2e-3
    m = L.Mlp(seed=0)
    m.params = p
    m.set_norm(mu, sd, np.full(7, -INF), np.full(7, INF), nu_scale)
    got = m.forward(X)
    xn = M.normalise(X, mu, sd)
    zb, _ = M.forward_bf16(p, xn)
    z64, _ = M.forward64(p, xn)
    nb = nu_scale * M.softplus(zb.astype(np.float64))
    n64 = nu_scale * M.softplus(z64)
    say("  rel vs bf16 ref", rl2(got, nb), "vs fp64", rl2(got, n64))
    assert rl2(got, nb) < 5e-3
    assert rl2(got, n64) < 3e-2


def _layers(g):
    return {k: v.reshape(-1) for k, v in M.unpack(g).items()}


@pytest.mark.parametrize("offset,count", [(0, 300), (100, 150), (0, 2048)])
def test_gradient_matches_reference(offset, count):
    say("mlp gradient", offset, count)
    n = 2048
    X = synthetic_inputs(2, n)
    mu, sd = norm_of(X)
    p = M.glorot(11)
    teacher = M.glorot(12)
    nu_scale = 1e-3
    y = (nu_scale * M.softplus(M.forward64(teacher, M.normalise(X, mu, sd))[0])).astype(np.float32)
    m = L.Mlp()
    m.params = p
    m.set_norm(mu, sd, np.full(7, -INF), np.full(7, INF), nu_scale)
    m.set_data(X, y)
    g, loss = m.gradient(offset, count)
    g2, loss2 = m.gradient(offset, count)
    assert np.array_equal(g, g2) and loss == loss2          # deterministic reduction
    xs, ys = M.normalise(X[offset:offset + count], mu, sd), y[offset:offset + count]
    gb, lb = M.grad_bf16(p, xs, ys, nu_scale)
    g64, l64 = M.grad64(p, xs, ys, nu_scale)
    say("  loss", loss, lb, l64)
    assert abs(loss - lb) < 1e-2 * lb and abs(loss - l64) < 5e-2 * l64
    G, GB, G64 = _layers(g), _layers(gb), _layers(g64)
    for k in G:
        say("  ", k, "vs bf16", rl2(G[k], GB[k]), "vs fp64", rl2(G[k], G64[k]))
        assert rl2(G[k], GB[k]) < 2e-2, k
        assert rl2(G[k], G64[k]) < 8e-2, k
    cos = float(g @ g64 / (np.linalg.norm(g) * np.linalg.norm(g64)))
    assert cos >= 0.995


def test_adam_first_step_and_training():
    say("mlp adam + training")
    n = 8192
    X = synthetic_inputs(3, n)
    mu, sd = norm_of(X)
    nu_scale = 1e-3
    teacher = M.glorot(21)
    y = (nu_scale * M.softplus(M.forward64(teacher, M.normalise(X, mu, sd))[0])).astype(np.float32)
    m = L.Mlp(seed=4)
    p0 = m.params
    m.set_norm(mu, sd, np.full(7, -INF), np.full(7, INF), nu_scale)
    m.set_data(X, y)
    lr = 1e-3
    g0, _ = m.gradient(0, 512)
    hist = m.train(1, 512, lr)
    p1 = m.params
    d = p1.astype(np.float64) - p0
    # the first Adam step: m/(1-b1) = g, v/(1-b2) = g^2, so each parameter
    # moves by -lr g / (|g| + eps)
    g64 = g0.astype(np.float64)
    want = -lr * g64 / (np.abs(g64) + 1e-8)
    say("  adam step max dev / lr", np.abs(d - want).max() / lr)
    assert np.abs(d - want).max() < 2e-3 * lr
    # training reaches a small fraction of the initial loss
    hist = np.concatenate([hist, m.train(600, 512, 3e-3)])
    say("  loss", hist[0], "->", hist[-16:].mean())
    assert np.all(np.isfinite(hist))
    assert hist[-16:].mean() < 0.05 * hist[0]
    # held-out samples from the same distribution
    Xt = synthetic_inputs(30, 2048)
    yt = nu_scale * M.softplus(M.forward64(teacher, M.normalise(Xt, mu, sd))[0])
    pt = m.forward(Xt)
    assert np.mean((pt - yt) ** 2) < 0.1 * np.mean((yt - yt.mean()) ** 2)


def _closure_case():
    c = L.default_config()
    c.nx, c.ny, c.nz = 40, 16, 12
    c.bc[0], c.bc[1], c.bc[2] = L.BC_PERIODIC, L.BC_WALL, L.BC_WALL
    c.halfway = 1
    c.model = L.MODEL_MRT_DHUMIERES
    c.omega = 1 / 0.51
    c.cs = 0.01
    return c


def _pre_collision(c, f):
    """Streaming and boundaries of one step with an identity collide (BGK
    at rate 0 leaves f unchanged)."""
    p = orc_params(c)
    p.collision, p.omega, p.cs, p.dpdy = O.BGK, 0.0, 0.0, 0.0
    fs, _ = O.step(p, f)
    return fs


def _node_features(c, fs, omega_last=None):
    """Closure inputs from the pre-collision populations: S = -3 omega /
    (2 rho) (P - rho/3 I - j j) with omega the rate each node was last
    relaxed at (the base rate omega0 before any), wall distance, |j|."""
    Cv = CV
    f = fs.astype(np.float64)
    rho = f.sum(0)
    j = np.einsum("ia,i...->a...", Cv, f)
    P = np.einsum("ia,ib,i...->ab...", Cv, Cv, f)
    om = float(np.float32(c.omega)) if omega_last is None else omega_last.reshape(rho.shape).astype(np.float64)
    k = -1.5 * om / rho
    S = k * (P - rho / 3.0 * np.eye(3)[:, :, None, None, None] - j[:, None] * j[None, :])
    S = np.moveaxis(S, (0, 1), (-2, -1))
    nz, ny, nx = rho.shape
    zz, yy, _ = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    dw = np.minimum(np.minimum(yy + 0.5, ny - 0.5 - yy), np.minimum(zz + 0.5, nz - 0.5 - zz))
    speed = np.sqrt((j ** 2).sum(0))
    return M.features(S, dw, speed).astype(np.float32)


def _gap_threshold(v, q):
    """A threshold near quantile q lying in a wide gap of the sorted values."""
    s = np.sort(v.reshape(-1))
    i0, i1 = int(len(s) * (q - 0.05)), int(len(s) * (q + 0.05))
    gaps = np.diff(s[i0:i1])
    i = i0 + int(np.argmax(gaps))
    return np.float32(0.5 * (s[i] + s[i + 1])), float(gaps.max())


def _closure_setup(seed):
    """A random near-equilibrium state of the closure case and a network
    whose training range leaves ~30% of the nodes to Smagorinsky."""
    c = _closure_case()
    rng = np.random.default_rng(seed)
    f = random_state(rng, (c.nz, c.ny, c.nx), amp=2e-3, umax=0.03)
    Xf = _node_features(c, _pre_collision(c, f)).reshape(-1, 7)
    mu, sd = norm_of(Xf)
    lo = np.full(7, -INF, np.float32)
    hi = np.full(7, INF, np.float32)
    lo[0], gap = _gap_threshold(Xf[:, 0], 0.3)
    assert gap > 1e-4
    net = dict(p=M.glorot(31), mu=mu, sd=sd, lo=lo, hi=hi, nu_scale=np.float32(0.01))
    m = L.Mlp()
    m.params = net["p"]
    m.set_norm(mu, sd, lo, hi, net["nu_scale"])
    return c, f, m, net


def _reference_rates(c, fs, net, omega_last=None):
    """(in-range flags, per-node rate) the NumPy network gives for the
    pre-collision populations fs."""
    Xf = _node_features(c, fs, omega_last).reshape(-1, 7)
    zb, _ = M.forward_bf16(net["p"], M.normalise(Xf, net["mu"], net["sd"]))
    nut = np.minimum(net["nu_scale"] * M.softplus(zb.astype(np.float64)), 0.05).astype(np.float32)
    inr = np.all((Xf >= net["lo"]) & (Xf <= net["hi"]), axis=1)
    om0 = np.float32(c.omega)
    omega = (np.float32(1) / (np.float32(1) / om0 + np.float32(3) * nut)).astype(np.float32)
    return inr, omega, nut


def _reference_collide(c, fs, inr, omega):
    """The oracle's collide node by node: at the given rate (no LES) where
    inr, at the Smagorinsky rate elsewhere."""
    pin = orc_params(c)
    pin.cs = 0.0
    pout = orc_params(c)
    ref = np.empty_like(fs)
    fr = fs.reshape(19, -1)
    rr = ref.reshape(19, -1)
    for n in range(fr.shape[1]):
        if inr[n]:
            pin.omega = float(omega[n])
            rr[:, n] = O.collide(pin, fr[:, n].copy())
        else:
            rr[:, n] = O.collide(pout, fr[:, n].copy())
    return ref


def test_fused_closure_collide_matches_reference():
    c, f, m, net = _closure_setup(5)
    say("fused closure collide", c.nx, c.ny, c.nz)
    with L.Solver(c) as s:
        s.set_populations(f)
        s.set_closure(m, nu_cap=0.05)
        s.step(1)
        g = s.populations()
    with L.Solver(c) as s:
        s.set_populations(f)
        s.step(1)
        g_smag = s.populations()
    fs = _pre_collision(c, f)
    inr, omega, nut = _reference_rates(c, fs, net)
    say("  in range", inr.mean(), "nu_t", nut[inr].min(), nut[inr].max())
    assert 0.5 < inr.mean() < 0.9
    ref = _reference_collide(c, fs, inr, omega)
    scale = np.abs(ref).max()
    err = np.abs(g - ref).max() / scale
    effect = np.abs(g_smag - ref).max() / scale
    say("  err", err, "closure effect", effect)
    assert err < 1e-5
    assert effect > 50 * 1e-5
    # nodes outside the range take the Smagorinsky rate: identical to the plain solver there
    out = ~inr.reshape(fs.shape[1:])
    assert np.abs(g[:, out] - g_smag[:, out]).max() < 1e-6 * scale


def test_closure_on_smagorinsky_base_matches_reference():
    """lbm_set_closure_base(1): in range, tau = tau_Smagorinsky + 3 nu_t,
    with tau_Smagorinsky the oracle's tau(Q) law at the configured CS
    (MRTtest/DynInlet3.cu:722-729) for the node's pre-collision populations;
    out of range the plain Smagorinsky rate.  One step against the NumPy
    network and the oracle's collide, and measurably apart from the
    tau0-based closure."""
    import ctypes
    c, f, m, net = _closure_setup(11)
    c.cs = 0.04  # Cs = 0.2: the two bases well apart on this near-equilibrium state
    say("closure on the Smagorinsky base")
    out = {}
    for base in (True, False):
        with L.Solver(c) as s:
            s.set_populations(f)
            s.set_closure(m, smagorinsky_base=base)
            s.step(1)
            out[base] = s.populations()
    fs = _pre_collision(c, f)
    inr, _, nut = _reference_rates(c, fs, net)
    pl = orc_params(c)
    fr = fs.reshape(19, -1)
    w = ctypes.c_float()
    omega = np.zeros(fr.shape[1], np.float32)
    for n in np.flatnonzero(inr):
        O.lib().orc_les_omega_f32(ctypes.byref(pl), np.ascontiguousarray(fr[:, n]), ctypes.byref(w), None)
        omega[n] = np.float32(1) / (np.float32(1) / np.float32(w.value) + np.float32(3) * nut[n])
    ref = _reference_collide(c, fs, inr, omega)
    scale = np.abs(ref).max()
    err = np.abs(out[True] - ref).max() / scale
    apart = np.abs(out[False] - ref).max() / scale
    say("  err", err, "vs tau0 base", apart)
    assert err < 1e-5
    assert apart > 20 * 1e-5


def test_closure_strain_uses_the_last_applied_rate():
    """Pi_neq before a collision scales with the rate the node was last
    relaxed at (tau0 + 3 nu_t), so from the second step on the closure's
    strain estimate S = -3 omega/(2 rho) Pi_neq takes that rate: two steps
    against the NumPy network and the oracle's collide with the step-1 rates
    (the network's, or the Smagorinsky rate where the inputs left the range)
    in the step-2 features.  With the base rate instead the two differ
    measurably."""
    import ctypes
    c, f, m, net = _closure_setup(9)
    say("closure strain from the last rate")
    with L.Solver(c) as s:
        s.set_populations(f)
        s.set_closure(m)
        s.step(2)
        g = s.populations()
    fs1 = _pre_collision(c, f)
    inr1, om1, _ = _reference_rates(c, fs1, net)
    # the Smagorinsky rate of the out-of-range nodes, from the oracle's tau(Q) law
    pl = orc_params(c)
    fr = fs1.reshape(19, -1)
    w = ctypes.c_float()
    last = om1.copy()
    for n in np.flatnonzero(~inr1):
        O.lib().orc_les_omega_f32(ctypes.byref(pl), np.ascontiguousarray(fr[:, n]), ctypes.byref(w), None)
        last[n] = w.value
    f1 = _reference_collide(c, fs1, inr1, om1)
    fs2 = _pre_collision(c, f1)
    inr2, om2, _ = _reference_rates(c, fs2, net, last)
    ref = _reference_collide(c, fs2, inr2, om2)
    inr_b, om_b, _ = _reference_rates(c, fs2, net)  # the base rate in the features
    ref_base = _reference_collide(c, fs2, inr_b, om_b)
    scale = np.abs(ref).max()
    err = np.abs(g - ref).max() / scale
    gap = np.abs(ref_base - ref).max() / scale
    say("  err", err, "base-rate reference differs by", gap)
    assert err < 1e-5 and gap > 5e-5


def test_closure_refresh_interval_matches_reference():
    """every = 3: the network runs on step 0 only; steps 1 and 2 collide at
    the rates of step 0 (Smagorinsky, recomputed each step, where step 0's
    inputs were out of range)."""
    c, f, m, net = _closure_setup(7)
    say("closure refresh interval")
    with L.Solver(c) as s:
        s.set_populations(f)
        s.set_closure(m, every=3)
        s.step(3)
        g = s.populations()
    with L.Solver(c) as s:
        s.set_populations(f)
        s.set_closure(m)
        s.step(3)
        g_every = s.populations()
    fs = _pre_collision(c, f)
    inr, omega, _ = _reference_rates(c, fs, net)
    ref = _reference_collide(c, fs, inr, omega)
    for _ in range(2):
        ref = _reference_collide(c, _pre_collision(c, ref), inr, omega)
    scale = np.abs(ref).max()
    err = np.abs(g - ref).max() / scale
    stale = np.abs(g_every - ref).max() / scale
    say("  err", err, "vs refreshing every step", stale)
    assert err < 1e-5
    assert stale > 10 * 1e-5


def test_closure_out_of_range_is_smagorinsky():
    c = _closure_case()
    say("closure fallback")
    rng = np.random.default_rng(6)
    f = random_state(rng, (c.nz, c.ny, c.nx), amp=2e-3, umax=0.03)
    m = L.Mlp(seed=1)
    m.set_norm(np.zeros(7), np.ones(7), np.full(7, 1e9, np.float32), np.full(7, 2e9, np.float32), 0.01)
    with L.Solver(c) as s:
        s.set_populations(f)
        s.set_closure(m)
        s.step(3)
        g = s.populations()
    with L.Solver(c) as s:
        s.set_populations(f)
        s.step(3)
        h = s.populations()
    assert np.abs(g - h).max() <= 1e-6 * np.abs(h).max()


def test_closure_rates_report_network_and_fallback():
    """lbm_get_closure_rates after one closure step: with every input inside
    the training range each node reports the network's rate, 1/(1/omega +
    3 nu_t) in (0, omega]; with every input outside it, minus the Smagorinsky
    rate, in [-omega, 0)."""
    c = _closure_case()
    say("closure rates")
    rng = np.random.default_rng(6)
    f = random_state(rng, (c.nz, c.ny, c.nx), amp=2e-3, umax=0.03)
    big = np.full(7, 3e38, np.float32)
    far = np.full(7, 1e9, np.float32)
    for lo, hi, sign in ((-big, big, 1.0), (far, 2 * far, -1.0)):
        m = L.Mlp(seed=1)
        m.set_norm(np.zeros(7), np.ones(7), lo, hi, 0.01)
        with L.Solver(c) as s:
            s.set_populations(f)
            s.set_closure(m)
            s.step(1)
            r = s.closure_rates()
        assert r.shape == (c.nz, c.ny, c.nx)
        w = sign * r
        assert (w > 0).all() and (w <= np.float32(c.omega) * (1 + 1e-6)).all(), (sign, w.min(), w.max())


@pytest.mark.parametrize("storage, bar", [("f16", 1e-3), ("bf16", 1e-2)])
def test_closure_on_16bit_storage_tracks_fp32(storage, bar):
    """The fused closure collide with 16-bit population storage against the
    same run in FP32 (pinned against the oracle above): the cylinder channel,
    20 steps with the network refreshed every step and every 3 steps, per-step
    forces recorded.  Velocity within the storage's precision bar (as in the
    storage precision tests); forces within 10 x that bar of max|F|, which
    also checks that the reduction reads exactly the partials the closure
    launch wrote (its workgroups differ from the 16-bit plain kernel's)."""
    say("closure on", storage)
    m = L.Mlp(seed=1)
    big = np.full(7, 3e38, np.float32)
    m.set_norm(np.zeros(7, np.float32), np.ones(7, np.float32), -big, big, 1e-3)
    for every in (1, 3):
        out = {}
        for st in ("f32", storage):
            cc = cfgmod.channel(nx=64, ny=32, nz=24, side=6, x_center=20, tau=0.52, storage=st, force_steps=20)
            with L.Solver(cc.to_c()) as s:
                s.set_geometry(None)
                s.init_equilibrium(L.INIT_BGK)
                s.set_closure(m, every=every)
                s.step(20)
                rho, u, v, w = s.moments()
                out[st] = (np.stack([u, v, w]), s.forces())
        (u32, F32), (u16, F16) = out["f32"], out[storage]
        du = np.sqrt(((u16 - u32) ** 2).sum() / (u32 ** 2).sum())
        dF = np.abs(F16 - F32).max() / np.abs(F32).max()
        say("  every", every, "velocity", du, "forces", dF, "max|F|", np.abs(F32).max())
        assert F32.shape == (20, 3) and np.abs(F32).max() > 0
        assert du < bar
        assert dF < 10 * bar


def test_closure_rejects_unsupported_contexts():
    m = L.Mlp()
    c = L.default_config()
    c.nx, c.ny, c.nz = 16, 8, 8
    c.model = L.MODEL_BGK
    with L.Solver(c) as s:
        with pytest.raises(L.LBMError):
            s.set_closure(m)
    with pytest.raises(L.LBMError):
        m.gradient(0, 10)                # no data set
    X = synthetic_inputs(0, 100)
    m.set_data(X, np.zeros(100, np.float32))
    with pytest.raises(L.LBMError):
        m.gradient(50, 60)               # range beyond the data


def _sampler_reference(f, mask, seed, keep, model=None):
    """Box filter by 2, central differences and the eddy-viscosity target,
    restated in FP64 from the populations and the node mask; model =
    (omega0, cs) adds the run's own filtered model stress to the target
    (mlp_ref.model_stress per fine node)."""
    Cv = CV
    solid = np.isin(mask, [L.NODE_SOLID, L.NODE_SOLID_FORCE])
    j = np.einsum("ia,i...->a...", Cv, f.astype(np.float64))
    j[:, solid] = 0
    nz, ny, nx = mask.shape
    cz, cy, cx = nz // 2, ny // 2, nx // 2

    def box(a):
        return a.reshape(cz, 2, cy, 2, cx, 2).mean(axis=(1, 3, 5))

    mb = None
    if model is not None:
        nu, Sf = M.model_stress(f.astype(np.float64), *model)
        m = nu[None, None] * Sf
        m[:, :, solid] = 0
        mb = np.array([[box(m[a, b]) for b in range(3)] for a in range(3)])

    ub = np.stack([box(j[a]) for a in range(3)])
    uu = {(a, b): box(j[a] * j[b]) for a in range(3) for b in range(a, 3)}
    sf = box(solid.astype(np.float64))
    dfine = M.wall_distance(solid, (False, True, True)).astype(np.float64)
    dc = box(dfine) * 0.5
    X, y, cells = [], [], []
    for k in range(1, cz - 1):
        for jj in range(1, cy - 1):
            for i in range(1, cx - 1):
                c = i + cx * (jj + cy * k)
                if M.hash_u32(seed, c) >= keep:
                    continue
                if sf[k - 1:k + 2, jj - 1:jj + 2, i - 1:i + 2].max() > 0:
                    continue
                g = np.empty((3, 3))
                for q in range(3):
                    g[q, 0] = 0.5 * (ub[q, k, jj, i + 1] - ub[q, k, jj, i - 1])
                    g[q, 1] = 0.5 * (ub[q, k, jj + 1, i] - ub[q, k, jj - 1, i])
                    g[q, 2] = 0.5 * (ub[q, k + 1, jj, i] - ub[q, k - 1, jj, i])
                S = 0.5 * (g + g.T)
                u = ub[:, k, jj, i]
                t = np.array([[uu[(min(a, b), max(a, b))][k, jj, i] - u[a] * u[b] for b in range(3)] for a in range(3)])
                t -= np.trace(t) / 3 * np.eye(3)
                ss = (S * S).sum()
                ms = (mb[:, :, k, jj, i] * S).sum() if mb is not None else 0.0
                nut = max(0.0, (2 * ms - (t * S).sum()) / (2 * ss)) if ss > 1e-20 else 0.0
                X.append(M.features(S, dc[k, jj, i], np.linalg.norm(u)))
                y.append(nut)
                cells.append(c)
    return np.array(X), np.array(y), np.array(cells)


def _les_state(cfg, shape, amp=0.02):
    """Post-collision populations of a smooth LES flow: u = U + a sin/cos
    modes, the non-equilibrium stress Pi = -(2 tau0/3) S of its strain rate
    (second-order Hermite part), collided node by node by the oracle's LES
    collide, so every node's stored stress carries the rate its collide
    applied (the sampler's model stress then matters: with random
    populations it does not)."""
    nz, ny, nx = shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    k = 2 * np.pi
    U = np.stack([cfg.umax + amp * np.sin(k * y / ny) * np.cos(k * z / nz),
                  amp * np.sin(k * x / nx) * np.cos(k * z / nz),
                  amp * np.cos(k * x / nx) * np.sin(k * y / ny)])
    g = np.stack([np.stack(np.gradient(U[a], axis=(2, 1, 0))) for a in range(3)])  # d u_a / d x_b
    S = 0.5 * (g + g.transpose(1, 0, 2, 3, 4))
    eye = np.eye(3)[:, :, None, None, None]
    Pi = -(2.0 / (3.0 * cfg.omega)) * S
    Pi = Pi - np.trace(Pi) / 3.0 * eye
    C = M.D3Q19_C
    w = np.array([12, 2, 2, 2, 2, 1, 1, 1, 1, 2, 1, 1, 1, 1, 2, 1, 1, 1, 1], np.float64) / 36.0
    cu = np.einsum("ia,a...->i...", C, U)
    feq = w[:, None, None, None] * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * (U * U).sum(0))
    H = np.einsum("ia,ib->iab", C, C) - np.eye(3) / 3.0
    f = (feq + w[:, None, None, None] * 4.5 * np.einsum("iab,ab...->i...", H, Pi)).astype(np.float32)
    p = orc_params(cfg.to_c())
    for idx in np.ndindex(nz, ny, nx):
        f[(slice(None),) + idx] = O.collide(p, np.ascontiguousarray(f[(slice(None),) + idx]))
    return f


@pytest.mark.parametrize("with_model", [False, True])
def test_sampler_matches_reference(with_model):
    cfg = cfgmod.channel(nx=32, ny=16, nz=16, side=4, x_center=10, tau=0.52, u0=0.05, seed=3)
    say("closure sampler, model stress" if with_model else "closure sampler")
    rng = np.random.default_rng(8)
    with L.Solver(cfg) as s:
        s.set_geometry()
        mask = s.get_mask()
        f = _les_state(cfg, mask.shape) if with_model else random_state(rng, mask.shape, amp=2e-3, umax=0.04)
        s.set_populations(f)
        s.set_closure_target(with_model)
        keep = 0.6
        X, y = s.closure_samples(seed=17, keep=keep, max_samples=4096)
    model = (float(np.float32(cfg.omega)), float(np.float32(cfg.cs))) if with_model else None
    Xr, yr, cells = _sampler_reference(f, mask, 17, int(keep * 2 ** 32), model)
    if with_model:
        _, y0, _ = _sampler_reference(f, mask, 17, int(keep * 2 ** 32))
        say("  model stress moves the target by", float(np.abs(yr - y0).max() / np.abs(y0).max()),
            "of its maximum; mean", float(y0.mean()), "->", float(yr.mean()))
        assert np.abs(yr - y0).max() > 0.05 * np.abs(y0).max()
    say("  samples", len(y), len(yr))
    assert len(y) == len(yr) and len(y) > 50
    assert np.abs(X[:, 0] - Xr[:, 0]).max() < 1e-4
    assert np.abs(X[:, 1:5] - Xr[:, 1:5]).max() < 2e-3
    assert np.abs(X[:, 5:] - Xr[:, 5:]).max() < 1e-5
    assert np.abs(y - yr).max() < 1e-3 * np.abs(yr).max() + 1e-9


def test_sampler_refuses_a_short_buffer_with_the_count():
    """More qualifying cells than max_samples: the call is refused with the
    number needed (the kept subset would depend on atomic order), and the
    binding's refetch at that size returns the same samples as a large buffer."""
    import ctypes
    cfg = cfgmod.channel(nx=32, ny=16, nz=16, side=4, x_center=10, tau=0.52, u0=0.05, seed=3)
    say("closure sampler, short buffer")
    rng = np.random.default_rng(8)
    with L.Solver(cfg) as s:
        s.set_geometry()
        s.set_populations(random_state(rng, s.get_mask().shape, amp=2e-3, umax=0.04))
        X, y = s.closure_samples(seed=17, keep=0.6, max_samples=4096)
        n = len(y)
        assert n > 50
        Xs = np.empty((n // 2, 7), np.float32)
        ys = np.empty(n // 2, np.float32)
        cnt = ctypes.c_int()
        rc = s.lib.lbm_closure_samples(s.h, 17, 0.6, Xs.reshape(-1), ys, n // 2, ctypes.byref(cnt))
        assert rc == -22 and cnt.value == n
        X2, y2 = s.closure_samples(seed=17, keep=0.6, max_samples=8)
    assert np.array_equal(X2, X) and np.array_equal(y2, y)


def test_model_stress_target_refused_near_tau_one():
    """The post-collision stress is (1 - omega) times the one that set a
    node's rate, so at |tau - 1| < 0.1 it no longer shows the rate and the
    model-stress target is refused; the plain target still samples there,
    and a flag other than 0 or 1 is rejected."""
    cfg = cfgmod.channel(nx=32, ny=16, nz=16, side=4, x_center=10, tau=1.05, u0=0.05, seed=3)
    say("closure target refusal at tau 1.05")
    with L.Solver(cfg) as s:
        s.set_geometry()
        s.init_equilibrium(L.INIT_BGK)
        assert s.lib.lbm_set_closure_target(s.h, 2) == -22
        s.set_closure_target(True)
        with pytest.raises(L.LBMError, match=r"\|tau - 1\|"):
            s.closure_samples(seed=1, keep=0.5, max_samples=64)
        s.set_closure_target(False)
        X, y = s.closure_samples(seed=1, keep=0.5, max_samples=4096)
    assert X.shape == (len(y), 7)


def test_trained_closure_beats_smagorinsky_apriori():
    """End to end at a quarter of the evaluation grids: a 128x32x32 fine run
    generates box-filtered samples, the network trains on the GPU, and on
    held-out snapshots its eddy viscosity fits the filtered-field target
    better than the best-fit Smagorinsky constant (and far better than the
    solver's Cs = 0.1); the learned coarse run stays close to the fine mean."""
    from lbm_multigpu_amd import closure as CL
    say("closure pipeline quarter scale")
    res = CL.run(scale=0.25, spinup_fine=4000, snapshots=40, every=50, keep=0.2, train_steps=1000,
                 resolutions=(), log=say)
    a = res["apriori"]
    say("  mse mlp", a["mse_mlp"], "smag best-fit", a["mse_smag_bestfit"], "var", a["var_target"])
    assert a["mse_mlp"] < 0.6 * a["mse_smag_bestfit"]
    assert a["mse_mlp"] < 0.6 * a["var_target"]
    assert a["mse_smag_cs0.1"] > 10 * a["mse_mlp"]
    r = res["runs"]
    assert r["coarse_learned"]["rel_l2"] < 0.1 and r["coarse_smagorinsky"]["rel_l2"] < 0.1
