"""NumPy reference of the learned eddy-viscosity closure (test infrastructure).

The closure has no counterpart in the reference solver (its LES closure is
constant-Cs Smagorinsky, MRTtest/DynInlet3.cu:628-735), so this module is
the CPU yardstick the HIP kernels are checked against:

* ``forward64`` / ``grad64``: the network in plain FP64 (no rounding), whose
  backward pass is pinned by central finite differences in
  tests/test_mlp_ref.py;
* ``forward_bf16`` / ``grad_bf16``: the same network with the BF16 rounding
  points of the HIP kernels (inputs, weights, pre-activations, activations and
  back-propagated deltas rounded to BF16; sums in FP64, standing in for the
  matrix cores' FP32 accumulation);
* ``features``: the seven closure inputs from a strain-rate tensor, wall
  distance and speed, checked against closed forms in the same test file;
* ``wall_distance``: the two-pass 3-D chamfer distance the library uses;
* ``model_stress``: an LES node's Smagorinsky viscosity and strain rate
  recovered from its stored post-collision populations (the sampler's
  consistent target), pinned in tests/test_closure_target.py against the
  oracle's own collide.

Parameter layout (flat FP32, 8897): W1[7][64] b1[64] W2[64][64] b2[64]
W3[64][64] b3[64] w4[64] b4.
"""
import math

import numpy as np

NIN, NH = 7, 64
SIZES = [("W1", (NIN, NH)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)),
         ("W3", (NH, NH)), ("b3", (NH,)), ("w4", (NH,)), ("b4", (1,))]
NPARAM = sum(int(np.prod(s)) for _, s in SIZES)
OFFSETS = {}
_o = 0
for _n, _s in SIZES:
    OFFSETS[_n] = (_o, _s)
    _o += int(np.prod(_s))


def unpack(p):
    return {n: np.asarray(p[o:o + int(np.prod(s))]).reshape(s) for n, (o, s) in OFFSETS.items()}


def pack(d):
    return np.concatenate([np.asarray(d[n], np.float64).reshape(-1) for n, _ in SIZES])


def glorot(seed):
    """A random parameter vector (Glorot-uniform weights, small biases)."""
    rng = np.random.default_rng(seed)
    d = {}
    for n, s in SIZES:
        if n.startswith("W") or n == "w4":
            fin, fout = (s[0], s[1]) if len(s) == 2 else (s[0], 1)
            a = math.sqrt(6.0 / (fin + fout))
            d[n] = rng.uniform(-a, a, s)
        else:
            d[n] = rng.uniform(-0.1, 0.1, s)
    return pack(d).astype(np.float32)


def bf16(x):
    """Round float32 values to BF16 (nearest even), returned as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


_C = math.sqrt(2.0 / math.pi)


def gelu(x):
    """GELU, tanh form: 0.5 x (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3)))."""
    return 0.5 * x * (1.0 + np.tanh(_C * (x + 0.044715 * x ** 3)))


def gelu_grad(x):
    t = np.tanh(_C * (x + 0.044715 * x ** 3))
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * _C * (1.0 + 3 * 0.044715 * x * x)


def softplus(z):
    return np.maximum(z, 0.0) + np.log1p(np.exp(-np.abs(z)))


def sigmoid(z):
    return 1.0 / (1.0 + np.exp(-z))


def normalise(X, mu, sd):
    """(x - mu) * (1/sd) in FP32, as the kernels apply it."""
    inv = (np.float32(1.0) / np.asarray(sd, np.float32)).astype(np.float32)
    return ((np.asarray(X, np.float32) - np.asarray(mu, np.float32)) * inv).astype(np.float32)


# ------------------------------------------------------------------ FP64

def forward64(p, xn):
    """Exact forward pass on normalised inputs xn[n, 7]; returns (zo, cache)."""
    P = unpack(np.asarray(p, np.float64))
    x = np.asarray(xn, np.float64)
    z1 = x @ P["W1"] + P["b1"]
    h1 = gelu(z1)
    z2 = h1 @ P["W2"] + P["b2"]
    h2 = gelu(z2)
    z3 = h2 @ P["W3"] + P["b3"]
    h3 = gelu(z3)
    zo = h3 @ P["w4"] + P["b4"][0]
    return zo, (x, z1, h1, z2, h2, z3, h3)


def loss64(p, xn, y, nu_scale=1.0):
    """mean((softplus(zo) - y/nu_scale)^2): the squared error of nu_t in
    units of nu_scale."""
    zo, _ = forward64(p, xn)
    r = softplus(zo) - np.asarray(y, np.float64) / nu_scale
    return float(np.mean(r * r))


def grad64(p, xn, y, nu_scale=1.0):
    """Gradient of loss64."""
    P = unpack(np.asarray(p, np.float64))
    zo, (x, z1, h1, z2, h2, z3, h3) = forward64(p, xn)
    n = x.shape[0]
    r = softplus(zo) - np.asarray(y, np.float64) / nu_scale
    g = 2.0 * r * sigmoid(zo) / n
    d = {"w4": h3.T @ g, "b4": np.array([g.sum()])}
    dz3 = np.outer(g, P["w4"]) * gelu_grad(z3)
    d["W3"], d["b3"] = h2.T @ dz3, dz3.sum(0)
    dz2 = (dz3 @ P["W3"].T) * gelu_grad(z2)
    d["W2"], d["b2"] = h1.T @ dz2, dz2.sum(0)
    dz1 = (dz2 @ P["W2"].T) * gelu_grad(z1)
    d["W1"], d["b1"] = x.T @ dz1, dz1.sum(0)
    return pack(d), float(np.mean(r * r))


# ------------------------------------------------------------------ BF16

def _f32(a):
    return np.asarray(a, np.float64).astype(np.float32)


def forward_bf16(p, xn):
    """Forward pass with the kernels' BF16 rounding points; xn FP32 [n, 7]."""
    P = unpack(np.asarray(p, np.float32))
    xb = bf16(xn).astype(np.float64)
    W1, W2, W3 = (bf16(P[k]).astype(np.float64) for k in ("W1", "W2", "W3"))
    z1 = bf16(_f32(xb @ W1) + P["b1"])
    h1 = bf16(_f32(gelu(z1.astype(np.float64))))
    z2 = bf16(_f32(h1.astype(np.float64) @ W2) + P["b2"])
    h2 = bf16(_f32(gelu(z2.astype(np.float64))))
    z3 = bf16(_f32(h2.astype(np.float64) @ W3) + P["b3"])
    h3 = bf16(_f32(gelu(z3.astype(np.float64))))
    zo = _f32(h3.astype(np.float64) @ P["w4"].astype(np.float64) + float(P["b4"][0]))
    return zo, (xb, z1, h1, z2, h2, z3, h3, (W1, W2, W3), P)


def grad_bf16(p, xn, y, nu_scale=1.0):
    """Backward pass with the kernels' BF16 rounding points."""
    zo, (xb, z1, h1, z2, h2, z3, h3, (W1, W2, W3), P) = forward_bf16(p, xn)
    n = xb.shape[0]
    zo64 = zo.astype(np.float64)
    r = softplus(zo64) - (np.asarray(y, np.float32) / np.float32(nu_scale)).astype(np.float64)
    g = _f32(2.0 * r * sigmoid(zo64) / n).astype(np.float64)
    h3d, h2d, h1d = (h.astype(np.float64) for h in (h3, h2, h1))
    d = {"w4": h3d.T @ g, "b4": np.array([g.sum()])}
    dz3 = bf16(_f32(np.outer(g, P["w4"].astype(np.float64)) * gelu_grad(z3.astype(np.float64)))).astype(np.float64)
    d["W3"], d["b3"] = h2d.T @ dz3, dz3.sum(0)
    dz2 = bf16(_f32((dz3 @ W3.T) * gelu_grad(z2.astype(np.float64)))).astype(np.float64)
    d["W2"], d["b2"] = h1d.T @ dz2, dz2.sum(0)
    dz1 = bf16(_f32((dz2 @ W2.T) * gelu_grad(z1.astype(np.float64)))).astype(np.float64)
    d["W1"], d["b1"] = xb.T @ dz1, dz1.sum(0)
    return pack(d), float(np.mean(r * r))


# ------------------------------------------------------------------ inputs

def features(S, dwall, speed):
    """Closure inputs from strain-rate tensors S[..., 3, 3] (FP64):
    log|S_dev|, 3 sqrt(6) det(S_dev/|S_dev|) clipped to [-1, 1], the three
    normalised off-diagonal components, log max(d_wall, 1/2), |u|."""
    S = np.asarray(S, np.float64)
    Sd = S - np.trace(S, axis1=-2, axis2=-1)[..., None, None] * np.eye(3) / 3.0
    n = np.sqrt(np.maximum(np.sum(Sd * Sd, axis=(-2, -1)), 1e-30))
    Sh = Sd / n[..., None, None]
    det = np.linalg.det(Sh)
    return np.stack([np.log(n + 1e-12), np.clip(3.0 * math.sqrt(6.0) * det, -1.0, 1.0),
                     Sh[..., 0, 1], Sh[..., 0, 2], Sh[..., 1, 2],
                     np.log(np.maximum(dwall, 0.5)), np.asarray(speed, np.float64)], axis=-1)


def wall_distance(solid, wall_axes):
    """Chamfer distance (steps 1, sqrt2, sqrt3) to the nearest solid node or
    wall face; solid[z, y, x] bool, wall_axes = (x, y, z) booleans for walls
    half a cell outside the domain.  Pure Python: small domains only."""
    nz, ny, nx = solid.shape
    d = np.full(solid.shape, 1e9, np.float32)
    d[solid] = 0
    zz, yy, xx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    for ax, (c, n) in enumerate(((xx, nx), (yy, ny), (zz, nz))):
        if wall_axes[ax]:
            d = np.minimum(d, np.minimum(c + 0.5, n - 0.5 - c).astype(np.float32))
    w = [np.float32(0), np.float32(1.0), np.float32(1.41421356), np.float32(1.73205081)]
    nb = [(dx, dy, dz) for dz in (-1, 0) for dy in (-1, 0, 1) for dx in (-1, 0, 1)
          if not (dz == 0 and (dy > 0 or (dy == 0 and dx >= 0)))]
    for s in (1, -1):
        xs = range(nx) if s == 1 else range(nx - 1, -1, -1)
        ys = range(ny) if s == 1 else range(ny - 1, -1, -1)
        zs = range(nz) if s == 1 else range(nz - 1, -1, -1)
        for z in zs:
            for y in ys:
                for x in xs:
                    v = d[z, y, x]
                    for dx, dy, dz in nb:
                        X, Y, Z = x + s * dx, y + s * dy, z + s * dz
                        if 0 <= X < nx and 0 <= Y < ny and 0 <= Z < nz:
                            v = min(v, np.float32(d[Z, Y, X] + w[abs(dx) + abs(dy) + abs(dz)]))
                    d[z, y, x] = v
    return d


D3Q19_C = np.array([(0, 0, 0), (1, 0, 0), (0, 1, 0), (-1, 0, 0), (0, -1, 0), (1, 1, 0), (-1, 1, 0), (-1, -1, 0),
                    (1, -1, 0), (0, 0, 1), (1, 0, 1), (0, 1, 1), (-1, 0, 1), (0, -1, 1), (0, 0, -1), (1, 0, -1),
                    (0, 1, -1), (-1, 0, -1), (0, -1, -1)], np.float64)


def model_stress(f, omega0, cs, iters=8):
    """(nu_f, S_f) of LES nodes from their post-collision populations
    f[19, ...] (FP64): Pi' = P - rho/3 I - j j is (1 - omega) times the
    pre-collision non-equilibrium stress Pi that set the node's rate through
    tau = tau0/2 + sqrt(tau0^2 + 18 sqrt2 cs |Pi|)/2, so tau solves
    tau = tau0/2 + sqrt(tau0^2 + 18 sqrt2 cs |Pi'_dev| tau/|tau - 1|)/2
    (fixed point from tau0; |tau - 1| floored at 0.05; the trace of Pi',
    relaxed at another rate, left out); nu_f = (tau - tau0)/3 and
    S_f = -3 Pi'_dev / (2 rho (tau - 1))."""
    f = np.asarray(f, np.float64)
    rho = f.sum(0)
    j = np.einsum("ia,i...->a...", D3Q19_C, f)
    P = np.einsum("ia,ib,i...->ab...", D3Q19_C, D3Q19_C, f)
    I = np.eye(3).reshape((3, 3) + (1,) * (f.ndim - 1))
    Pi = P - rho / 3.0 * I - j[:, None] * j[None, :]
    Pi = Pi - np.trace(Pi) / 3.0 * I
    q = np.sqrt((Pi * Pi).sum(axis=(0, 1)))
    tau0 = 1.0 / omega0
    k = 18.0 * math.sqrt(2.0) * cs * q
    tau = np.full_like(q, tau0)
    for _ in range(iters):
        tau = 0.5 * tau0 + 0.5 * np.sqrt(tau0 * tau0 + k * tau / np.maximum(np.abs(tau - 1.0), 0.05))
    nu = (tau - tau0) / 3.0
    S = -1.5 * Pi / (rho * (tau - 1.0))
    return nu, S


def hash_u32(seed, counter):
    """The library's counter hash (splitmix64 finaliser over seed ^ counter * golden)."""
    M = (1 << 64) - 1
    x = (seed ^ ((counter * 0x9E3779B97F4A7C15) & M)) & M
    x = (x + 0x632BE59BD9B4E019) & M
    x ^= x >> 30
    x = (x * 0xBF58476D1CE4E5B9) & M
    x ^= x >> 27
    x = (x * 0x94D049BB133111EB) & M
    x ^= x >> 31
    return x >> 32


def adam_reference(p, grads, lr, b1=0.9, b2=0.999, eps=1e-8):
    """Adam in FP64 over a sequence of gradients."""
    p = np.asarray(p, np.float64).copy()
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    for t, g in enumerate(grads, 1):
        m = b1 * m + (1 - b1) * g
        v = b2 * v + (1 - b2) * g * g
        p -= lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return p
