"""Pins the NumPy closure reference (tests/mlp_ref.py) before the HIP
kernels are compared with it: finite differences for the backward pass,
closed forms for the inputs and the chamfer distance, the Adam first step,
and the counter hash against the oracle and the library's host export."""
import math

import numpy as np
import pytest

import mlp_ref as M


def _batch(seed, n=48):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, M.NIN)).astype(np.float32)
    y = rng.uniform(0.0, 2.0, n)
    return x, y


@pytest.mark.parametrize("nu_scale", [1.0, 3e-3])
def test_grad64_matches_finite_differences(nu_scale):
    p = M.glorot(1).astype(np.float64)
    x, y = _batch(2)
    y = y * nu_scale
    g, _ = M.grad64(p, x, y, nu_scale)
    rng = np.random.default_rng(3)
    for name, (o, s) in M.OFFSETS.items():
        idx = o + rng.choice(int(np.prod(s)), size=min(6, int(np.prod(s))), replace=False)
        for i in idx:
            h = 1e-6 * max(1.0, abs(p[i]))
            pp, pm = p.copy(), p.copy()
            pp[i] += h
            pm[i] -= h
            fd = (M.loss64(pp, x, y, nu_scale) - M.loss64(pm, x, y, nu_scale)) / (2 * h)
            scale = np.abs(g[o:o + int(np.prod(s))]).max()
            assert abs(fd - g[i]) <= 1e-6 * scale + 1e-12, (name, i, fd, g[i])


def test_pack_unpack_and_sizes():
    assert M.NPARAM == 8897
    p = np.arange(M.NPARAM, dtype=np.float64)
    assert np.array_equal(M.pack(M.unpack(p)), p)
    assert M.OFFSETS["b4"][0] == M.NPARAM - 1


def test_bf16_rounding_closed_forms():
    # exactly representable values stay, ties go to even, the 8-bit mantissa is kept
    v = np.array([1.0, -2.5, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, 3.0e-39], np.float32)
    r = M.bf16(v)
    assert r[0] == 1.0 and r[1] == -2.5
    assert r[2] == 1.0                       # tie -> even (mantissa 0)
    assert r[3] == np.float32(1.0 + 4 * 2 ** -8)  # tie -> even (mantissa 2)
    assert r[4] == 1.0
    x = np.random.default_rng(0).normal(size=1000).astype(np.float32)
    assert np.max(np.abs(M.bf16(x) - x) / np.abs(x)) <= 2 ** -8


def test_bf16_forward_tracks_fp64():
    p = M.glorot(4)
    x, y = _batch(5, 256)
    z64, _ = M.forward64(p, x)
    zb, _ = M.forward_bf16(p, x)
    assert np.linalg.norm(zb - z64) / np.linalg.norm(z64) < 2e-2
    g64, _ = M.grad64(p, x, y)
    gb, _ = M.grad_bf16(p, x, y)
    assert np.linalg.norm(gb - g64) / np.linalg.norm(g64) < 5e-2


def test_features_closed_forms():
    s = 1e-3
    axi = np.diag([2.0, -1.0, -1.0]) * s
    f = M.features(axi, 3.0, 0.04)
    assert f[1] == pytest.approx(1.0, abs=1e-12)
    assert f[0] == pytest.approx(math.log(math.sqrt(6.0) * s), rel=1e-9)
    assert f[5] == pytest.approx(math.log(3.0)) and f[6] == pytest.approx(0.04)
    assert M.features(-axi, 3.0, 0.0)[1] == pytest.approx(-1.0, abs=1e-12)
    shear = np.zeros((3, 3))
    shear[0, 1] = shear[1, 0] = s
    f = M.features(shear, 0.1, 0.0)
    assert f[1] == pytest.approx(0.0, abs=1e-12)
    assert f[2] == pytest.approx(1.0 / math.sqrt(2.0)) and f[3] == 0 and f[4] == 0
    assert f[0] == pytest.approx(math.log(math.sqrt(2.0) * s))
    assert f[5] == pytest.approx(math.log(0.5))  # distances below half a cell are clamped
    # the isotropic part does not enter
    rng = np.random.default_rng(1)
    a = rng.normal(size=(3, 3))
    a = a + a.T
    assert np.allclose(M.features(a, 2.0, 0.1), M.features(a + 0.7 * np.eye(3), 2.0, 0.1))
    # rotation invariance of the two invariants
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    fa, fr = M.features(a, 2.0, 0.1), M.features(q @ a @ q.T, 2.0, 0.1)
    assert np.allclose(fa[:2], fr[:2])


def test_wall_distance_chamfer_closed_form():
    n = 9
    solid = np.zeros((n, n, n), bool)
    solid[4, 4, 4] = True
    d = M.wall_distance(solid, (False, False, False))
    zz, yy, xx = np.meshgrid(range(n), range(n), range(n), indexing="ij")
    o = np.sort(np.abs(np.stack([xx - 4, yy - 4, zz - 4])), axis=0)
    c, b, a = o[0], o[1], o[2]
    want = (a - b) + (b - c) * math.sqrt(2.0) + c * math.sqrt(3.0)
    assert np.allclose(d, want, atol=1e-5)
    # walls only: the distance to the nearest face
    d = M.wall_distance(np.zeros((6, 5, 4), bool), (False, True, True))
    face = np.minimum(np.minimum(yy[:6, :5, :4] + 0.5, 4.5 - yy[:6, :5, :4]),
                      np.minimum(zz[:6, :5, :4] + 0.5, 5.5 - zz[:6, :5, :4]))
    assert np.allclose(d, face)


def test_hash_matches_oracle_and_library():
    import oracle_lib as O
    lib = O.lib()
    for seed, c in [(0, 0), (1, 2), (12345, 987654321), (2 ** 63 + 5, 2 ** 40 + 3)]:
        assert M.hash_u32(seed, c) == lib.orc_hash_u32(seed, c)
    import lbm_multigpu_amd as L
    L.load_library()
    for seed, c in [(7, 0), (99, 123456)]:
        assert M.hash_u32(seed, c) == L.load_library().lbm_hash_u32(seed, c)


def test_adam_first_step_is_sign_step():
    g = np.array([3.0, -1e-2, 0.5, -7.0])
    p = M.adam_reference(np.zeros(4), [g], lr=0.01)
    assert np.allclose(p, -0.01 * np.sign(g), rtol=1e-5)
