"""Host-side pieces of the closure pipeline (lbm_multigpu_amd/closure.py):
resampling between grids, the evaluation region, the a-priori Smagorinsky
model and the error norms, against closed forms."""
import math

import numpy as np
import pytest

import mlp_ref as M
from lbm_multigpu_amd import closure as CL


def test_resampling_is_box_filter_and_exact_on_linear_fields():
    rng = np.random.default_rng(0)
    f = rng.normal(size=(8, 6, 10))
    box = f.reshape(4, 2, 3, 2, 5, 2).mean(axis=(1, 3, 5))
    assert np.allclose(CL.at_coarse_centres(f, 1.0, 0.5), box)
    # a linear field of position (fine units) is reproduced at the coarse
    # node positions from a 1.25x grid (interior nodes; edges clamp)
    r = 0.625
    n = (20, 15, 25)
    z, y, x = np.meshgrid(*((np.arange(k) + 0.5) / r for k in n), indexing="ij")
    g = 0.3 * x - 0.2 * y + 0.1 * z
    out = CL.at_coarse_centres(g, r, 0.5)
    zc, yc, xc = np.meshgrid(*((np.arange(k) + 0.5) / 0.5 for k in out.shape), indexing="ij")
    want = 0.3 * xc - 0.2 * yc + 0.1 * zc
    assert np.allclose(out[1:-1, 1:-1, 1:-1], want[1:-1, 1:-1, 1:-1])


def test_smagorinsky_apriori_model():
    S = np.diag([1.0, -0.5, -0.5]) * 2e-3
    S[0, 1] = S[1, 0] = 1e-3
    X = M.features(S, 4.0, 0.05)[None, :].astype(np.float32)
    want = 0.01 * math.sqrt(2.0 * np.sum(S * S))
    assert CL.smagorinsky_nu(X, 0.01)[0] == pytest.approx(want, rel=1e-5)


def test_eval_region_and_errors():
    mask = np.zeros((6, 6, 12), np.uint8)
    mask[3, 3, 5] = 1
    reg = CL.eval_region(mask)
    assert not reg[3, 3, 5] and not reg[3, 3, 7] and reg[3, 3, 8] and not reg[1, 3, 5] and reg[0, 0, 0]
    u_ref = np.full(mask.shape, 0.05) - 0.01 * (np.arange(12) == 9)
    u = u_ref + 0.001
    e = CL.l2_errors(u, u_ref, reg, 0.05)
    assert e["rel_l2"] == pytest.approx(np.linalg.norm(np.full(reg.sum(), 0.001)) / np.linalg.norm(u_ref[reg]))
    assert e["wake_l2"] > e["rel_l2"]


def test_channel_family_keeps_physical_set_up():
    for r in (0.5, 0.625, 0.75, 1.0):
        c = CL.channel_at(r)
        assert (c.nx, c.ny, c.nz) == (round(512 * r), round(128 * r), round(128 * r))
        assert c.obst1_r * 2 == round(16 * r)
        nu = (1 / c.omega - 0.5) / 3
        assert nu == pytest.approx(0.02 / 3 * r, rel=1e-9)
