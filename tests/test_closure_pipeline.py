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


def test_time_for_error_on_a_resolution_curve():
    """Cheapest GPU time at which the curve reaches a target error: a crossing
    from above is interpolated log-log (exact for a power law), a cheapest
    point already at or below the target is returned as is, and an unreached
    target gives None; a non-monotone curve uses its cheapest crossing from
    above, not a later rise through the target."""
    tms = [1.0, 2.0, 4.0, 8.0]
    errs = [0.08 * t ** -0.5 for t in tms]                      # power law e = 0.08 t^-1/2
    t, cheapest = CL.time_for_error(tms, errs, 0.08 * 3.0 ** -0.5)
    assert not cheapest and abs(t - 3.0) < 1e-12
    assert CL.time_for_error(tms, errs, 0.09) == (1.0, True)
    assert CL.time_for_error(tms, errs, 0.01) == (None, False)
    # the cheapest point already meets the target; a later bump above it
    # and back down is not a crossing that counts
    t, cheapest = CL.time_for_error([1.0, 1.5, 2.4], [0.19, 0.23, 0.14], 0.228)
    assert cheapest and t == 1.0
    t, cheapest = CL.time_for_error([1.0, 1.5, 2.4], [0.24, 0.26, 0.14], 0.2)
    assert not cheapest and 1.5 < t < 2.4
    assert CL.time_for_error([1.0, 2.0], [float("nan"), 0.1], 0.2) == (None, False)
