"""The learned-closure workload at its full size: a 512x128x128 fine
channel generates the box-filtered samples, the network trains on the GPU,
and 256x64x64 coarse runs with the learned closure, Smagorinsky Cs = 0.1 and
no sub-grid model are compared with the fine run's mean velocity.

At fine tau 0.505 (coarse tau 0.5025) the coarse grid needs a model: without
one the run diverges.  What the study shows there, and what this test holds
it to: the learned run stays stable and within 10 % relative L2 of the fine
mean, and within 35 % of the Smagorinsky run's error (the full study: 7.3 %
against 6.4 %).  It does not beat
Smagorinsky: the box-filtered eddy viscosity it learns (mean 2.5e-5 in
lattice units, against 3e-4 for Cs = 0.1) adds little dissipation, and an a
posteriori scale on it does not close the gap (the full study with
calibration: python -m lbm_multigpu_amd.closure --tau 0.505 --calibrate).
Fewer snapshots and training steps than the full study keep this to about
a minute on one GPU.
"""
import math

import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def test_config5_learned_closure_a_posteriori():
    from lbm_multigpu_amd import closure as CL
    say("closure workload 256x64x64 vs 512x128x128, fine tau 0.505")
    res = CL.run(scale=1.0, spinup_fine=16000, snapshots=160, every=100, keep=0.02, train_steps=1500,
                 resolutions=(), refresh=(1,), tau_fine=0.505, log=say)
    assert res["fine"]["grid"] == [512, 128, 128]
    r = res["runs"]
    assert r["coarse_learned"]["grid"] == [256, 64, 64]
    learned, smag, none = (r[k]["rel_l2"] for k in ("coarse_learned", "coarse_smagorinsky", "coarse_no_model"))
    say("  rel L2 learned", learned, "smagorinsky", smag, "no model", none)
    assert math.isfinite(learned) and learned < 0.10
    # no model: diverged, or no better than the learned closure
    assert not math.isfinite(none) or none > learned
    assert learned < 1.35 * smag
