"""The closure sampler's consistent target: an LES node's Smagorinsky
viscosity and strain rate recovered from its stored post-collision
populations (mlp_ref.model_stress), pinned against the oracle's own collide.

A pre-collision node f = f_eq(rho, j) + f_neq with a chosen traceless
non-equilibrium stress Pi (second-order Hermite part, so sum c c f_neq = Pi,
zero mass and momentum) is collided by the oracle (MRT with the LES rate,
both rate presets, FP64).  From the post-collision populations alone the
inversion must return the rate the collide applied (nu = (tau - tau0)/3 with
tau from the oracle's own LES rate of the pre-collision node) and the strain
rate that shaped Pi (S = -3 Pi / (2 rho tau)).  A sign slip in the (tau - 1)
factor, or the base rate in place of the node's own, fails it.
"""
import ctypes

import numpy as np
import pytest

import mlp_ref as M
import oracle_lib as O

W = np.array([12, 2, 2, 2, 2, 1, 1, 1, 1, 2, 1, 1, 1, 1, 2, 1, 1, 1, 1], np.float64) / 36.0


def _pre_collision(rng, rho, j, pi):
    C = M.D3Q19_C
    feq = O.feq_bgk(rho, *j, dtype=np.float64)
    H = np.einsum("ia,ib->iab", C, C) - np.eye(3) / 3.0
    fneq = W * 4.5 * np.einsum("iab,ab->i", H, pi)
    return feq + fneq


@pytest.mark.parametrize("collision", [O.MRT_LES, O.MRT_DHUM])
def test_model_stress_recovers_the_applied_rate_and_strain(collision):
    rng = np.random.default_rng(11)
    omega0, cs = 1.0 / 0.52, 0.01
    p = O.params(collision=collision, omega=omega0, cs=cs)
    if collision == O.MRT_DHUM:
        for k, v in {1: 1.19, 2: 1.4, 4: 1.2, 10: 1.4, 16: 1.98}.items():
            p.s_dhum[k] = v
    C = M.D3Q19_C
    for trial in range(40):
        a = rng.normal(size=(3, 3))
        pi = 0.5 * (a + a.T)
        pi -= np.trace(pi) / 3.0 * np.eye(3)
        pi *= 10 ** rng.uniform(-5, -2) / np.sqrt((pi * pi).sum())
        rho = 1.0 + rng.uniform(-0.02, 0.02)
        j = rng.uniform(-0.06, 0.06, 3)
        f = _pre_collision(rng, rho, j, pi)
        assert np.allclose(np.einsum("ia,ib,i->ab", C, C, f - O.feq_bgk(rho, *j, dtype=np.float64)), pi, atol=1e-15)
        om = ctypes.c_double()
        O.lib().orc_les_omega_f64(ctypes.byref(p), f, ctypes.byref(om), None)
        tau = 1.0 / om.value
        g = O.collide(p, f)
        nu, S = M.model_stress(g, omega0, cs, iters=30)
        nu_exp = (tau - 1.0 / omega0) / 3.0
        S_exp = -3.0 * pi / (2.0 * rho * tau)
        assert nu_exp > 0
        assert abs(nu - nu_exp) <= 1e-9 * nu_exp + 1e-15, (trial, nu, nu_exp)
        assert np.abs(S - S_exp).max() <= 1e-9 * np.abs(S_exp).max(), (trial, S, S_exp)


def test_model_stress_fixed_point_converges_in_the_sampler_steps():
    """The sampler's 8 fixed-point steps reach the converged rate to 1e-6 over
    the range of stresses the channel runs show (|Pi| up to 1e-2)."""
    omega0, cs = 1.0 / 0.505, 0.01
    rng = np.random.default_rng(3)
    for q in np.logspace(-6, -2, 13):
        pi = np.diag([1.0, -1.0, 0.0]) * q / np.sqrt(2.0)
        f = _pre_collision(rng, 1.0, np.zeros(3), pi)[:, None]
        nu8, _ = M.model_stress(f, omega0, cs, iters=8)
        nu60, _ = M.model_stress(f, omega0, cs, iters=60)
        assert abs(nu8[0] - nu60[0]) <= 1e-6 * nu60[0] + 1e-14
