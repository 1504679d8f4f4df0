"""Pins the oracle's nested integrator with the gauge force on the fine scale (no
reference counterpart) by what fixes it: at beta = 0 every inner integrator is one
drift, so the trajectory is the plain one of the fermion force alone; at phi = 0 the
fermion force vanishes, so the trajectory is the plain one of the gauge force with the
inner step; the plain trajectories are pinned to the reference's golden vectors.  Also
time reversibility and the second-order scaling of the energy violation."""
import numpy as np
import pytest

import oracle as O
from schwinger_mg_amd.fields import gaussian_vec, random_gauge

L, MASS, R, CSW, MU = 8, 0.1, 1.0, 1.0, 0.0


def _setup(seed):
    U = random_gauge(L, 0.4, seed)
    pi = np.random.default_rng(seed + 1).normal(size=2 * L * L)
    phi = gaussian_vec(2 * L * L, seed + 2)
    return U, pi, phi


@pytest.mark.parametrize("om_out,om_in", [(0, 0), (1, 0), (1, 1), (0, 1)])
def test_nested_beta0_is_plain_fermion_trajectory(om_out, om_in):
    U, pi, phi = _setup(5)
    Un, pn, _ = O.hmc_nested_evolve(L, U, MASS, R, CSW, MU, 0.0, 0.6, 3, 4, om_out, om_in, 500, 1e-13, pi, phi)
    Up, pp, _ = O.hmc_evolve(L, U, MASS, R, CSW, MU, 0.0, 0.6, 3, om_out, 500, 1e-13, pi, phi)
    assert np.abs(Un - Up).max() < 1e-12 and np.abs(pn - pp).max() < 1e-11


@pytest.mark.parametrize("om_out,om_in", [(0, 0), (1, 0), (1, 1), (0, 1)])
def test_nested_phi0_is_plain_gauge_trajectory(om_out, om_in):
    U, pi, _ = _setup(7)
    phi = np.zeros(2 * L * L, np.complex128)
    n_outer, n_gauge = 2, 3
    Un, pn, _ = O.hmc_nested_evolve(L, U, MASS, R, CSW, MU, 2.0, 0.8, n_outer, n_gauge, om_out, om_in, 500, 1e-13,
                                    pi, phi)
    n_plain = n_outer * n_gauge * (2 if om_out else 1)
    Up, pp, _ = O.hmc_evolve(L, U, MASS, R, CSW, MU, 2.0, 0.8, n_plain, om_in, 500, 1e-13, pi, phi)
    assert np.abs(Un - Up).max() < 1e-12 and np.abs(pn - pp).max() < 1e-11


def test_nested_reversible_and_second_order():
    U, pi, phi = _setup(9)
    beta, tau = 2.0, 0.5
    U1, p1, _ = O.hmc_nested_evolve(L, U, MASS, R, CSW, MU, beta, tau, 2, 4, 1, 1, 1000, 1e-14, pi, phi)
    U2, p2, _ = O.hmc_nested_evolve(L, U1, MASS, R, CSW, MU, beta, tau, 2, 4, 1, 1, 1000, 1e-14, -p1, phi)
    assert np.abs(U2 - U).max() < 1e-10 and np.abs(-p2 - pi).max() < 1e-9

    tau = 0.1

    def dH(n_outer):
        h0, h1 = np.zeros(4), np.zeros(4)
        O.lib().orc_hmc_hamiltonian(L, O._ptr(U), MASS, R, CSW, MU, beta, 1000, 1e-14, O._ptr(pi), O._ptr(phi),
                                    O._ptr(h0))
        Ue, pe, _ = O.hmc_nested_evolve(L, U, MASS, R, CSW, MU, beta, tau, n_outer, 4, 0, 0, 1000, 1e-14, pi, phi)
        O.lib().orc_hmc_hamiltonian(L, O._ptr(Ue), MASS, R, CSW, MU, beta, 1000, 1e-14, O._ptr(pe), O._ptr(phi),
                                    O._ptr(h1))
        return h1[:3].sum() - h0[:3].sum()
    d1, d2, d4 = dH(4), dH(8), dH(16)
    # leapfrog: the energy violation falls as h^2 (ratios near 4)
    assert 3.8 < d1 / d2 < 4.2 and 3.8 < d2 / d4 < 4.2, (d1, d2, d4)
