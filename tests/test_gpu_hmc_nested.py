"""The nested integrator with the gauge force on the fine scale on the device
(smg_hmc_nested_evolve / smg_hmc_nested_trajectory) against the oracle's, with FP64
CG, the MG-preconditioned CG and the mixed-precision MG solver."""
import numpy as np
import pytest

import oracle as O
from helpers import banner
from schwinger_mg_amd.fields import gaussian_vec, gaussian_vecs, random_gauge

pytestmark = pytest.mark.gpu
MASS, R, CSW, MU = 0.1, 1.0, 1.0, 0.0


def _case(L, seed):
    U = random_gauge(L, 0.3, seed)
    pi = np.random.default_rng(seed + 1).normal(size=2 * L * L)
    phi = gaussian_vec(2 * L * L, seed + 2)
    return U, pi, phi


def _mg(smg, D, L, seed):
    starts = gaussian_vecs(4, 2 * L * L, seed).ravel()
    return smg.MGHierarchy(D, starts, n_levels=2, blocks=((4, 4),), k=(4,), null_iters=10, coarse_maxit=500,
                           coarse_tol=1e-12)


@pytest.mark.parametrize("om_out,om_in", [(1, 1), (0, 0), (1, 0)])
def test_nested_evolve_vs_oracle(smg, om_out, om_in):
    banner(f"nested evolve L=16 omelyan outer={om_out} inner={om_in}")
    L = 16
    U, pi, phi = _case(L, 30 + 2 * om_out + om_in)
    D = smg.DiracOp(L, U, MASS, R, CSW, MU)
    p = smg.nested_params(2.0, 0.4, 2, 4, om_out, om_in, smg.F64, 2000, 1e-13)
    pg, it = smg.hmc_nested_evolve(D, p, pi, phi)
    Uo, po, ito = O.hmc_nested_evolve(L, U, MASS, R, CSW, MU, 2.0, 0.4, 2, 4, om_out, om_in, 2000, 1e-13, pi, phi)
    Ug = smg.gauge_of(D)
    assert np.abs(Ug - Uo).max() < 1e-10 and np.abs(pg - po).max() < 1e-9
    assert abs(it - ito) <= 5   # at most one iteration per solve


@pytest.mark.parametrize("solver", ["mg64", "mixed"])
def test_nested_evolve_mg_solvers_vs_oracle(smg, solver):
    banner(f"nested evolve L=32 solver={solver}")
    L = 32
    U, pi, phi = _case(L, 41)
    D = smg.DiracOp(L, U, MASS, R, CSW, MU)
    mg = _mg(smg, D, L, 43)
    p = smg.nested_params(2.0, 0.3, 1, 8, True, True, smg.F64 if solver == "mg64" else smg.F32, 500, 1e-12)
    pg, it = smg.hmc_nested_evolve(D, p, pi, phi, mg=mg)
    Uo, po, ito = O.hmc_nested_evolve(L, U, MASS, R, CSW, MU, 2.0, 0.3, 1, 8, 1, 1, 5000, 1e-13, pi, phi)
    # the solves agree to their 1e-12 tolerance, amplified by the condition number
    assert np.abs(smg.gauge_of(D) - Uo).max() < 1e-8 and np.abs(pg - po).max() < 1e-7
    assert it < ito   # the MG solves take fewer iterations than the oracle's plain CG


def test_nested_trajectory_dH_and_reversibility(smg):
    banner("nested trajectory L=16 dH + reversibility")
    L = 16
    beta, tau = 2.0, 0.5
    U, pi, _ = _case(L, 50)
    eta = gaussian_vec(2 * L * L, 53)
    D = smg.DiracOp(L, U, MASS, R, CSW, MU)
    p = smg.nested_params(beta, tau, 2, 6, True, True, smg.F64, 2000, 1e-13, force_accept=True)
    res = smg.hmc_trajectory_nested(D, p, pi, eta, 0.5)
    # the oracle's: phi = D^dag eta, H before and after its own trajectory
    Do = O.Dirac(L, U, MASS, R, CSW, MU)
    phi = Do.apply(1, eta)
    h0, h1 = np.zeros(4), np.zeros(4)
    O.lib().orc_hmc_hamiltonian(L, O._ptr(U), MASS, R, CSW, MU, beta, 2000, 1e-13, O._ptr(pi), O._ptr(phi),
                                O._ptr(h0))
    Uo, po, _ = O.hmc_nested_evolve(L, U, MASS, R, CSW, MU, beta, tau, 2, 6, 1, 1, 2000, 1e-13, pi, phi)
    O.lib().orc_hmc_hamiltonian(L, O._ptr(Uo), MASS, R, CSW, MU, beta, 2000, 1e-13, O._ptr(po), O._ptr(phi),
                                O._ptr(h1))
    dHo = h1[:3].sum() - h0[:3].sum()
    assert res.accepted == 1 and abs(res.dH - dHo) < 1e-8 * max(1.0, abs(h0[:3]).max()), (res.dH, dHo)
    assert np.abs(smg.gauge_of(D) - Uo).max() < 1e-10
    # reversibility on the device: evolve back with the flipped momenta
    D2 = smg.DiracOp(L, U, MASS, R, CSW, MU)
    p1, _ = smg.hmc_nested_evolve(D2, p, pi, phi)
    p2, _ = smg.hmc_nested_evolve(D2, p, -p1, phi)
    assert np.abs(smg.gauge_of(D2) - U).max() < 1e-10 and np.abs(-p2 - pi).max() < 1e-9


def test_nested_reject_restores_gauge(smg):
    banner("nested trajectory reject L=16")
    L = 16
    U, pi, _ = _case(L, 60)
    eta = gaussian_vec(2 * L * L, 61)
    D = smg.DiracOp(L, U, MASS, R, CSW, MU)
    # a coarse step makes dH > 0 large; u = 1 rejects any dH > 0
    p = smg.nested_params(2.0, 1.0, 1, 1, False, False, smg.F64, 2000, 1e-12)
    res = smg.hmc_trajectory_nested(D, p, pi, eta, 1.0 - 1e-16)
    assert res.dH > 0 and res.accepted == 0
    assert np.array_equal(smg.gauge_of(D), U)
