"""HMC molecular dynamics on the GPU (the gauge field evolving in the device-resident
operator) against the oracle's restatement of plain_leapfrog_evolve
(src/hmc_trajectory.cpp:243-409), which the CPU suite pins bit-for-bit to the
compiled reference; plus what the mathematics fixes: time reversibility, second
order in dt, and an exact restore of the gauge field on rejection."""
import os

import numpy as np
import pytest

import oracle as O
from helpers import banner, relerr
from schwinger_mg_amd.fields import gaussian_vec, random_gauge

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_vectors.npz")
MASS, R_W, BETA = 0.1, 1.0, 2.0


@pytest.mark.parametrize("L,csw,omelyan", [(8, 1.0, 0), (8, 1.0, 1), (16, 0.0, 0), (16, 1.0, 1)])
def test_md_evolve_vs_oracle(smg, L, csw, omelyan):
    banner(f"hmc evolve L={L} csw={csw} omelyan={omelyan}")
    V = L * L
    U = random_gauge(L, 0.3, 70 + L)
    pi = np.random.default_rng(71).standard_normal(2 * V)
    phi = gaussian_vec(2 * V, 72)
    D = smg.DiracOp(L, U, MASS, R_W, csw, 0.0)
    p = smg.hmc_params(BETA, 0.5, 4, omelyan=omelyan, cg_maxiter=2000, cg_tol=1e-12)
    pg, itg = smg.hmc_evolve(D, p, pi, phi)
    Ug = smg.gauge_of(D)
    Uo, po, ito = O.hmc_evolve(L, U, MASS, R_W, csw, 0.0, BETA, 0.5, 4, omelyan, 2000, 1e-12, pi, phi)
    print(f"U relerr {relerr(Ug, Uo):.2e}, pi relerr {relerr(pg, po):.2e}, CG {itg} vs {ito}")
    assert relerr(Ug, Uo) < 1e-11 and relerr(pg, po) < 1e-10
    assert abs(itg - ito) <= 2 * (4 * (omelyan + 1) + 1)   # each solve within a couple of iterations
    # the evolved operator is the operator of the evolved links (clover and FP32 copies follow)
    psi = gaussian_vec(2 * V, 73)
    assert relerr(D.apply(psi), O.Dirac(L, Uo, MASS, R_W, csw, 0.0).apply(0, psi)) < 1e-10


def test_md_evolve_golden(smg):
    banner("hmc evolve golden L=6")
    g = np.load(GOLD)
    for om in (0, 1):
        D = smg.DiracOp(6, g["L6_U"], MASS, R_W, 1.0, 0.0)
        pg, _ = smg.hmc_evolve(D, smg.hmc_params(BETA, 0.5, 3, omelyan=om, cg_maxiter=500, cg_tol=1e-12),
                               g["L6_pi"], g["L6_phi"])
        assert relerr(smg.gauge_of(D), g[f"L6_md{om}_U"]) < 1e-11 and relerr(pg, g[f"L6_md{om}_pi"]) < 1e-10


def test_hamiltonian_vs_oracle(smg):
    banner("hmc hamiltonian L=16")
    L, V = 16, 256
    U = random_gauge(L, 0.4, 80)
    pi = np.random.default_rng(81).standard_normal(2 * V)
    phi = gaussian_vec(2 * V, 82)
    D = smg.DiracOp(L, U, MASS, R_W, 1.0, 0.0)
    hg = smg.hmc_hamiltonian(D, BETA, pi, phi, 2000, 1e-12)
    ho = O.hmc_hamiltonian(L, U, MASS, R_W, 1.0, 0.0, BETA, 2000, 1e-12, pi, phi)
    assert np.all(np.abs(hg[:3] - ho) <= 1e-10 * np.abs(ho)) and hg[3] == 0.0
    # even-odd: S_f on the Schur operator and LD = -2 log det D_ee
    phi_o = gaussian_vec(V, 83)
    hg = smg.hmc_hamiltonian(D, BETA, pi, phi_o, 2000, 1e-12, use_eo=True)
    ho = O.hmc_hamiltonian_eo(L, U, MASS, R_W, 1.0, 0.0, BETA, 2000, 1e-12, pi, phi_o)
    assert np.all(np.abs(hg - ho) <= 1e-10 * np.abs(ho))


@pytest.mark.parametrize("omelyan", [0, 1])
def test_md_reversible_and_second_order(smg, omelyan):
    banner(f"hmc reversibility omelyan={omelyan}")
    L, V = 16, 256
    U = random_gauge(L, 0.3, 90)
    pi = np.random.default_rng(91).standard_normal(2 * V)
    phi = gaussian_vec(2 * V, 92)
    D = smg.DiracOp(L, U, MASS, R_W, 1.0, 0.0)
    p = smg.hmc_params(BETA, 0.5, 4, omelyan=omelyan, cg_maxiter=4000, cg_tol=1e-14)
    p1, _ = smg.hmc_evolve(D, p, pi, phi)
    p2, _ = smg.hmc_evolve(D, p, -p1, phi)
    print(f"reversibility: U {relerr(smg.gauge_of(D), U):.2e}, pi {relerr(-p2, pi):.2e}")
    assert relerr(smg.gauge_of(D), U) < 1e-11 and relerr(-p2, pi) < 1e-10
    dH = []
    for n in (8, 16):
        Dn = smg.DiracOp(L, U, MASS, R_W, 1.0, 0.0)
        h0 = smg.hmc_hamiltonian(Dn, BETA, pi, phi, 4000, 1e-14).sum()
        pf, _ = smg.hmc_evolve(Dn, smg.hmc_params(BETA, 0.5, n, omelyan=omelyan, cg_maxiter=4000, cg_tol=1e-14),
                               pi, phi)
        dH.append(abs(smg.hmc_hamiltonian(Dn, BETA, pf, phi, 4000, 1e-14).sum() - h0))
    print(f"dH at 8 / 16 steps: {dH[0]:.3e} / {dH[1]:.3e}")
    assert 3.0 < dH[0] / dH[1] < 5.0


def test_trajectory_metropolis(smg):
    """hmc_trajectory: dH is H_final - H_initial of the evolution; a rejected
    trajectory leaves the gauge field exactly as it was, an accepted one holds the
    evolved field."""
    banner("hmc trajectory L=8")
    L, V = 8, 64
    U = random_gauge(L, 0.3, 100)
    pi = np.random.default_rng(101).standard_normal(2 * V)
    eta = gaussian_vec(2 * V, 102)
    phi = O.Dirac(L, U, MASS, R_W, 1.0, 0.0).apply(1, eta)   # D^dag eta
    p = smg.hmc_params(BETA, 0.5, 4, cg_maxiter=2000, cg_tol=1e-12)
    ho = O.hmc_hamiltonian(L, U, MASS, R_W, 1.0, 0.0, BETA, 2000, 1e-12, pi, phi).sum()
    Uo, po, _ = O.hmc_evolve(L, U, MASS, R_W, 1.0, 0.0, BETA, 0.5, 4, 0, 2000, 1e-12, pi, phi)
    dHo = O.hmc_hamiltonian(L, Uo, MASS, R_W, 1.0, 0.0, BETA, 2000, 1e-12, po, phi).sum() - ho
    D = smg.DiracOp(L, U, MASS, R_W, 1.0, 0.0)
    res = smg.hmc_trajectory(D, p, pi, eta, 2.0)   # u = 2 > exp(-dH) unless dH < 0
    assert abs(res.dH - dHo) < 1e-8
    assert bool(res.accepted) == (dHo < 0)
    if res.accepted:
        assert relerr(smg.gauge_of(D), Uo) < 1e-11
    else:
        assert np.array_equal(smg.gauge_of(D), U)
    D2 = smg.DiracOp(L, U, MASS, R_W, 1.0, 0.0)
    p.force_accept = 1
    res2 = smg.hmc_trajectory(D2, p, pi, eta, 2.0)
    assert res2.accepted == 1 and relerr(smg.gauge_of(D2), Uo) < 1e-11
    assert abs(res2.dH - dHo) < 1e-8


@pytest.mark.parametrize("csw", [0.0, 1.0])
def test_eo_evolve_and_trajectory_vs_oracle(smg, csw):
    """The even-odd trajectory (src/hmc_trajectory.cpp:53-135): leapfrog on the Schur
    force, phi_o = M^dag eta_o, dH with the log-det term."""
    banner(f"hmc even-odd csw={csw}")
    L, V = 16, 256
    U = random_gauge(L, 0.3, 110)
    pi = np.random.default_rng(111).standard_normal(2 * V)
    eta = gaussian_vec(V, 112)
    Do = O.Dirac(L, U, MASS, R_W, csw, 0.0)
    phi_o = Do.eo(6, eta)   # M^dag eta_o
    D = smg.DiracOp(L, U, MASS, R_W, csw, 0.0)
    p = smg.hmc_params(BETA, 0.5, 4, cg_maxiter=2000, cg_tol=1e-12, use_eo=True)
    pg, _ = smg.hmc_evolve(D, p, pi, phi_o)
    Uo, po, _ = O.hmc_evolve_eo(L, U, MASS, R_W, csw, 0.0, BETA, 0.5, 4, 2000, 1e-12, pi, phi_o)
    print(f"eo evolve: U {relerr(smg.gauge_of(D), Uo):.2e}, pi {relerr(pg, po):.2e}")
    assert relerr(smg.gauge_of(D), Uo) < 1e-11 and relerr(pg, po) < 1e-10
    h0 = O.hmc_hamiltonian_eo(L, U, MASS, R_W, csw, 0.0, BETA, 2000, 1e-12, pi, phi_o)
    h1 = O.hmc_hamiltonian_eo(L, Uo, MASS, R_W, csw, 0.0, BETA, 2000, 1e-12, po, phi_o)
    D2 = smg.DiracOp(L, U, MASS, R_W, csw, 0.0)
    p.force_accept = 1
    res = smg.hmc_trajectory(D2, p, pi, eta, 0.5)
    assert res.accepted == 1 and relerr(smg.gauge_of(D2), Uo) < 1e-11
    assert abs(res.dH - (h1 - h0).sum()) < 1e-8 and abs(res.dLD - (h1[3] - h0[3])) < 1e-9
    if csw == 0.0:
        assert res.dLD == 0.0


def test_trajectory_golden(smg):
    """hmc_trajectory of the compiled reference (full lattice and even-odd) on its own
    draws, reproduced from the same momenta and Gaussians."""
    banner("hmc trajectory golden L=6")
    g = np.load(GOLD)
    for eo in (0, 1):
        D = smg.DiracOp(6, g["L6_U"], MASS, R_W, 1.0, 0.0)
        p = smg.hmc_params(BETA, 0.5, 3, cg_maxiter=500, cg_tol=1e-12, force_accept=True, use_eo=bool(eo))
        res = smg.hmc_trajectory(D, p, g[f"L6_traj{eo}_pi"], g[f"L6_traj{eo}_eta"], 0.5)
        want = g[f"L6_traj{eo}_out"]
        assert relerr(smg.gauge_of(D), g[f"L6_traj{eo}_U"]) < 1e-11
        assert abs(res.dH - want[0]) < 1e-9 and abs(res.dLD - want[4]) < 1e-9
