"""The multi-timescale MG HMC on the GPU (hmc_trajectory_mg_multiscale,
src/hmc_multiscale.cpp:340-888, and coarse_lowmode_force, :237-276) against the
oracle's restatement, which the CPU suite pins to the compiled reference, and against
the reference's own trajectory (golden fixture)."""
import os

import numpy as np
import pytest

import oracle as O
from helpers import banner, relerr
from schwinger_mg_amd.fields import gaussian_vec, gaussian_vecs, random_gauge

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_vectors.npz")
MASS, BETA = 0.1, 2.0


def _setup(smg, L, seed, n_defl=6, b=4, k=4):
    U = random_gauge(L, 0.3, seed)
    nulls = gaussian_vecs(k, 2 * L * L, seed + 1)
    D = smg.DiracOp(L, U, MASS, 1.0, 1.0, 0.0)
    P = smg.Prolongator(L, L, 2, b, b, k)
    P.build_from_vectors(nulls)
    C = smg.SparseCoarseOp.build(P, D)
    C.setup_deflation(n_defl)
    Do = O.Dirac(L, U, MASS, 1.0, 1.0, 0.0)
    Po = O.Prolongator(L, L, 2, b, b, k).build(nulls.ravel())
    return U, D, P, C, Do, Po


@pytest.mark.parametrize("smooth", [0, 3])
def test_lowmode_force_vs_oracle(smg, smooth):
    banner(f"coarse_lowmode_force smooth={smooth}")
    L = 16
    U, D, P, C, Do, Po = _setup(smg, L, 300)
    phi = gaussian_vec(2 * L * L, 302)
    F = smg.coarse_lowmode_force(D, P, C, phi, smooth)
    Fo = O.coarse_lowmode_force(Do, Po, C.defl_vecs, C.defl_vals, phi, smooth)
    print(f"relerr {relerr(F, Fo):.2e}")
    assert relerr(F, Fo) < 1e-10


def test_evolve_deflation_vs_oracle(smg):
    """evolve_coarse_deflation: the space of one gauge field Rayleigh-Ritz rotated on the
    Galerkin operator of another."""
    banner("evolve_coarse_deflation")
    L = 16
    U, D, P, C, Do, Po = _setup(smg, L, 310)
    U2 = U * np.exp(0.05j * np.random.default_rng(311).standard_normal(U.shape))
    D2 = smg.DiracOp(L, U2, MASS, 1.0, 1.0, 0.0)
    C2 = smg.SparseCoarseOp.build(P, D2)
    V0, v0 = C.defl_vecs, C.defl_vals
    smg.evolve_coarse_deflation(C, C2)
    Co2 = O.Coarse(O.Dirac(L, U2, MASS, 1.0, 1.0, 0.0).op(), Po, 1)
    Vo, vo = O.rr_evolve(Co2.op(), V0)
    assert np.abs(C.defl_vals - vo).max() < 1e-11 * vo.max()
    for i in range(len(vo)):   # Ritz vectors fixed up to a phase
        a, b_ = Vo[i], C.defl_vecs[i]
        ov = np.vdot(b_, a)
        assert np.linalg.norm(b_ * ov / abs(ov) - a) < 1e-9


@pytest.mark.parametrize("omelyan,smooth,refresh", [(0, 0, 0), (1, 3, 0), (1, 3, 2), (0, 2, 1)])
def test_ms_evolve_vs_oracle(smg, omelyan, smooth, refresh):
    banner(f"ms evolve omelyan={omelyan} smooth={smooth} refresh={refresh}")
    L = 16
    U, D, P, C, Do, Po = _setup(smg, L, 320)
    V0, v0 = C.defl_vecs, C.defl_vals
    pi = np.random.default_rng(321).standard_normal(2 * L * L)
    phi = Do.apply(1, gaussian_vec(2 * L * L, 322))
    p = smg.ms_params(BETA, 0.3, 3, 3, omelyan=omelyan, inner_smooth=smooth, defl_refresh=refresh, cg_maxiter=1000,
                      cg_tol=1e-12)
    pg, itg, nlg = smg.hmc_ms_evolve(D, P, C, p, pi, phi)
    Uo, po, Vo, vo, ito, nlo = O.hmc_ms_evolve(L, U, MASS, 1.0, 1.0, 0.0, BETA, 0.3, 3, 3, omelyan, smooth, refresh,
                                               1000, 1e-12, Po, V0, v0, pi, phi)
    print(f"U {relerr(smg.gauge_of(D), Uo):.2e} pi {relerr(pg, po):.2e} vals {np.abs(C.defl_vals - vo).max():.2e} "
          f"cg {itg} vs {ito}")
    assert nlg == nlo
    assert relerr(smg.gauge_of(D), Uo) < 1e-10 and relerr(pg, po) < 1e-9
    assert np.abs(C.defl_vals - vo).max() < 1e-10 * vo.max()


def test_ms_trajectory_golden(smg):
    """The reference's trajectory (golden fixture): the same dH, final gauge field and
    evolved deflation space from its draws and its initial space."""
    banner("ms trajectory golden L=16")
    g = np.load(GOLD)
    L = 16
    D = smg.DiracOp(L, g["MS_U"], MASS, 1.0, 1.0, 0.0)
    P = smg.Prolongator(L, L, 2, 4, 4, 4)
    P.build_from_vectors(g["MS_nulls"])
    C = smg.SparseCoarseOp.build(P, D)
    C.set_deflation(g["MS_vals0"], g["MS_vecs0"])
    out = g["MS_out"]
    p = smg.ms_params(BETA, 0.3, 3, 3, omelyan=1, inner_smooth=3, defl_refresh=2, cg_maxiter=500, cg_tol=1e-12,
                      force_accept=bool(out[0]))
    res = smg.hmc_trajectory_mg_multiscale(D, P, C, p, g["MS_pi"], g["MS_eta2"], 0.5)
    print(f"dH {res.dH:.10e} vs {out[1]:.10e}; low-mode evaluations {res.lowmode_force_evals} vs {int(out[6])}")
    assert res.accepted == int(out[0]) and res.lowmode_force_evals == int(out[6])
    assert abs(res.dH - out[1]) < 1e-8
    assert relerr(smg.gauge_of(D), g["MS_Uf"]) < 1e-10
    assert np.abs(C.defl_vals - g["MS_vals1"]).max() < 1e-10


def test_ms_trajectory_mg_preconditioned_and_reject(smg):
    """With the MG cycle preconditioning the outer solves the trajectory is the same to
    the solve tolerance; a rejected trajectory restores the gauge field exactly (the
    deflation space keeps its evolution, as in the reference)."""
    banner("ms trajectory with MG, reject")
    L = 16
    U, D, P, C, Do, Po = _setup(smg, L, 330)
    pi = np.random.default_rng(331).standard_normal(2 * L * L)
    eta = gaussian_vec(2 * L * L, 332)
    starts = gaussian_vecs(4, 2 * L * L, 333).ravel()
    mg = smg.MGHierarchy(D, starts, n_levels=2, blocks=((4, 4),), k=(4,), null_iters=10)
    p = smg.ms_params(BETA, 0.3, 3, 3, omelyan=1, inner_smooth=3, cg_maxiter=1000, cg_tol=1e-12, force_accept=True)
    D1 = smg.DiracOp(L, U, MASS, 1.0, 1.0, 0.0)
    C1 = smg.SparseCoarseOp.build(P, D1)
    C1.set_deflation(C.defl_vals, C.defl_vecs)
    r0 = smg.hmc_trajectory_mg_multiscale(D1, P, C1, p, pi, eta, 0.5)
    r1 = smg.hmc_trajectory_mg_multiscale(D, P, C, p, pi, eta, 0.5, mg=mg)
    print(f"dH plain {r0.dH:.10e}, MG {r1.dH:.10e}; CG {r0.highmode_cg_iters} vs {r1.highmode_cg_iters}")
    assert abs(r0.dH - r1.dH) < 1e-8 and relerr(smg.gauge_of(D), smg.gauge_of(D1)) < 1e-9
    assert r1.highmode_cg_iters < r0.highmode_cg_iters
    D2 = smg.DiracOp(L, U, MASS, 1.0, 1.0, 0.0)
    C2 = smg.SparseCoarseOp.build(P, D2)
    C2.set_deflation(C1.defl_vals, C1.defl_vecs)
    p.force_accept = 0
    r2 = smg.hmc_trajectory_mg_multiscale(D2, P, C2, p, pi, eta, 2.0)   # u = 2 rejects unless dH < 0
    if not r2.accepted:
        assert np.array_equal(smg.gauge_of(D2), U)
