"""Parity of the HIP fermion/gauge forces with the oracle and the golden vectors,
and the adjoint force against central finite differences of the action computed
with the cancellation-free identity S(U') - S(U) = -X'^dag (A' - A) X
(A = D^dag D, X = A^-1 phi, X' = A'^-1 phi) to <= 1e-7 relative."""
import os

import numpy as np
import pytest

import oracle as O
from helpers import CONFIGS, banner, relerr
from schwinger_mg_amd.fields import gaussian_vec, random_gauge

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_vectors.npz")


@pytest.mark.parametrize("ci", range(4))
def test_forces_vs_oracle(smg, ci):
    csw, mu = CONFIGS[ci]
    L = 16
    banner(f"forces L={L} csw={csw} mu={mu}")
    V = L * L
    U = random_gauge(L, 0.5, 60)
    X, chi = gaussian_vec(2 * V, 61), gaussian_vec(2 * V, 62)
    Dg, Do = smg.DiracOp(L, U, 0.1, 1.0, csw, mu), O.Dirac(L, U, 0.1, 1.0, csw, mu)
    assert relerr(smg.fermion_force(Dg, X), Do.force(X)) < 1e-12
    assert relerr(smg.force_forward(Dg, X), Do.apply(0, X)) < 1e-12
    assert relerr(smg.fermion_force_bilinear(Dg, chi, X), Do.force_bilinear(chi, X)) < 1e-12
    assert relerr(smg.hopping_force_bilinear(Dg, chi, X), Do.hopping_force(chi, X)) < 1e-12
    if csw:
        assert relerr(smg.logdet_ee_force(Dg), Do.logdet_force()) < 1e-12
    assert relerr(smg.eo_fermion_force(Dg, X[:V], chi[:V]), Do.eo_force(X[:V], chi[:V])) < 1e-12
    assert relerr(smg.gauge_force(L, U, 4.0), O.gauge_force(L, U, 4.0)) < 1e-12
    assert abs(smg.gauge_action(L, U, 4.0) - O.gauge_action(L, U, 4.0)) < 1e-11 * abs(O.gauge_action(L, U, 4.0))


@pytest.mark.parametrize("L", [2, 6, 8, 12, 16, 30, 48])
@pytest.mark.parametrize("ci", range(4))
def test_fused_force_recomputes_chi(smg, L, ci):
    """fermion_force runs one kernel that recomputes chi = D X per tile in LDS: it equals
    the two-pass forward (chi stored) + backward pair and the oracle, on lattices smaller
    than, equal to and not a multiple of the 16 x 16 tile."""
    csw, mu = CONFIGS[ci]
    banner(f"fused force L={L} csw={csw} mu={mu}")
    V = L * L
    U = random_gauge(L, 0.5, 70 + L)
    X = gaussian_vec(2 * V, 71 + L)
    Dg = smg.DiracOp(L, U, 0.1, 1.0, csw, mu)
    fused = smg.fermion_force(Dg, X)
    two_pass = smg.fermion_force_bilinear(Dg, smg.force_forward(Dg, X), X)
    assert relerr(fused, two_pass) < 1e-14
    assert relerr(fused, O.Dirac(L, U, 0.1, 1.0, csw, mu).force(X)) < 1e-12


def test_forces_golden(smg):
    banner("forces golden L=8")
    g = np.load(GOLD)
    L, V = 8, 64
    for ci, (csw, mu) in enumerate(CONFIGS):
        D = smg.DiracOp(L, g["L8_U"], 0.1, 1.0, csw, mu)
        assert relerr(smg.fermion_force_bilinear(D, g["L8_chi"], g["L8_psi"]), g[f"L8_c{ci}_force1"]) < 1e-12
        assert relerr(smg.eo_fermion_force(D, g["L8_psi"][:V], g["L8_chi"][:V]), g[f"L8_c{ci}_force_eo"]) < 1e-12


@pytest.mark.parametrize("csw", [0.0, 1.0])
def test_adjoint_force_vs_finite_differences(smg, csw):
    """64 random links, eps = 1e-5, FP64 solves to 1e-14."""
    L, m, eps = 32, 0.1, 1e-5
    banner(f"force finite differences L={L} csw={csw}")
    V = L * L
    U = random_gauge(L, 0.35, 70)
    eta = gaussian_vec(2 * V, 71)
    D = smg.DiracOp(L, U, m, 1.0, csw, 0.0)
    phi = D.apply_dag(eta)
    X = smg.cg_solve(D, phi, 5000, 1e-14).solution
    F = smg.fermion_force(D, X)
    AX = D.apply_DdagD(X)
    links = np.random.default_rng(72).choice(2 * V, size=64, replace=False)
    worst = 0.0
    for li in links:
        dS = []
        for sgn in (1, -1):
            Up = U.copy()
            Up[li] *= np.exp(1j * sgn * eps)
            Dp = smg.DiracOp(L, Up, m, 1.0, csw, 0.0)
            Xp = smg.cg_solve(Dp, phi, 5000, 1e-14).solution
            dS.append(-np.vdot(Xp, Dp.apply_DdagD(X) - AX).real)
        fd = -(dS[0] - dS[1]) / (2 * eps)
        worst = max(worst, abs(fd - F[li]))
    assert worst / np.abs(F).max() < 1e-7, worst / np.abs(F).max()
