"""FGMRES with the Ritz harvest and intra-solve deflation (fgmres_solve_generic,
src/solvers.cpp:555-772) on the device against the oracle and the reference's own
output (tests/golden), through smg_fgmres_solve_ritz."""
import os

import numpy as np
import pytest

import oracle as O
from helpers import CONFIGS, banner, relerr
from schwinger_mg_amd.fields import gaussian_vec, random_gauge

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_vectors.npz")


@pytest.mark.parametrize("ci", (0, 2))
@pytest.mark.parametrize("nr", (2, 4))
def test_golden_fgmres_ritz(smg, ci, nr):
    """Two restart cycles: the pairs agree with the reference's to rounding.  At
    convergence the solution and iteration count agree; the pairs of the last cycles
    start from a residual formed with 1e-9 cancellation and agree to 1e-4 relative."""
    banner(f"fgmres ritz golden L=8 config={ci} n_ritz={nr}")
    g = np.load(GOLD)
    csw, mu = CONFIGS[ci]
    D = smg.DiracOp(8, g["L8_U"], 0.1, 1.0, csw, mu)
    key = f"L8_c{ci}_fgr{nr}_20"
    r = smg.fgmres_solve(D, g["L8_psi"], None, 10, 20, 1e-10, n_ritz_harvest=nr)
    assert r.iterations == 20 and relerr(r.solution, g[key + "_x"]) < 1e-11
    vals = np.array([p.value for p in r.ritz_pairs])
    vecs = np.array([p.vector for p in r.ritz_pairs])
    assert len(vals) == 2 * nr
    assert relerr(vals, g[key + "_vals"]) < 1e-10 and relerr(vecs, g[key + "_vecs"]) < 1e-8
    key = f"L8_c{ci}_fgr{nr}_500"
    r = smg.fgmres_solve(D, g["L8_psi"], None, 10, 500, 1e-10, n_ritz_harvest=nr)
    assert r.iterations == g[key + "_it"][0] and relerr(r.solution, g[key + "_x"]) < 1e-11
    vals = np.array([p.value for p in r.ritz_pairs])
    assert len(vals) == len(g[key + "_vals"])
    assert np.all(np.abs(vals - g[key + "_vals"]) <= 1e-4 * np.abs(g[key + "_vals"]))
    for p in r.ritz_pairs:
        assert abs(np.linalg.norm(p.vector) - 1) < 1e-12


@pytest.mark.parametrize("ci", range(4))
def test_fgmres_ritz_vs_oracle(smg, ci):
    """L = 16, identity preconditioner, 3 pairs from restart 8: the harvested pairs of
    the first two cycles and the converged solution against the oracle."""
    banner(f"fgmres ritz vs oracle L=16 config={ci}")
    L = 16
    csw, mu = CONFIGS[ci]
    U = random_gauge(L, 0.4, 140 + ci)
    b = gaussian_vec(2 * L * L, 150 + ci)
    D, A = smg.DiracOp(L, U, 0.1, 1.0, csw, mu), O.Dirac(L, U, 0.1, 1.0, csw, mu).op()
    r = smg.fgmres_solve(D, b, None, 8, 16, 1e-10, n_ritz_harvest=3)
    xo, ito, _, vo, Vo = O.fgmres_ritz(A, b, 8, 16, 1e-10, 3)
    assert r.iterations == ito == 16 and relerr(r.solution, xo) < 1e-11
    assert relerr([p.value for p in r.ritz_pairs], vo) < 1e-10
    assert relerr(np.array([p.vector for p in r.ritz_pairs]), Vo) < 1e-8
    r = smg.fgmres_solve(D, b, None, 8, 2000, 1e-10, n_ritz_harvest=3)
    xo, ito, reso, vo, _ = O.fgmres_ritz(A, b, 8, 2000, 1e-10, 3)
    # the late cycles' pairs carry the cancellation of their start residuals into the
    # deflation, so the iteration paths agree to a few steps, not to rounding
    assert abs(r.iterations - ito) <= max(3, ito // 20) and relerr(r.solution, xo) < 1e-6
    assert r.final_residual < 1.1e-10 and 1 <= len(r.ritz_pairs) <= 6


def test_fgmres_ritz_mg(smg):
    """The default path of fgmres_solve_mg: MG preconditioner, 4 harvested pairs; the
    deflation by earlier cycles' pairs does not change the converged solution."""
    banner("fgmres ritz with mg L=16")
    g = np.load(GOLD)
    L = 16
    D = smg.DiracOp(L, g["L16_U"], 0.05, 1.0, 1.0, 0.0)
    mg = smg.MGHierarchy(D, g["L16_starts"], n_levels=2, blocks=((4, 4),), k=(4,), null_iters=20,
                         coarse_maxit=5000, coarse_tol=1e-15)
    r4 = smg.fgmres_solve(D, g["L16_rhs"], mg, 4, 100, 1e-10)
    r0 = smg.fgmres_solve(D, g["L16_rhs"], mg, 4, 100, 1e-10, n_ritz_harvest=0)
    assert r4.final_residual < 1.1e-10 and r0.final_residual < 1.1e-10
    assert relerr(r4.solution, r0.solution) < 1e-8
    assert 1 <= len(r4.ritz_pairs) <= 8 and r0.ritz_pairs == []
