"""The coarsest-level Chebyshev iteration of the batched FP32 cycle (coarse_solver 2,
BF16 coarsest vectors): its Lanczos spectral interval against the dense coarsest
operator's eigenvalues, its solve against the oracle's Chebyshev restatement on the
FP64 operator, graph replays of the cycle, and mixed solves that reach the 1e-10 true
residual with it."""
import numpy as np
import pytest

import oracle as O
from helpers import banner, relerr
from oracle.cheb import cheb_solve, links_to_dense
from schwinger_mg_amd.fields import gaussian_vecs, random_gauge

pytestmark = pytest.mark.gpu

L, K, STEPS = 32, 16, 20


@pytest.fixture(scope="module")
def cheb_mg(smg):
    banner(f"Chebyshev coarsest hierarchy L={L} k={K} steps={STEPS}")
    V = L * L
    U = random_gauge(L, 0.4, 300)
    D = smg.DiracOp(L, U, 0.05, 1.0, 1.0, 0.0)
    starts = gaussian_vecs(K // 2, 2 * V, 301).ravel()
    mg = smg.MGHierarchy(D, starts, n_levels=2, blocks=((4, 4),), k=(K,), null_iters=5, coarse_maxit=STEPS,
                         coarse_tol=0.1, coarse_half=2, coarse_solver=2, chiral=True, w_cycle=False)
    nc = L // 4
    M = links_to_dense(mg.coarse_links(0), nc, nc, K)
    return dict(U=U, D=D, mg=mg, M=M, starts=starts)


def test_cheb_interval_brackets_the_spectrum(smg, cheb_mg):
    """lo is the smallest Ritz value (at or above the smallest eigenvalue, and close to
    it), hi = 1.1 x the largest Ritz value (at or above the largest eigenvalue)."""
    lo, hi = cheb_mg["mg"].cheb_interval()
    ev = np.linalg.eigvalsh(cheb_mg["M"])
    print(f"interval [{lo:.6g}, {hi:.6g}], spectrum [{ev[0]:.6g}, {ev[-1]:.6g}]")
    assert ev[0] * (1 - 1e-9) <= lo <= 1.5 * ev[0]
    assert ev[-1] <= hi <= 1.1 * ev[-1] * (1 + 1e-9)


def test_cheb_coarsest_solve_vs_oracle(smg, cheb_mg):
    """16 right-hand sides at once: the device iteration (FP16 links, BF16 vectors)
    against the oracle's FP64 iteration on the same interval; the residual falls as
    the FP64 polynomial's does."""
    mg, M = cheb_mg["mg"], cheb_mg["M"]
    lo, hi = mg.cheb_interval()
    N = 16
    B = gaussian_vecs(N, M.shape[0], 302)
    X = mg.coarsest_batched(B)
    errs, res_dev, res_ref = [], [], []
    for j in range(N):
        x_ref = cheb_solve(lambda v: M @ v, B[j], lo, hi, STEPS)
        errs.append(np.linalg.norm(X[j] - x_ref) / np.linalg.norm(x_ref))
        res_dev.append(np.linalg.norm(B[j] - M @ X[j]) / np.linalg.norm(B[j]))
        res_ref.append(np.linalg.norm(B[j] - M @ x_ref) / np.linalg.norm(B[j]))
    print(f"max relative error vs the FP64 iteration {max(errs):.2e}; residual device {max(res_dev):.3f}, "
          f"FP64 {max(res_ref):.3f}")
    assert max(errs) < 3e-2
    assert max(res_dev) < max(res_ref) + 3e-2 and max(res_ref) < 0.9


def test_cheb_cycle_graph_replays_equal_eager(smg, cheb_mg):
    """The batched FP32 cycle with the Chebyshev coarsest solve has no reductions there;
    captured as a hipGraph, three replays (eager work between them) equal the eager
    cycle bit for bit."""
    mg, D = cheb_mg["mg"], cheb_mg["D"]
    B = gaussian_vecs(32, 2 * L * L, 303)
    mg.set_cycle(3, 3, STEPS, 0.1, use_graph=False)
    eager = mg.precondition_batched(B, smg.F32)
    mg.set_cycle(3, 3, STEPS, 0.1, use_graph=True)
    for rep in range(3):
        assert np.array_equal(mg.precondition_batched(B, smg.F32), eager), rep
        D.apply(B[rep])
    mg.set_cycle(3, 3, STEPS, 0.1, use_graph=False)


@pytest.mark.parametrize("levels", [2, 3])
def test_mixed_solve_with_cheb_coarsest(smg, levels):
    """The mixed batched solve with the Chebyshev coarsest solve reaches the 1e-10 FP64
    true residual (checked with the oracle's operator) in about as many iterations as
    with the BF16 coarsest CG, and the same solution."""
    banner(f"mixed solve, Chebyshev coarsest, {levels} levels")
    k, N = 32, 32
    V = L * L
    U = random_gauge(L, 0.4, 310)
    D = smg.DiracOp(L, U, 0.05, 1.0, 1.0, 0.0)
    dims = [2 * V, (L // 4) ** 2 * k]
    starts = np.concatenate([gaussian_vecs(k // 2, n, 311 + i).ravel() for i, n in enumerate(dims[:levels - 1])])
    blocks = ((4, 4), (2, 2))[:levels - 1]
    B = gaussian_vecs(N, 2 * V, 312)
    A = O.Dirac(L, U, 0.05, 1.0, 1.0, 0.0).op()
    out = {}
    for solver, maxit in ((0, 30), (2, 40)):
        mg = smg.MGHierarchy(D, starts, n_levels=levels, blocks=blocks, k=(k,) * (levels - 1), null_iters=5,
                             coarse_maxit=maxit, coarse_tol=0.1, coarse_solver=solver, coarse_half=2, chiral=True,
                             w_cycle=False)
        X, info, tr = smg.gcr_solve_batched(D, B, mg, 200, 1e-10, 10, prec=smg.F32, inner_tol=1e-5)
        o = np.array([np.linalg.norm(B[j] - A(X[j])) / np.linalg.norm(B[j]) for j in range(N)])
        assert o.max() < 1e-10, (solver, o.max())
        out[solver] = (X, info.iterations)
    print(f"iterations: BF16 CG {out[0][1]}, Chebyshev {out[2][1]}")
    assert out[2][1] <= out[0][1] + 3
    assert relerr(out[2][0], out[0][0]) < 1e-8


def test_set_cheb_interval(smg, cheb_mg):
    """A caller's interval replaces the Lanczos one (the solve then follows the oracle's
    iteration on it), invalid ones are refused, and a rebuild estimates it again."""
    mg, M = cheb_mg["mg"], cheb_mg["M"]
    lo, hi = mg.cheb_interval()
    mg.set_cheb_interval(2.0 * lo, hi)
    assert mg.cheb_interval() == (2.0 * lo, hi)
    B = gaussian_vecs(16, M.shape[0], 304)
    X = mg.coarsest_batched(B)
    errs = [np.linalg.norm(X[j] - cheb_solve(lambda v: M @ v, B[j], 2.0 * lo, hi, STEPS))
            / np.linalg.norm(cheb_solve(lambda v: M @ v, B[j], 2.0 * lo, hi, STEPS)) for j in range(16)]
    assert max(errs) < 3e-2
    for bad in ((0.0, hi), (hi, lo), (-1.0, 1.0)):
        with pytest.raises(smg.SMGError):
            mg.set_cheb_interval(*bad)
    mg.rebuild()
    lo2, hi2 = mg.cheb_interval()
    assert abs(lo2 - lo) <= 1e-9 * lo and abs(hi2 - hi) <= 1e-9 * hi
