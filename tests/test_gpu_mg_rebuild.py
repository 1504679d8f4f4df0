"""MGHierarchy::rebuild_deeper_levels (with the level-0 Ac.build) and
setup_sparse_coarse (src/multigrid.cpp:163-194, 754-794) on the GPU, against the
oracle's Galerkin products and dense coarsest solve."""
import numpy as np
import pytest

import oracle as O
from helpers import banner, relerr
from schwinger_mg_amd.fields import gaussian_vec, gaussian_vecs, random_gauge

pytestmark = pytest.mark.gpu


def test_rebuild_after_gauge_change(smg):
    """After the operator's links change, every level's links equal the Galerkin
    products of the new operator with the hierarchy's (unchanged) prolongators, and the
    rebuilt hierarchy preconditions the new system."""
    banner("MG rebuild after a gauge change")
    L, k = 16, 4
    U1 = random_gauge(L, 0.3, 400)
    U2 = U1 * np.exp(0.2j * np.random.default_rng(401).standard_normal(U1.shape))
    D = smg.DiracOp(L, U1, 0.05, 1.0, 1.0, 0.0)
    dims = [2 * L * L, (L // 4) ** 2 * k]
    starts = np.concatenate([gaussian_vecs(k, n, 402 + i).ravel() for i, n in enumerate(dims)])
    mg = smg.MGHierarchy(D, starts, n_levels=3, blocks=((4, 4), (2, 2)), k=(k, k), null_iters=5, coarse_solver=1)
    bv0, bv1 = mg.block_vecs(0), mg.block_vecs(1)
    old0 = mg.coarse_links(0)
    D.set_gauge(U2)
    mg.rebuild()
    assert np.array_equal(mg.block_vecs(0), bv0) and np.array_equal(mg.block_vecs(1), bv1)
    P0 = O.Prolongator(L, L, 2, 4, 4, k)
    P0.set_block_vecs(bv0)
    C0 = O.Coarse(O.Dirac(L, U2, 0.05, 1.0, 1.0, 0.0).op(), P0, 1)
    assert relerr(mg.coarse_links(0), C0.links()) < 1e-12 and relerr(old0, C0.links()) > 1e-3
    P1 = O.Prolongator(L // 4, L // 4, k, 2, 2, k)
    P1.set_block_vecs(bv1)
    C1 = O.Coarse(C0.op(), P1, 1)
    assert relerr(mg.coarse_links(1), C1.links()) < 1e-12
    b = gaussian_vec(2 * L * L, 403)
    r = smg.fcg_solve(D, b, mg, 200, 1e-10, 10)
    A = O.Dirac(L, U2, 0.05, 1.0, 1.0, 0.0).op()
    assert np.linalg.norm(A(r.solution) - b) / np.linalg.norm(b) < 1e-10


def test_setup_sparse_coarse(smg):
    """The deflated coarsest CG to 1e-12 solves the level-1 system exactly: the cycle
    equals the oracle's with the dense coarsest solve; deeper hierarchies are refused,
    as the reference's wiring only fits two levels."""
    banner("setup_sparse_coarse")
    L, k = 16, 4
    U = random_gauge(L, 0.3, 410)
    D = smg.DiracOp(L, U, 0.05, 1.0, 1.0, 0.0)
    starts = gaussian_vecs(k, 2 * L * L, 411).ravel()
    mg = smg.MGHierarchy(D, starts, n_levels=2, blocks=((4, 4),), k=(k,), null_iters=10, coarse_maxit=3,
                         coarse_tol=0.5)
    mgo = O.MG(O.Dirac(L, U, 0.05, 1.0, 1.0, 0.0).op(), L, L, 2, starts, n_levels=2, block=((4, 4),), k=(k,),
               null_iters=10, coarse_solver=1)
    b = gaussian_vec(2 * L * L, 412)
    loose = mg.precondition(b)
    mg.setup_sparse_coarse(8, 1e-12, 500)
    x = mg.precondition(b)
    xo = mgo(b)
    print(f"cycle vs oracle (dense coarsest): {relerr(x, xo):.2e}; before: {relerr(loose, xo):.2e}")
    assert relerr(x, xo) < 1e-9 and relerr(loose, xo) > 1e-4
    mg3 = smg.MGHierarchy(D, np.concatenate([starts, gaussian_vecs(k, (L // 4) ** 2 * k, 413).ravel()]),
                          n_levels=3, blocks=((4, 4), (2, 2)), k=(k, k), null_iters=5)
    with pytest.raises(smg.SMGError):
        mg3.setup_sparse_coarse(4)


def test_refresh_is_warm_started_setup(smg):
    """MGHierarchy refresh (build_full_mg with warm_start, src/mg_builder.cpp:3-31) after
    a gauge change equals a new hierarchy set up on the new field from the current
    near-null vectors of every level with the same iteration count; refresh(0) equals
    the Galerkin rebuild."""
    banner("MG refresh = warm-started setup")
    L, k = 16, 4
    U1 = random_gauge(L, 0.3, 410)
    U2 = U1 * np.exp(0.3j * np.random.default_rng(411).standard_normal(U1.shape))
    D = smg.DiracOp(L, U1, 0.05, 1.0, 1.0, 0.0)
    dims = [2 * L * L, (L // 4) ** 2 * k]
    starts = np.concatenate([gaussian_vecs(k, n, 412 + i).ravel() for i, n in enumerate(dims)])
    kw = dict(n_levels=3, blocks=((4, 4), (2, 2)), k=(k, k), coarse_solver=1)
    mg = smg.MGHierarchy(D, starts, null_iters=5, **kw)
    D.set_gauge(U2)
    warm = np.concatenate([mg.null_vecs(0).ravel(), mg.null_vecs(1).ravel()])
    # two more hierarchies set up on the old field: one rebuilt, one refreshed with 0 iterations
    D.set_gauge(U1)
    mg_rb = smg.MGHierarchy(D, starts, null_iters=5, **kw)
    mg0 = smg.MGHierarchy(D, starts, null_iters=5, **kw)
    D.set_gauge(U2)
    mg_rb.rebuild()
    mg0.refresh(0)
    b = gaussian_vec(2 * L * L, 415)
    assert relerr(mg0.precondition(b), mg_rb.precondition(b)) < 1e-13
    mg.refresh(3)
    fresh = smg.MGHierarchy(D, warm, null_iters=3, **kw)
    assert relerr(mg.null_vecs(0), fresh.null_vecs(0)) < 1e-12
    assert relerr(mg.precondition(b), fresh.precondition(b)) < 1e-11
    r = smg.fcg_solve(D, b, mg, 100, 1e-10, 10)
    assert r.true_residual < 1.1e-10
