"""Parity of the HIP aggregation, Galerkin, coarse operator, smoothers, multigrid
cycle and Krylov solvers with the oracle and the reference's golden vectors."""
import os

import numpy as np
import pytest

import oracle as O
from helpers import banner, relerr
from schwinger_mg_amd.fields import gaussian_vec, gaussian_vecs, random_gauge

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_vectors.npz")


def test_golden_aggregation_galerkin(smg):
    banner("aggregation + galerkin golden L=16")
    g = np.load(GOLD)
    L, b, k = 16, 4, 4
    P = smg.Prolongator(L, L, 2, b, b, k).build_from_vectors(g["L16_nulls"])
    assert relerr(P.block_vecs, g["L16_bv"]) < 1e-13
    assert relerr(P.prolong(g["L16_c"]), g["L16_prolong"]) < 1e-13
    assert relerr(P.restrict_vec(g["L16_rhs"]), g["L16_restrict"]) < 1e-13
    D = smg.DiracOp(L, g["L16_U"], 0.05, 1.0, 1.0, 0.0)
    C = smg.SparseCoarseOp.build(P, D)
    assert relerr(C.link, g["L16_links"]) < 1e-12
    dense = g["L16_dense"]          # dense[j] = column j
    c = g["L16_c"]
    assert relerr(C.apply_to(c), dense.T @ c) < 1e-12


@pytest.mark.parametrize("L,b,k", [(8, 4, 3), (24, 4, 8), (32, 2, 4)])
def test_aggregation_and_coarse_vs_oracle(smg, L, b, k):
    banner(f"aggregation/coarse L={L} b={b} k={k}")
    V = L * L
    U = random_gauge(L, 0.4, L + k)
    nulls = gaussian_vecs(k, 2 * V, 5)
    Pg = smg.Prolongator(L, L, 2, b, b, k).build_from_vectors(nulls)
    Po = O.Prolongator(L, L, 2, b, b, k).build(nulls.ravel())
    assert relerr(Pg.block_vecs, Po.block_vecs()) < 1e-12
    D, Do = smg.DiracOp(L, U, 0.05, 1.0, 1.0, 0.0), O.Dirac(L, U, 0.05, 1.0, 1.0, 0.0)
    C = smg.SparseCoarseOp.build(Pg, D)
    Co = O.Coarse(Do.op(), Po, 1)
    assert relerr(C.link, Co.links()) < 1e-12
    c = gaussian_vec(C.dim, 6)
    want = Co.op()(c)
    assert relerr(C.apply_to(c, smg.F64), want) < 1e-12
    assert relerr(C.apply_to(c, smg.F32), want) < 1e-5
    assert relerr(C.apply_to(c, smg.F16), want) < 5e-3
    # next level: 2 x 2 aggregates of the coarse lattice with nd = k
    nb = L // b
    if nb % 2 == 0:
        cn = gaussian_vecs(k, C.dim, 7)
        P2 = smg.Prolongator(nb, nb, k, 2, 2, k).build_from_vectors(cn)
        P2o = O.Prolongator(nb, nb, k, 2, 2, k).build(cn.ravel())
        C2 = smg.SparseCoarseOp.build_from_coarse(P2, C)
        C2o = O.Coarse(Co.op(), P2o, 1)
        assert relerr(C2.link, C2o.links()) < 1e-11


@pytest.mark.parametrize("L,b,k", [(16, 4, 4), (32, 4, 8)])
def test_dense_coarse_op_vs_oracle(smg, L, b, k):
    """CoarseOp (dense): the matrix equals the oracle's column-by-column Galerkin
    build, and the dense solve equals its LU solve."""
    banner(f"dense coarse op L={L} b={b} k={k}")
    U = random_gauge(L, 0.4, 70 + k)
    nulls = gaussian_vecs(k, 2 * L * L, 71)
    Pg = smg.Prolongator(L, L, 2, b, b, k).build_from_vectors(nulls)
    Po = O.Prolongator(L, L, 2, b, b, k).build(nulls.ravel())
    D, Do = smg.DiracOp(L, U, 0.05, 1.0, 1.0, 0.0), O.Dirac(L, U, 0.05, 1.0, 1.0, 0.0)
    C = smg.SparseCoarseOp.build(Pg, D)
    want = O.dense_galerkin(Do.op(), Po)
    assert relerr(C.dense(), want) < 1e-12
    rhs = gaussian_vec(C.dim, 72)
    assert relerr(C.dense_solve(rhs), O.dense_solve_galerkin(Do.op(), Po, rhs)) < 1e-10


@pytest.mark.parametrize("nv", [2, 4])
def test_chiral_aggregation_and_cycle_vs_oracle(smg, nv):
    """Chiral aggregation (k = 2 nv coarse dof) on the GPU equals the oracle's:
    block vectors, Galerkin links, and a 3-level W-cycle built from the same starts."""
    banner(f"chiral aggregation nv={nv}")
    L, k = 32, 2 * nv
    V = L * L
    U = random_gauge(L, 0.4, 80 + nv)
    nulls = gaussian_vecs(nv, 2 * V, 81)
    Pg = smg.Prolongator(L, L, 2, 4, 4, k, chiral=True).build_from_vectors(nulls)
    Po = O.Prolongator(L, L, 2, 4, 4, k, chiral=True).build(nulls.ravel())
    assert relerr(Pg.block_vecs, Po.block_vecs()) < 1e-12
    D, Do = smg.DiracOp(L, U, 0.05, 1.0, 1.0, 0.0), O.Dirac(L, U, 0.05, 1.0, 1.0, 0.0)
    C = smg.SparseCoarseOp.build(Pg, D)
    Co = O.Coarse(Do.op(), Po, 1)
    assert relerr(C.link, Co.links()) < 1e-12
    dims = [2 * V, (L // 4) ** 2 * k]
    starts = np.concatenate([gaussian_vecs(nv, n, 82 + i).ravel() for i, n in enumerate(dims)])
    mg = smg.MGHierarchy(D, starts, n_levels=3, blocks=((4, 4), (2, 2)), k=(k, k), null_iters=4, coarse_solver=1,
                         chiral=True)
    mgo = O.MG(Do.op(), L, L, 2, starts, n_levels=3, block=((4, 4), (2, 2)), k=(k, k), null_iters=4,
               coarse_solver=1, chiral=True)
    b = gaussian_vec(2 * V, 83)
    assert relerr(mg.precondition(b), mgo(b)) < 1e-10
    r = smg.fcg_solve(D, b, mg, 200, 1e-10, 10, prec=smg.F32)
    assert np.linalg.norm(b - Do.op()(r.solution)) / np.linalg.norm(b) < 1e-10


def test_coarse_solve_and_smoothers(smg):
    banner("coarse solve + smoothers L=16")
    L, b, k = 16, 4, 4
    V = L * L
    U = random_gauge(L, 0.4, 9)
    nulls = gaussian_vecs(k, 2 * V, 10)
    D, Do = smg.DiracOp(L, U, 0.05, 1.0, 1.0, 0.0), O.Dirac(L, U, 0.05, 1.0, 1.0, 0.0)
    P = smg.Prolongator(L, L, 2, b, b, k).build_from_vectors(nulls)
    Po = O.Prolongator(L, L, 2, b, b, k).build(nulls.ravel())
    C, Co = smg.SparseCoarseOp.build(P, D), O.Coarse(Do.op(), Po, 1)
    bc = gaussian_vec(C.dim, 11)
    xg = C.solve(bc, 300, 1e-12)
    xo, ito = O.coarse_cg(Co.op(), bc, 300, 1e-12)
    assert abs(C.last_iterations - ito) <= 1
    assert relerr(xg, xo) < 1e-9
    x0, bb = gaussian_vec(2 * V, 12), gaussian_vec(2 * V, 13)
    assert relerr(smg.mr_smooth(D, x0, bb, 5), O.mr_smooth(Do.op(), x0, bb, 5)) < 1e-11
    assert relerr(smg.richardson_smooth(D, x0, bb, 4, 0.1), O.richardson(Do.op(), x0, bb, 4, 0.1)) < 1e-12
    x0c = gaussian_vec(C.dim, 14)
    assert relerr(smg.mr_smooth_op(C, x0c, bc, 4), O.mr_smooth(Co.op(), x0c, bc, 4)) < 1e-11
    v0 = gaussian_vec(2 * V, 15)
    assert abs(smg.estimate_lambda_max(D, v0, 30) - O.lambda_max(Do.op(), v0, 30)) < 1e-11


def _mg_pair(smg, L, U, levels, blocks, k, null_iters, seed, coarse_tol=1e-12, coarse_maxit=200):
    V = L * L
    dims = [2 * V]
    nx, nd = L, 2
    for (bx, by), kk in zip(blocks[: levels - 1], k):
        dims.append((nx // bx) * (nx // by) * kk)
        nx = nx // bx
    starts = np.concatenate([gaussian_vecs(kk, n, seed + i).ravel() for i, (kk, n) in enumerate(zip(k, dims))])
    D, Do = smg.DiracOp(L, U, 0.05, 1.0, 1.0, 0.0), O.Dirac(L, U, 0.05, 1.0, 1.0, 0.0)
    mg = smg.MGHierarchy(D, starts, n_levels=levels, blocks=blocks, k=k, null_iters=null_iters,
                         coarse_maxit=coarse_maxit, coarse_tol=coarse_tol)
    A = Do.op()
    mgo = O.MG(A, L, L, 2, starts, n_levels=levels, block=blocks, k=k, null_iters=null_iters, coarse_solver=0,
               coarse_maxit=coarse_maxit, coarse_tol=coarse_tol)
    return D, Do, A, mg, mgo


@pytest.mark.parametrize("levels,L,blocks", [(2, 16, ((4, 4),)), (3, 32, ((4, 4), (2, 2)))])
def test_mg_cycle_and_solvers(smg, levels, L, blocks):
    banner(f"mg cycle levels={levels} L={L}")
    k = (4, 4)
    V = L * L
    U = random_gauge(L, 0.4, 20 + L)
    D, Do, A, mg, mgo = _mg_pair(smg, L, U, levels, blocks, k, 10, 30, coarse_tol=1e-14, coarse_maxit=2000)
    # setup: the null vectors and Galerkin links follow the oracle's
    assert relerr(mg.null_vecs(0), mgo_null(mgo, k[0], 2 * V)) < 1e-8
    b = gaussian_vec(2 * V, 40)
    assert relerr(mg.precondition(b), mgo(b)) < 1e-8
    x32 = mg.precondition(b, smg.F32)
    assert relerr(x32, mgo(b)) < 1e-3
    # CG (no preconditioner): same iteration count and solution
    r = smg.cg_solve(D, b, 1000, 1e-10)
    xo, ito, _ = O.cg(A, b, 1000, 1e-10)
    assert abs(r.iterations - ito) <= 2 and relerr(r.solution, xo) < 1e-8
    assert r.true_residual < 1.1e-10
    # GCR with the MG preconditioner (textbook projection) to 1e-10
    r = smg.fcg_solve(D, b, mg, 100, 1e-10, 20)
    xo, ito, _ = O.gcr(A, b, 100, 1e-10, 20, mgo, ref_projection=False)
    assert abs(r.iterations - ito) <= 1 and r.true_residual < 1.2e-10
    assert relerr(r.solution, xo) < 1e-8
    # the reference's conjugated projection
    r = smg.fcg_solve(D, b, mg, 12, 1e-10, 20, ref_projection=True)
    xo, ito, _ = O.gcr(A, b, 12, 1e-10, 20, mgo, ref_projection=True)
    assert r.iterations == ito and relerr(r.solution, xo) < 1e-7
    # FGMRES, with the reference's default harvest of 4 Ritz pairs and without
    r = smg.fgmres_solve(D, b, mg, 20, 100, 1e-10)
    xo, ito, _, vals, _ = O.fgmres_ritz(A, b, 20, 100, 1e-10, 4, mgo)
    assert abs(r.iterations - ito) <= 1 and relerr(r.solution, xo) < 1e-8
    assert len(r.ritz_pairs) == len(vals)
    r = smg.fgmres_solve(D, b, mg, 20, 100, 1e-10, n_ritz_harvest=0)
    xo, ito, _ = O.fgmres(A, b, 20, 100, 1e-10, mgo)
    assert abs(r.iterations - ito) <= 1 and relerr(r.solution, xo) < 1e-8 and r.ritz_pairs == []
    # mixed precision: FP32 GCR + FP32 cycle inside FP64 refinement reaches 1e-10
    r = smg.fcg_solve(D, b, mg, 200, 1e-10, 10, prec=smg.F32, inner_tol=1e-5)
    assert r.true_residual < 1e-10
    assert relerr(r.solution, xo) < 1e-8


def mgo_null(mgo, k, n):
    import oracle
    out = np.zeros(k * n, np.complex128)
    oracle.lib().orc_mg_null_vecs(mgo.h, 0, out.ctypes.data_as(oracle._P))
    return out.reshape(k, n)


def test_symmetric_mg_pcg(smg):
    banner("symmetric mg pcg L=16")
    L, k = 16, (4,)
    V = L * L
    U = random_gauge(L, 0.4, 77)
    D, Do, A, mg, mgo = _mg_pair(smg, L, U, 2, ((4, 4),), k, 10, 78, coarse_tol=1e-14, coarse_maxit=2000)
    sym = gaussian_vec(2 * V, 79)
    mg.set_symmetric(0.8, sym)
    mgo.set_symmetric(0.8, sym)
    assert abs(mg.omega(0) - mgo.omega(0)) < 1e-10 * mgo.omega(0)
    b = gaussian_vec(2 * V, 80)
    r = smg.cg_solve(D, b, 200, 1e-10, mg=mg)
    xo, ito, _ = O.cg(A, b, 200, 1e-10, mg=mgo)
    assert abs(r.iterations - ito) <= 1 and relerr(r.solution, xo) < 1e-8


def test_golden_multigrid(smg):
    banner("mg golden L=16 two-level")
    g = np.load(GOLD)
    L, k = 16, 4
    D = smg.DiracOp(L, g["L16_U"], 0.05, 1.0, 1.0, 0.0)
    mg = smg.MGHierarchy(D, g["L16_starts"], n_levels=2, blocks=((4, 4),), k=(k,), null_iters=20,
                         coarse_maxit=5000, coarse_tol=1e-15)
    # the reference solves the coarsest level by dense LU; a CG to 1e-15 reproduces it
    assert relerr(mg.precondition(g["L16_rhs"]), g["L16_mg2_mode0_x"]) < 1e-8
    r = smg.fgmres_solve(D, g["L16_rhs"], mg, 20, 100, 1e-10, n_ritz_harvest=0)
    assert abs(r.iterations - g["L16_mg2_mode3_it"][0]) <= 1
    assert relerr(r.solution, g["L16_mg2_mode3_x"]) < 1e-8


def test_schur_cg(smg):
    banner("cg on M^dag M L=16")
    L = 16
    V = L * L
    U = random_gauge(L, 0.4, 90)
    D, Do = smg.DiracOp(L, U, 0.1, 1.0, 1.0, 0.05), O.Dirac(L, U, 0.1, 1.0, 1.0, 0.05)
    b = gaussian_vec(V, 91)
    r = smg.cg_solve(D, b, 1000, 1e-10, op=smg.OP_SCHUR)
    xo, ito, _ = O.cg(Do.op(O.SCHUR_NORMAL), b, 1000, 1e-10)
    assert abs(r.iterations - ito) <= 2 and relerr(r.solution, xo) < 1e-8


def test_dense_coarsest_matches_reference_lu(smg):
    """The dense inverse coarsest solve reproduces the reference's dense-LU cycle."""
    banner("mg dense coarsest vs golden L=16")
    g = np.load(GOLD)
    L, k = 16, 4
    D = smg.DiracOp(L, g["L16_U"], 0.05, 1.0, 1.0, 0.0)
    mg = smg.MGHierarchy(D, g["L16_starts"], n_levels=2, blocks=((4, 4),), k=(k,), null_iters=20, coarse_solver=1)
    assert relerr(mg.precondition(g["L16_rhs"]), g["L16_mg2_mode0_x"]) < 1e-10
    L = 32
    D3 = smg.DiracOp(L, g["L32_U"], 0.05, 1.0, 1.0, 0.0)
    mg3 = smg.MGHierarchy(D3, g["L32_starts"], n_levels=3, blocks=((4, 4), (4, 1)), k=(k, k), null_iters=10,
                          coarse_solver=1)
    assert relerr(mg3.precondition(g["L32_rhs"]), g["L32_mg3_x"]) < 1e-10


@pytest.mark.parametrize("coarse_solver", [0, 1])
def test_graph_replay_equals_eager(smg, coarse_solver):
    """A captured cycle replays to the eager result, on repeated replays with other
    library work (and host synchronisation) between them."""
    banner(f"graph replay coarse_solver={coarse_solver}")
    L, k = 32, (4, 4)
    V = L * L
    U = random_gauge(L, 0.4, 55)
    dims = [2 * V, (L // 4) ** 2 * 4]
    starts = np.concatenate([gaussian_vecs(4, n, 56 + i).ravel() for i, n in enumerate(dims)])
    D = smg.DiracOp(L, U, 0.05, 1.0, 1.0, 0.0)
    mg = smg.MGHierarchy(D, starts, n_levels=3, blocks=((4, 4), (2, 2)), k=k, null_iters=5, coarse_maxit=20,
                         coarse_tol=1e-3, coarse_solver=coarse_solver)
    bs = [gaussian_vec(2 * V, 60 + i) for i in range(3)]
    for prec, tol in ((smg.F64, 0.0), (smg.F32, 0.0)):
        mg.set_cycle(3, 3, 20, 1e-3, use_graph=False)
        eager = [mg.precondition(b, prec) for b in bs]
        mg.set_cycle(3, 3, 20, 1e-3, use_graph=True)
        for rep in range(3):
            for b, want in zip(bs, eager):
                got = mg.precondition(b, prec)
                D.apply(b)            # unrelated eager work between replays
                assert relerr(got, want) <= tol + 1e-14, (prec, rep)


def test_coarse_apply_strided_large_lattice(smg):
    """L = 128, 4 x 4 aggregates, k = 32 (16 null vectors, chiral): 32 x 32 coarse sites
    of 32 x 32 link entries per direction, 1,048,576 lanes against the 2048 x 256 of one
    pass, so the single-RHS apply's workgroups take two strides and its fused dot
    products (the coarse CG's) finish over the rows each workgroup accumulated."""
    banner("strided single-RHS coarse apply L=128 b=4 k=32")
    L, b, k = 128, 4, 32
    U = random_gauge(L, 0.4, 90)
    Po = O.Prolongator(L, L, 2, b, b, k, chiral=True).build(gaussian_vecs(k // 2, 2 * L * L, 91).ravel())
    Co = O.Coarse(O.Dirac(L, U, 0.05, 1.0, 1.0, 0.0).op(), Po, 1)
    C = smg.SparseCoarseOp.from_links(L // b, L // b, k, Co.links())
    assert C.dim * k > 2048 * 256
    A = Co.op()
    c = gaussian_vec(C.dim, 92)
    want = A(c)
    assert relerr(C.apply_to(c, smg.F64), want) < 1e-12
    assert relerr(C.apply_to(c, smg.F32), want) < 1e-5
    assert relerr(C.apply_to(c, smg.F16), want) < 5e-3
    x = C.solve(c, 300, 1e-10)
    xo, ito = O.coarse_cg(A, c, 300, 1e-10)
    print(f"coarse CG: {C.last_iterations} its (oracle {ito}), relerr {relerr(x, xo):.2e}")
    assert abs(C.last_iterations - ito) <= 1
    assert relerr(x, xo) < 1e-8
    # both stop at the iteration cap on this ill-conditioned operator: equal residuals
    rg, ro = (np.linalg.norm(c - A(v)) / np.linalg.norm(c) for v in (x, xo))
    assert abs(rg - ro) < 1e-3 * ro
