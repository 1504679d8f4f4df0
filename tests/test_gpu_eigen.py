"""Eigensolvers and deflation on the GPU, through the C-ABI, against the oracle's
restatements (lanczos_eigen, trlm_eigensolver, SparseCoarseOp::setup_deflation +
solve, cg_solve_deflated) that the CPU suite pins to the reference's own outputs,
and against the golden fixtures the reference wrote.  The device TRLM orthogonalises
by two classical Gram-Schmidt passes where the reference runs modified Gram-Schmidt
twice; both are exact in exact arithmetic, so eigenvalues agree to rounding
(relative 1e-10), eigenvectors up to a phase, restart counts exactly."""
import os

import numpy as np
import pytest

import oracle as O
from helpers import banner, relerr
from schwinger_mg_amd.fields import gaussian_vec, gaussian_vecs, random_gauge

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_vectors.npz")


def _aligned_err(a, b):
    """|b e^{i phi} - a| / |a| for the phase phi that aligns b with a."""
    ov = np.vdot(b, a)
    return np.linalg.norm(b * (ov / abs(ov)) - a) / np.linalg.norm(a)


def test_random_vec_matches_reference(smg):
    banner("random_vec")
    g = np.load(GOLD)
    assert np.array_equal(smg.random_vec(128, 12345), g["L8_pv12345"])
    assert np.array_equal(smg.random_vec(128, 42), g["rv42_128"])
    assert np.array_equal(smg.random_vec(64, 42), g["rv42_64"])


def test_lanczos_eigen_vs_oracle_and_golden(smg):
    banner("lanczos_eigen")
    g = np.load(GOLD)
    for name in ("H12", "H8deg"):
        A = g[name]
        ev, vecs = smg.lanczos_eigen(np.ascontiguousarray(A.T))   # columns A_cols[j][i] = A_ij
        evo, vo = O.lanczos_eigen(A)   # both [row][col]
        assert np.abs(ev - evo).max() < 1e-13 and np.abs(ev - g[f"{name}_evals"]).max() < 1e-13
        assert np.abs(A @ vecs - vecs * ev).max() < 1e-12
        if name == "H12":
            assert np.abs(vecs - vo).max() < 1e-12
    ev, vecs = smg.lanczos_eigen(np.array([[2.5]]))
    assert ev[0] == 2.5 and vecs[0, 0] == 1.0


@pytest.mark.parametrize("L,n_ev,n_kr,csw", [(8, 6, 24, 1.0), (8, 4, 0, 1.0), (16, 8, 32, 0.0)])
def test_trlm_ddagd_vs_oracle(smg, L, n_ev, n_kr, csw):
    banner(f"trlm D^dag D L={L} n_ev={n_ev} n_kr={n_kr} csw={csw}")
    U = random_gauge(L, 0.3, 200 + L)
    D = smg.DiracOp(L, U, 0.05, 1.0, csw, 0.0)
    res = smg.trlm_eigensolver(D, n_ev, n_kr, 100, 1e-10)
    Do = O.Dirac(L, U, 0.05, 1.0, csw, 0.0)
    conv, rs, evo, vo, its = O.trlm(Do.op(), n_ev, n_kr, 100, 1e-10, smg.random_vec(2 * L * L, 42))
    print(f"restarts {res.num_restarts} vs {rs}, iterations {res.iterations} vs {its}, "
          f"evals relerr {relerr(res.eigvals, evo):.2e}")
    assert res.converged and conv and res.num_restarts == rs and res.iterations == its
    assert np.abs(res.eigvals - evo).max() < 1e-10 * evo.max()
    for i in range(n_ev):
        assert _aligned_err(vo[i], res.eigvecs[i]) < 1e-7
    # eigenpair residuals of the device vectors under the oracle operator
    A = Do.op()
    for i in range(n_ev):
        r = A(res.eigvecs[i]) - res.eigvals[i] * res.eigvecs[i]
        assert np.linalg.norm(r) < 1e-8 * res.eigvals[-1]


def test_trlm_golden(smg):
    banner("trlm golden L=8")
    g = np.load(GOLD)
    D = smg.DiracOp(8, g["E8_U"], 0.05, 1.0, 1.0, 0.0)
    for tag, (ne, nk) in (("a", (6, 24)), ("b", (4, 0))):
        res = smg.trlm_eigensolver(D, ne, nk, 100, 1e-10)
        assert res.converged and res.num_restarts == int(g[f"E8_trlm{tag}_rs"][0])
        ref = g[f"E8_trlm{tag}_evals"]
        assert np.abs(res.eigvals - ref).max() < 1e-10 * ref.max()
        for i in range(ne):
            assert _aligned_err(g[f"E8_trlm{tag}_evecs"][i], res.eigvecs[i]) < 1e-7


def test_trlm_schur_vs_oracle(smg):
    banner("trlm M^dag M L=16")
    L = 16
    U = random_gauge(L, 0.3, 230)
    D = smg.DiracOp(L, U, 0.05, 1.0, 1.0, 0.0)
    res = smg.trlm_eigensolver(D, 6, 24, 100, 1e-10, op=smg.OP_SCHUR)
    conv, rs, evo, vo, _ = O.trlm(O.Dirac(L, U, 0.05, 1.0, 1.0, 0.0).op(O.SCHUR_NORMAL), 6, 24, 100, 1e-10,
                                  smg.random_vec(L * L, 42))
    assert res.converged and res.num_restarts == rs
    assert np.abs(res.eigvals - evo).max() < 1e-10 * evo.max()


def test_trlm_unconverged_reports_it(smg):
    """max_restarts = 1 with a tight tolerance: TRLMResult::converged is false and the
    returned pairs are the current Ritz pairs, as in the reference."""
    banner("trlm unconverged")
    L = 16
    U = random_gauge(L, 0.3, 240)
    D = smg.DiracOp(L, U, 0.05, 1.0, 1.0, 0.0)
    res = smg.trlm_eigensolver(D, 8, 20, 1, 1e-14)
    conv, rs, evo, _, _ = O.trlm(O.Dirac(L, U, 0.05, 1.0, 1.0, 0.0).op(), 8, 20, 1, 1e-14,
                                 smg.random_vec(2 * L * L, 42))
    assert not res.converged and not conv and res.num_restarts == rs == 1


def _coarse16(smg, g):
    L, b, k = 16, 4, 4
    D = smg.DiracOp(L, g["L16_U"], 0.05, 1.0, 1.0, 0.0)
    P = smg.Prolongator(L, L, 2, b, b, k)
    P.build_from_vectors(g["L16_nulls"])
    return smg.SparseCoarseOp.build(P, D), D, P


def test_coarse_deflation_golden_and_oracle(smg):
    banner("coarse deflation L=16")
    g = np.load(GOLD)
    C, _, _ = _coarse16(smg, g)
    info = C.setup_deflation(8)
    assert info.converged
    ev, vecs = C.defl_vals, C.defl_vecs
    assert np.abs(ev - g["L16_defl_evals"]).max() < 1e-10 * ev.max()
    x = C.solve(g["L16_c"], 200, 1e-12)
    print(f"deflated coarse CG: {C.last_iterations} iterations")
    assert relerr(x, g["L16_defl_x"]) < 1e-10 and relerr(x, g["L16_dense_solve"]) < 1e-9
    # the oracle's deflated CG from the device's eigenpairs: same iterations
    Co = O.Coarse(O.Dirac(16, g["L16_U"], 0.05, 1.0, 1.0, 0.0).op(),
                  O.Prolongator(16, 16, 2, 4, 4, 4).build(g["L16_nulls"].ravel()), 1)
    xo, ito = O.deflated_coarse_cg(Co.op(), vecs, ev, g["L16_c"], 200, 1e-12)
    assert C.last_iterations == ito and relerr(x, xo) < 1e-10
    # set / clear the space: no deflation is the plain coarse CG from x0 = 0
    C.set_deflation([], [])
    x0 = C.solve(g["L16_c"], 200, 1e-12)
    assert len(C.defl_vals) == 0 and C.last_iterations > ito and relerr(x0, xo) < 1e-9
    C.set_deflation(ev, vecs)
    C.solve(g["L16_c"], 200, 1e-12)
    assert C.last_iterations == ito


@pytest.mark.parametrize("with_mg", [False, True])
def test_cg_deflated_vs_oracle(smg, with_mg):
    banner(f"cg_solve_deflated mg={with_mg}")
    g = np.load(GOLD)
    L = 8
    D = smg.DiracOp(L, g["E8_U"], 0.05, 1.0, 1.0, 0.0)
    vecs = g["E8_trlma_evecs"]
    if not with_mg:
        r = smg.cg_solve_deflated(D, g["E8_rhs"], None, vecs, g["E8_trlma_evals"], 500, 1e-10)
        print(f"iterations {r.iterations} vs {int(g['E8_cgdefl_it'][0])}")
        assert abs(r.iterations - int(g["E8_cgdefl_it"][0])) <= 1
        assert relerr(r.solution, g["E8_cgdefl_x"]) < 1e-9 and r.true_residual < 1e-9
        # no deflation vectors: plain CG
        r0 = smg.cg_solve_deflated(D, g["E8_rhs"], None, vecs[:0], [], 500, 1e-10)
        rc = smg.cg_solve(D, g["E8_rhs"], 500, 1e-10)
        assert r0.iterations == rc.iterations and r.iterations < r0.iterations
        return
    L, k = 16, 4
    U = random_gauge(L, 0.3, 250)
    starts = gaussian_vecs(k, 2 * L * L, 251).ravel()
    D = smg.DiracOp(L, U, 0.05, 1.0, 1.0, 0.0)
    mg = smg.MGHierarchy(D, starts, n_levels=2, blocks=((4, 4),), k=(k,), null_iters=10, coarse_solver=1)
    Do = O.Dirac(L, U, 0.05, 1.0, 1.0, 0.0)
    mgo = O.MG(Do.op(), L, L, 2, starts, n_levels=2, block=((4, 4),), k=(k,), null_iters=10, coarse_solver=1)
    tr = smg.trlm_eigensolver(D, 4, 20, 100, 1e-10)
    b = gaussian_vec(2 * L * L, 252)
    r = smg.cg_solve_deflated(D, b, mg, tr.eigvecs, tr.eigvals, 200, 1e-10)
    xo, ito, _ = O.cg_deflated(Do.op(), b, tr.eigvecs, 200, 1e-10, mg=mgo)
    print(f"MG-preconditioned deflated CG: {r.iterations} vs oracle {ito}")
    assert r.true_residual < 1e-9
    assert abs(r.iterations - ito) <= 1 and relerr(r.solution, xo) < 1e-8


def test_trlm_large_properties(smg):
    """L = 64 (n = 8192): each returned pair satisfies |A v - l v| <= tol-scale, the
    vectors are orthonormal and the eigenvalues ascend; the lowest equals the
    oracle's."""
    banner("trlm L=64 properties")
    L, n_ev = 64, 16
    U = random_gauge(L, 0.3, 260)
    D = smg.DiracOp(L, U, 0.05, 1.0, 1.0, 0.0)
    res = smg.trlm_eigensolver(D, n_ev, 0, 200, 1e-10)
    assert res.converged
    A = O.Dirac(L, U, 0.05, 1.0, 1.0, 0.0).op()
    V = res.eigvecs
    assert np.abs(V.conj() @ V.T - np.eye(n_ev)).max() < 1e-10
    assert np.all(np.diff(res.eigvals) >= 0)
    for i in range(n_ev):
        r = A(V[i]) - res.eigvals[i] * V[i]
        assert np.linalg.norm(r) < 1e-6 * res.eigvals[-1]
    conv, rs, evo, _, _ = O.trlm(A, n_ev, 0, 200, 1e-10, smg.random_vec(2 * L * L, 42))
    print(f"restarts {res.num_restarts} vs {rs}; lowest {res.eigvals[0]:.12e} vs {evo[0]:.12e}")
    assert abs(res.eigvals[0] - evo[0]) < 1e-10 * evo[-1]


# ------------------------------------------------------------------ LOBPCG
@pytest.mark.parametrize("its", [10, 40])
def test_lobpcg_golden_and_oracle(smg, its):
    """lobpcg_update on D^dag D as one 8-column block (k = 6 padded): Cholesky QR
    and block Gram matrices in place of the reference's sequential Gram-Schmidt and
    dots, equal in exact arithmetic."""
    banner(f"lobpcg golden L=8 iterations={its}")
    g = np.load(GOLD)
    D = smg.DiracOp(8, g["E8_U"], 0.05, 1.0, 1.0, 0.0)
    res = smg.lobpcg_update(D, g["E8_lob_X0"], None, its, 1e-10)
    ref = g[f"E8_lob{its}_evals"]
    print(f"evals relerr {relerr(res.eigvals, ref):.2e}, X relerr {relerr(res.eigvecs, g[f'E8_lob{its}_X']):.2e}")
    assert res.iterations == int(g[f"E8_lob{its}_it"][0])
    assert np.abs(res.eigvals - ref).max() < 1e-9 * np.abs(ref).max()
    assert relerr(res.eigvecs, g[f"E8_lob{its}_X"]) < 1e-7


@pytest.mark.parametrize("k", [4, 16])
def test_lobpcg_mg_preconditioned_vs_oracle(smg, k):
    banner(f"lobpcg with the MG preconditioner L=16 k={k}")
    L, kn = 16, 4
    U = random_gauge(L, 0.3, 270)
    starts = gaussian_vecs(kn, 2 * L * L, 271).ravel()
    D = smg.DiracOp(L, U, 0.05, 1.0, 1.0, 0.0)
    mg = smg.MGHierarchy(D, starts, n_levels=2, blocks=((4, 4),), k=(kn,), null_iters=10, coarse_solver=1)
    Do = O.Dirac(L, U, 0.05, 1.0, 1.0, 0.0)
    mgo = O.MG(Do.op(), L, L, 2, starts, n_levels=2, block=((4, 4),), k=(kn,), null_iters=10, coarse_solver=1)
    X0 = gaussian_vecs(k, 2 * L * L, 272)
    res = smg.lobpcg_update(D, X0, mg, 5, 1e-10)
    evo, Xo, ito = O.lobpcg(Do.op(), X0, 5, 1e-10, mg=mgo)
    print(f"evals {res.eigvals[:4]} vs {evo[:4]}")
    assert res.iterations == ito
    assert np.abs(res.eigvals - evo).max() < 1e-8 * np.abs(evo).max()
    assert relerr(res.eigvecs, Xo) < 1e-6


@pytest.mark.parametrize("precond", [False, True])
def test_lobpcg_exact_ax_vs_oracle(smg, precond):
    """The LOBPCG variant that keeps AX = A X, on the device against the oracle's, from
    the same block; its returned quotients are those of the returned vectors."""
    banner(f"lobpcg exact AX L=16 precond={precond}")
    L, kn, k = 16, 4, 8
    U = random_gauge(L, 0.3, 280)
    starts = gaussian_vecs(kn, 2 * L * L, 281).ravel()
    D = smg.DiracOp(L, U, 0.05, 1.0, 1.0, 0.0)
    Do = O.Dirac(L, U, 0.05, 1.0, 1.0, 0.0)
    mg = mgo = None
    if precond:
        mg = smg.MGHierarchy(D, starts, n_levels=2, blocks=((4, 4),), k=(kn,), null_iters=10, coarse_solver=1)
        mgo = O.MG(Do.op(), L, L, 2, starts, n_levels=2, block=((4, 4),), k=(kn,), null_iters=10, coarse_solver=1)
    X0 = gaussian_vecs(k, 2 * L * L, 282)
    res = smg.lobpcg_update(D, X0, mg, 6, 1e-10, exact_ax=True)
    evo, Xo, ito = O.lobpcg(Do.op(), X0, 6, 1e-10, mg=mgo, exact_ax=True)
    assert res.iterations == ito
    assert np.abs(res.eigvals - evo).max() < 1e-8 * np.abs(evo).max()
    # each vector agrees up to its phase (the Ritz vectors' phases are arbitrary)
    ph = np.array([np.vdot(g, o) for g, o in zip(res.eigvecs, Xo)])
    assert relerr(res.eigvecs * (ph / np.abs(ph))[:, None], Xo) < 1e-6
    A = Do.op()
    rq = np.array([np.vdot(x, A(x)).real for x in res.eigvecs])
    assert np.abs(res.eigvals - rq).max() < 1e-10 * rq.max()


def test_lobpcg_coarse_and_coarsest_deflation(smg):
    """lobpcg_update on a coarse operator, and MGHierarchy::update_coarsest_deflation:
    cold start from random_vec(dim) drawn n_defl times from mt19937(42), then a warm
    start from the stored space."""
    banner("lobpcg coarse / update_coarsest_deflation")
    L, kn = 16, 4
    U = random_gauge(L, 0.3, 280)
    starts = gaussian_vecs(kn, 2 * L * L, 281).ravel()
    D = smg.DiracOp(L, U, 0.05, 1.0, 1.0, 0.0)
    mg = smg.MGHierarchy(D, starts, n_levels=2, blocks=((4, 4),), k=(kn,), null_iters=10)
    dim = mg.level_dim(1)
    Po = O.Prolongator(L, L, 2, 4, 4, kn)
    Po.set_block_vecs(mg.block_vecs(0))
    Co = O.Coarse(O.Dirac(L, U, 0.05, 1.0, 1.0, 0.0).op(), Po, 1)
    n_defl = 8
    mg.update_coarsest_deflation(n_defl, 3)
    vals, vecs = mg.coarsest_deflation()
    X0 = smg.random_vec(n_defl * dim, 42).reshape(n_defl, dim)
    evo, Xo, _ = O.lobpcg(Co.op(), X0, 3, 1e-10)
    assert np.abs(vals - evo).max() < 1e-9 * np.abs(evo).max() and relerr(vecs, Xo) < 1e-7
    mg.update_coarsest_deflation(n_defl, 3)   # warm start
    vals2, vecs2 = mg.coarsest_deflation()
    evo2, Xo2, _ = O.lobpcg(Co.op(), Xo, 3, 1e-10)
    assert np.abs(vals2 - evo2).max() < 1e-9 * np.abs(evo2).max() and relerr(vecs2, Xo2) < 1e-6
    # the coarse operator directly
    P = smg.Prolongator(L, L, 2, 4, 4, kn)
    P.set_block_vecs(mg.block_vecs(0))
    C = smg.SparseCoarseOp.build(P, D)
    res = smg.lobpcg_update(C, X0, None, 3, 1e-10)
    assert np.abs(res.eigvals - evo).max() < 1e-9 * np.abs(evo).max()


def test_lobpcg_large_block_of_16(smg):
    """L = 64 (n = 8192), a 16-vector block: X stays orthonormal, its Ritz values lie
    above the 16 lowest eigenvalues (Cauchy interlacing), and the Rayleigh quotients
    follow the oracle's from the same start over 30 iterations."""
    banner("lobpcg L=64 k=16")
    L, k = 64, 16
    U = random_gauge(L, 0.3, 290)
    D = smg.DiracOp(L, U, 0.05, 1.0, 1.0, 0.0)
    tr = smg.trlm_eigensolver(D, k, 0, 200, 1e-10)
    X0 = gaussian_vecs(k, 2 * L * L, 291)
    res = smg.lobpcg_update(D, X0, None, 30, 1e-10)
    X = res.eigvecs
    assert np.abs(X.conj() @ X.T - np.eye(k)).max() < 1e-10
    A = O.Dirac(L, U, 0.05, 1.0, 1.0, 0.0).op()
    AX = np.array([A(x) for x in X])
    ritz = np.linalg.eigvalsh(X.conj() @ AX.T)
    print(f"ritz[:4] {ritz[:4]}, trlm[:4] {tr.eigvals[:4]}")
    assert np.all(ritz >= tr.eigvals - 1e-10 * tr.eigvals[-1])
    evo, _, ito = O.lobpcg(A, X0, 30, 1e-10)
    assert res.iterations == ito and np.abs(res.eigvals - evo).max() < 1e-7 * np.abs(evo).max()
