"""The many-right-hand-side path: batched stencil, batched multigrid cycle and
batched GCR agree with the single-right-hand-side path (which is pinned to the
oracle and the reference), and every right-hand side reaches the 1e-10 FP64
true residual in FP64 and in mixed precision."""
import numpy as np
import pytest

import oracle as O
from helpers import banner, relerr
from schwinger_mg_amd.fields import gaussian_vec, gaussian_vecs, random_gauge

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N", [1, 4, 32])
def test_batched_stencil(smg, N):
    import torch
    banner(f"batched stencil N={N}")
    L, V = 32, 1024
    U = random_gauge(L, 0.5, 3)
    D = smg.DiracOp(L, U, 0.1, 1.0, 1.0, 0.07)
    Do = O.Dirac(L, U, 0.1, 1.0, 1.0, 0.07)
    B = gaussian_vecs(N, 2 * V, 4)
    Bb = torch.from_numpy(np.ascontiguousarray(B.T)).cuda()        # [i][j]
    for kind in range(4):
        want = np.stack([Do.apply(kind, B[j]) for j in range(N)])
        for prec, tol, dt in ((smg.F64, 1e-12, torch.complex128), (smg.F32, 2e-6, torch.complex64)):
            src = Bb.to(dt).contiguous()
            dst = torch.empty_like(src)
            smg._chk(smg.lib().smg_dirac_apply_batched_dev(D.h, kind, prec, src.data_ptr(), dst.data_ptr(), N,
                                                           torch.cuda.current_stream().cuda_stream))
            torch.cuda.synchronize()
            got = dst.cpu().numpy().astype(np.complex128).T
            assert relerr(got, want) < tol, (kind, prec)


@pytest.mark.parametrize("k", [8, 16, 32])
@pytest.mark.parametrize("N", [4, 16, 32, 64])
def test_batched_coarse_apply(smg, k, N):
    """Galerkin links from the oracle (chiral aggregation for k = 32); N = 16, 32, 64
    with k = 8, 16, 32 run on the matrix cores, N = 4 on the plain kernel."""
    banner(f"batched coarse apply k={k} N={N}")
    L, b = 32, 4
    U = random_gauge(L, 0.4, 40 + k)
    chiral = k == 32
    nv = k // 2 if chiral else k
    Po = O.Prolongator(L, L, 2, b, b, k, chiral=chiral).build(gaussian_vecs(nv, 2 * L * L, 41).ravel())
    Co = O.Coarse(O.Dirac(L, U, 0.05, 1.0, 1.0, 0.0).op(), Po, 1)
    C = smg.SparseCoarseOp.from_links(L // b, L // b, k, Co.links())
    S = gaussian_vecs(N, C.dim, 42)
    A = Co.op()
    want = np.stack([A(S[j]) for j in range(N)])
    assert relerr(C.apply_batched(S, smg.F64), want) < 1e-12
    assert relerr(C.apply_batched(S, smg.F32), want) < 2e-6
    # FP16 links; with k = 8, 16, 32 and N >= 16 the batched product runs on the bf16
    # matrix cores with the FP32 sources split into two bf16 terms (about 2^-16)
    want16 = np.stack([C.apply_to(S[j], smg.F16) for j in range(N)])
    err16 = relerr(C.apply_batched(S, smg.F16), want16)
    print(f"k={k} N={N}: FP16-link batched vs single relerr {err16:.2e}, single vs FP64 {relerr(want16, want):.2e}")
    assert err16 < (1e-4 if N >= 16 else 2e-6)


@pytest.mark.parametrize("N", [16, 32])
def test_batched_coarse_apply_rectangular(smg, N):
    """Random links on a 6 x 4 coarse lattice (and 2 x 5, where +x and -x meet the
    same neighbour): the batched apply equals the single-RHS apply per RHS."""
    banner(f"batched coarse apply rectangular N={N}")
    k = 8
    for nx, ny in ((6, 4), (2, 5)):
        links = gaussian_vec(nx * ny * 9 * k * k, 43 + nx)
        C = smg.SparseCoarseOp.from_links(nx, ny, k, links)
        S = gaussian_vecs(N, C.dim, 44)
        want = np.stack([C.apply_to(S[j], smg.F64) for j in range(N)])
        assert relerr(C.apply_batched(S, smg.F64), want) < 1e-13, (nx, ny)


def _mg(smg, L, coarse_solver, seed=10, k=4, chiral=False, w_cycle=True):
    V = L * L
    U = random_gauge(L, 0.4, seed)
    D = smg.DiracOp(L, U, 0.05, 1.0, 1.0, 0.0)
    dims = [2 * V, (L // 4) ** 2 * k]
    nv = k // 2 if chiral else k
    starts = np.concatenate([gaussian_vecs(nv, n, seed + 1 + i).ravel() for i, n in enumerate(dims)])
    mg = smg.MGHierarchy(D, starts, n_levels=3, blocks=((4, 4), (2, 2)), k=(k, k), null_iters=5, coarse_maxit=30,
                         coarse_tol=1e-4, coarse_solver=coarse_solver, coarse_half=True, chiral=chiral,
                         w_cycle=w_cycle)
    return D, U, mg


@pytest.mark.parametrize("coarse_solver,k,N,chiral", [(0, 4, 8, False), (1, 4, 8, False), (1, 8, 32, False),
                                                      (0, 8, 16, False), (1, 16, 16, False), (0, 32, 32, True)])
def test_batched_cycle_equals_single(smg, coarse_solver, k, N, chiral):
    """k = 8, 16, 32 with N >= 16 runs the coarse levels and the dense coarsest
    product on the matrix cores; k = 32 with chiral aggregation."""
    banner(f"batched cycle coarse_solver={coarse_solver} k={k} N={N} chiral={chiral}")
    L = 32
    D, U, mg = _mg(smg, L, coarse_solver, k=k, chiral=chiral, w_cycle=not chiral)
    B = gaussian_vecs(N, 2 * L * L, 20)
    want64 = np.stack([mg.precondition(B[j], smg.F64) for j in range(N)])
    want32 = np.stack([mg.precondition(B[j], smg.F32) for j in range(N)])
    assert relerr(mg.precondition_batched(B, smg.F64), want64) < 1e-10
    # both FP32 cycles keep FP16 coarse links; for N >= 16 the batched
    # coarse products run on the bf16 matrix cores (sources split in two bf16 terms)
    err32 = relerr(mg.precondition_batched(B, smg.F32), want32)
    print(f"batched vs single FP32 cycle relerr {err32:.2e}, FP32 vs FP64 cycle {relerr(want32, want64):.2e}")
    assert err32 < (5e-4 if N >= 16 else 2e-5)


@pytest.mark.parametrize("prec,k,N,chiral", [("f64", 4, 8, False), ("mixed", 4, 8, False), ("f64", 8, 32, False),
                                              ("mixed", 8, 32, False), ("f64", 32, 32, True),
                                              ("mixed", 32, 32, True)])
def test_batched_gcr_solves(smg, prec, k, N, chiral):
    banner(f"batched gcr {prec} k={k} N={N} chiral={chiral}")
    L = 32
    D, U, mg = _mg(smg, L, 0 if chiral else 1, k=k, chiral=chiral, w_cycle=not chiral)
    B = gaussian_vecs(N, 2 * L * L, 30)
    p = smg.F64 if prec == "f64" else smg.F32
    X, info, tr = smg.gcr_solve_batched(D, B, mg, 200, 1e-10, 8, prec=p, inner_tol=1e-5)
    assert (tr < 1e-10).all(), tr
    A = O.Dirac(L, U, 0.05, 1.0, 1.0, 0.0).op()
    for j in (0, N - 1):
        assert np.linalg.norm(B[j] - A(X[j])) / np.linalg.norm(B[j]) < 1e-10
        r = smg.fcg_solve(D, B[j], mg, 200, 1e-12, 8)
        assert relerr(X[j], r.solution) < 1e-8


def _bf16(x):
    """Round each real component to BF16 (nearest even), as the device stores it."""
    f = np.ascontiguousarray(x.view(np.float64).astype(np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    r = (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16).astype(np.uint32)
    return r.view(np.float32).astype(np.float64).view(np.complex128).reshape(x.shape)


@pytest.mark.parametrize("k,N", [(8, 16), (16, 32), (32, 32), (32, 64)])
def test_coarse_apply_bf16_vectors(smg, k, N):
    """FP16 links with BF16 sources and results (the storage of the coarsest CG with
    coarse_half = 2): each result is the BF16 rounding of the FP16-link product of
    the BF16-rounded sources, to the nearest BF16 or its neighbour."""
    banner(f"coarse apply BF16 vectors k={k} N={N}")
    L, b = 32, 4
    U = random_gauge(L, 0.4, 60 + k)
    chiral = k == 32
    nv = k // 2 if chiral else k
    Po = O.Prolongator(L, L, 2, b, b, k, chiral=chiral).build(gaussian_vecs(nv, 2 * L * L, 61).ravel())
    Co = O.Coarse(O.Dirac(L, U, 0.05, 1.0, 1.0, 0.0).op(), Po, 1)
    C = smg.SparseCoarseOp.from_links(L // b, L // b, k, Co.links())
    S = gaussian_vecs(N, C.dim, 62)
    Sr = _bf16(S)
    exact = np.stack([C.apply_to(Sr[j], smg.F16) for j in range(N)])   # FP16 links, FP32 arithmetic
    got = C.apply_batched(S, smg.BF16V)
    assert np.array_equal(got, _bf16(got))   # every result is a BF16 value
    want = _bf16(exact)
    g, w, e = got.view(np.float64), want.view(np.float64), exact.view(np.float64)
    same = np.mean(g == w)
    ulp = np.maximum(np.abs(e), np.abs(g)) * 2.0 ** -7   # one BF16 step at the value (upper bound)
    print(f"k={k} N={N}: {100 * same:.2f}% bit-equal to the rounded product")
    assert same > 0.99 and np.all(np.abs(g - w) <= ulp + 1e-30)


@pytest.mark.parametrize("w_cycle", [False, True])
def test_mixed_solve_with_bf16_coarsest_cg(smg, w_cycle):
    """coarse_half = 2 (BF16 vectors in the coarsest CG of the batched FP32 cycle):
    the mixed batched solve reaches the same 1e-10 true residual as with FP32
    vectors, in about as many iterations, and the same solution."""
    banner(f"mixed solve, BF16 coarsest CG, w_cycle={w_cycle}")
    L, k, N = 32, 32, 32
    V = L * L
    U = random_gauge(L, 0.4, 70)
    D = smg.DiracOp(L, U, 0.05, 1.0, 1.0, 0.0)
    dims = [2 * V, (L // 4) ** 2 * k]
    starts = np.concatenate([gaussian_vecs(k // 2, n, 71 + i).ravel() for i, n in enumerate(dims)])
    B = gaussian_vecs(N, 2 * V, 72)
    out = {}
    for half in (1, 2):
        mg = smg.MGHierarchy(D, starts, n_levels=3, blocks=((4, 4), (2, 2)), k=(k, k), null_iters=5,
                             coarse_maxit=30, coarse_tol=0.1, coarse_solver=0, coarse_half=half, chiral=True,
                             w_cycle=w_cycle)
        X, info, tr = smg.gcr_solve_batched(D, B, mg, 200, 1e-10, 10, prec=smg.F32, inner_tol=1e-5)
        assert (tr < 1e-10).all(), tr
        out[half] = (X, info.iterations)
    print(f"iterations: FP32 coarsest CG {out[1][1]}, BF16 coarsest CG {out[2][1]}")
    assert out[2][1] <= out[1][1] + 3
    assert relerr(out[2][0], out[1][0]) < 1e-8


@pytest.mark.parametrize("half", [1, 2])
def test_batched_cycle_graph_replays_equal_eager(smg, half):
    """The batched cycle captured as a hipGraph (use_graph): three replays per precision,
    with eager work between them, equal the eager cycle bit for bit (FP32 with FP16 links
    and, for half = 2, BF16 coarsest-CG vectors; FP64), and a batched mixed solve through
    the graph equals the eager solve."""
    banner(f"batched cycle hipGraph replays, coarse_half={half}")
    L, k, N = 32, 32, 32
    V = L * L
    U = random_gauge(L, 0.4, 90)
    D = smg.DiracOp(L, U, 0.05, 1.0, 1.0, 0.0)
    dims = [2 * V, (L // 4) ** 2 * k]
    starts = np.concatenate([gaussian_vecs(k // 2, n, 91 + i).ravel() for i, n in enumerate(dims)])
    mg = smg.MGHierarchy(D, starts, n_levels=3, blocks=((4, 4), (2, 2)), k=(k, k), null_iters=5, coarse_maxit=30,
                         coarse_tol=0.1, coarse_half=half, chiral=True, w_cycle=False)
    B = gaussian_vecs(N, 2 * V, 92)
    eager = {p: mg.precondition_batched(B, p) for p in (smg.F32, smg.F64)}
    Xe, ie, te = smg.gcr_solve_batched(D, B, mg, 200, 1e-10, 10, prec=smg.F32)
    mg.set_cycle(3, 3, 30, 0.1, use_graph=True)
    for rep in range(3):
        for p in (smg.F32, smg.F64):
            assert np.array_equal(mg.precondition_batched(B, p), eager[p]), (rep, p)
        D.apply(B[rep])     # eager work between the replays
    Xg, ig, tg = smg.gcr_solve_batched(D, B, mg, 200, 1e-10, 10, prec=smg.F32)
    assert np.array_equal(Xg, Xe) and ig.iterations == ie.iterations
    assert (tg < 1e-10).all()


@pytest.mark.parametrize("L", [32, 48])
def test_fused_ddagd_equals_two_launches(smg, L):
    """D^dag D of the batched path in one launch (the tile's D x kept in LDS) equals the
    two stencil launches entry for entry (the same arithmetic per entry), with ordinary
    and non-temporal stores, in FP32 and FP64, and the oracle's D^dag D."""
    import torch
    banner(f"fused D^dag D L={L}")
    V = L * L
    U = random_gauge(L, 0.5, 500 + L)
    D = smg.DiracOp(L, U, 0.1, 1.0, 1.0, 0.07)
    Do = O.Dirac(L, U, 0.1, 1.0, 1.0, 0.07)
    try:
        for N, prec, dt, nbytes, tol in ((16, smg.F32, torch.complex64, 4, 2e-6), (32, smg.F32, torch.complex64, 4, 2e-6),
                                         (8, smg.F64, torch.complex128, 8, 1e-12), (32, smg.F64, torch.complex128, 8, 1e-12)):
            B = gaussian_vecs(N, 2 * V, 501)
            src = torch.from_numpy(np.ascontiguousarray(B.T)).cuda().to(dt).contiguous()
            want = np.stack([Do.apply(2, B[j]) for j in range(N)])
            got = {}
            for choice in (0, 1, 2):   # two launches; fused; fused with non-temporal stores
                smg.tune_set(5, (L, N, nbytes, 0, 0), choice)
                dst = torch.empty_like(src)
                smg._chk(smg.lib().smg_dirac_apply_batched_dev(D.h, 2, prec, src.data_ptr(), dst.data_ptr(), N,
                                                               torch.cuda.current_stream().cuda_stream))
                torch.cuda.synchronize()
                got[choice] = dst.cpu().numpy().astype(np.complex128).T
                assert relerr(got[choice], want) < tol, (N, prec, choice)
            same = np.mean(got[1] == got[0])
            print(f"N={N} prec={prec}: fused equals the two launches on {100 * same:.3f}% of entries, "
                  f"max relative difference {relerr(got[1], got[0]):.1e}")
            assert np.array_equal(got[1], got[2])
            assert relerr(got[1], got[0]) < (1e-6 if prec == smg.F32 else 1e-14)
    finally:
        smg.tune_clear()


def test_fused_ddagd_cycle_and_solve(smg):
    """The batched cycle and the mixed and FP64 batched solves with the fused D^dag D
    (every epilogue: plain, residual, dot products) against the two-launch form: the
    same iterations and the 1e-10 true residual, the cycle to roundoff.  The coarsest
    solve is the dense inverse, so the cycle is a smooth function of the rounding (a
    coarsest CG's stopping test, in BF16 above all, turns last-bit differences into
    1e-3 ones)."""
    banner("fused D^dag D in the batched cycle and solves")
    L, k, N = 32, 32, 32
    V = L * L
    U = random_gauge(L, 0.4, 510)
    D = smg.DiracOp(L, U, 0.05, 1.0, 1.0, 0.0)
    dims = [2 * V, (L // 4) ** 2 * k]
    starts = np.concatenate([gaussian_vecs(k // 2, n, 511 + i).ravel() for i, n in enumerate(dims)])
    mg = smg.MGHierarchy(D, starts, n_levels=3, blocks=((4, 4), (2, 2)), k=(k, k), null_iters=5, coarse_maxit=30,
                         coarse_tol=0.1, coarse_half=1, coarse_solver=1, chiral=True, w_cycle=False)
    B = gaussian_vecs(N, 2 * V, 512)
    A = O.Dirac(L, U, 0.05, 1.0, 1.0, 0.0).op()
    out = {}
    try:
        for choice in (0, 1):
            for nbytes in (4, 8):
                for epi in (0, 1, 2):
                    smg.tune_set(5, (L, N, nbytes, epi, 0), choice)
            mg.set_cycle(3, 3, 30, 0.1)   # drop the captured cycles
            cyc = {p: mg.precondition_batched(B, p) for p in (smg.F32, smg.F64)}
            sol = {}
            for p in (smg.F32, smg.F64):
                X, info, tr = smg.gcr_solve_batched(D, B, mg, 200, 1e-10, 10, prec=p)
                o = max(np.linalg.norm(B[j] - A(X[j])) / np.linalg.norm(B[j]) for j in range(N))
                assert o < 1e-10, (choice, p, o)
                sol[p] = (X, info.iterations)
            out[choice] = (cyc, sol)
    finally:
        smg.tune_clear()
    (c0, s0), (c1, s1) = out[0], out[1]
    print(f"cycle FP32 {relerr(c1[smg.F32], c0[smg.F32]):.1e}, FP64 {relerr(c1[smg.F64], c0[smg.F64]):.1e}; "
          f"iterations {s0[smg.F32][1]}/{s1[smg.F32][1]} (mixed), {s0[smg.F64][1]}/{s1[smg.F64][1]} (FP64)")
    assert relerr(c1[smg.F32], c0[smg.F32]) < 1e-5 and relerr(c1[smg.F64], c0[smg.F64]) < 1e-12
    assert s1[smg.F32][1] == s0[smg.F32][1] and s1[smg.F64][1] == s0[smg.F64][1]
    assert relerr(s1[smg.F64][0], s0[smg.F64][0]) < 1e-9
