"""The Hermitian link reads of the FP16-link matrix-core coarse apply and the launch-shape
autotuner.  A Galerkin operator's FP16 links are made exactly Hermitian, so the apply that
reads 5 of a site's 9 link blocks returns, bit for bit, what the 9-block apply returns, and
both agree with the oracle's FP64 operator to the FP16 rounding.  A tuned table, saved and
loaded again, reproduces a batched mixed-precision solve bit for bit."""
import numpy as np
import pytest

import oracle as O
from helpers import banner, relerr
from schwinger_mg_amd.fields import gaussian_vecs, random_gauge

pytestmark = pytest.mark.gpu


@pytest.fixture
def clean_table(smg):
    smg.tune_clear()
    yield
    smg.tune_clear()


@pytest.mark.parametrize("k,N", [(16, 16), (32, 32), (32, 64)])
def test_hermitian_reads_bit_identical(smg, clean_table, k, N):
    banner(f"hermitian coarse reads k={k} N={N}")
    L, b = 32, 4
    V = L * L
    U = random_gauge(L, 0.4, 300 + k)
    D = smg.DiracOp(L, U, 0.05, 1.0, 1.0, 0.0)
    nulls = gaussian_vecs(k // 2, 2 * V, 301)
    P = smg.Prolongator(L, L, 2, b, b, k, chiral=True).build_from_vectors(nulls)
    C = smg.SparseCoarseOp.build(P, D)            # Galerkin: Hermitian, 8 x 8 coarse sites
    S = gaussian_vecs(N, C.dim, 302)
    nb = L // b
    out = {}
    for prec in (smg.F16, smg.BF16V):
        hv = 4 if prec == smg.BF16V else 0
        for herm in (0, 1):
            smg.tune_set(smg.TUNE_COARSE_BF16, [nb, nb, k, N, hv + 0], herm)
            out[prec, herm] = C.apply_batched(S, prec)
        assert np.array_equal(out[prec, 0], out[prec, 1]), prec
    Po = O.Prolongator(L, L, 2, b, b, k, chiral=True)
    Po.set_block_vecs(P.block_vecs)
    A = O.Coarse(O.Dirac(L, U, 0.05, 1.0, 1.0, 0.0).op(), Po, 1).op()
    want = np.stack([A(S[j]) for j in range(N)])
    e16, ebf = relerr(out[smg.F16, 1], want), relerr(out[smg.BF16V, 1], want)
    print(f"FP16 links vs FP64 oracle: {e16:.2e}; with BF16 vectors {ebf:.2e}")
    assert e16 < 2e-3 and ebf < 2e-2


def test_tuned_table_reproduces_solve(smg, clean_table, tmp_path):
    banner("autotuner: tune, save, load, bit-identical batched solve L=64")
    L, k, N = 64, 32, 32
    V = L * L
    U = random_gauge(L, 0.4, 310)
    D = smg.DiracOp(L, U, 0.05, 1.0, 1.0, 0.0)
    dims = [2 * V, (L // 4) ** 2 * k]
    starts = np.concatenate([gaussian_vecs(k // 2, n, 311 + i).ravel() for i, n in enumerate(dims)])
    mg = smg.MGHierarchy(D, starts, n_levels=3, blocks=((4, 4), (2, 2)), k=(k, k), null_iters=10, w_cycle=False,
                         coarse_maxit=30, coarse_tol=0.1, coarse_half=2, chiral=True)
    added = smg.tune_mg(mg, N)
    assert added >= 6
    path = str(tmp_path / "tune.txt")
    smg.tune_save(path)
    B = gaussian_vecs(N, 2 * V, 312)
    X1, i1, t1 = smg.gcr_solve_batched(D, B, mg, 500, 1e-10, trunc=10, prec=smg.F32)
    smg.tune_clear()
    smg.tune_load(path)
    X2, i2, t2 = smg.gcr_solve_batched(D, B, mg, 500, 1e-10, trunc=10, prec=smg.F32)
    print(f"tuned: {added} shapes, {i1.iterations} iterations, max true residual {t1.max():.2e}")
    print(open(path).read())
    assert np.array_equal(X1, X2) and i1.iterations == i2.iterations
    assert t1.max() < 1e-10


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_stencil_store_hints_bit_identical(smg, clean_table, prec):
    """The batched stencil with non-temporal stores (a tuner choice) writes exactly what
    the one with ordinary stores writes, for every epilogue the solvers use."""
    import torch
    banner(f"stencil store hints {prec}")
    L, N = 64, 32
    V = L * L
    U = random_gauge(L, 0.4, 320)
    D = smg.DiracOp(L, U, 0.05, 1.0, 1.0, 0.0)
    dt = torch.complex64 if prec == "f32" else torch.complex128
    p = smg.F32 if prec == "f32" else smg.F64
    g = torch.Generator().manual_seed(321)
    src = torch.randn(2 * V, N, dtype=dt, generator=g).cuda()
    stream = torch.cuda.current_stream().cuda_stream
    outs = []
    for nt in (0, 1):
        for epi in (0, 1, 2):
            smg.tune_set(smg.TUNE_DSLASH_STORE, [L, N, 4 if prec == "f32" else 8, epi, 0], nt)
        dst = torch.empty_like(src)
        for kind in range(4):
            smg._chk(smg.lib().smg_dirac_apply_batched_dev(D.h, kind, p, src.data_ptr(), dst.data_ptr(), N, stream))
            torch.cuda.synchronize()
            outs.append(dst.cpu().clone())
    for kind in range(4):
        assert torch.equal(outs[kind], outs[4 + kind]), kind


def test_stencil_store_hints_same_solve(smg, clean_table):
    """A batched mixed solve with non-temporal stencil stores in every epilogue equals,
    bit for bit, the solve with ordinary stores."""
    banner("stencil store hints: batched solve")
    L, k, N = 32, 16, 16
    V = L * L
    U = random_gauge(L, 0.4, 330)
    D = smg.DiracOp(L, U, 0.05, 1.0, 1.0, 0.0)
    starts = gaussian_vecs(k // 2, 2 * V, 331)
    mg = smg.MGHierarchy(D, starts, n_levels=2, blocks=((4, 4),), k=(k,), null_iters=10, chiral=True,
                         coarse_half=True, coarse_maxit=50, coarse_tol=1e-2)
    B = gaussian_vecs(N, 2 * V, 332)
    X = []
    for nt in (0, 1):
        for bytes_ in (4, 8):
            for epi in (0, 1, 2):
                smg.tune_set(smg.TUNE_DSLASH_STORE, [L, N, bytes_, epi, 0], nt)
        Xn, info, tr = smg.gcr_solve_batched(D, B, mg, 200, 1e-10, trunc=10, prec=smg.F32)
        assert tr.max() < 1e-10
        X.append(Xn)
    assert np.array_equal(X[0], X[1])


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_staged_transfers_same_cycle(smg, clean_table, prec):
    """The transfers with a workgroup per aggregate and its blocks staged in LDS form the
    same sums in the same order as the per-output kernels: the batched three-level cycle
    (both levels' restrictions and prolongations) agrees to roundoff, and the batched
    solve takes the same iterations to the same FP64 true residual."""
    banner(f"staged transfers: batched cycle {prec}")
    L, k, N = 32, 16, 16
    V = L * L
    U = random_gauge(L, 0.4, 340)
    D = smg.DiracOp(L, U, 0.05, 1.0, 1.0, 0.0)
    dims = [2 * V, (L // 4) ** 2 * k]
    starts = np.concatenate([gaussian_vecs(k // 2, n, 341 + i).ravel() for i, n in enumerate(dims)])
    mg = smg.MGHierarchy(D, starts, n_levels=3, blocks=((4, 4), (2, 2)), k=(k, k), null_iters=6, w_cycle=False,
                         coarse_maxit=30, coarse_tol=0.1, coarse_half=True, chiral=True)
    B = gaussian_vecs(N, 2 * V, 342)
    p = smg.F64 if prec == "f64" else smg.F32
    shapes = [(L, L, 128 * 2 + k), (L // 4, L // 4, 128 * k + k)]   # fine lattice of each transfer
    out = []
    for choice in (0, 1):
        for nx, ny, w in shapes:
            for direction in (0, 1):
                smg.tune_set(smg.TUNE_TRANSFER, [nx, ny, w, N, 2 * (prec == "f64") + direction], choice)
        X = mg.precondition_batched(B, p)
        Xs, info, tr = smg.gcr_solve_batched(D, B, mg, 200, 1e-10, trunc=10, prec=p)
        out.append((X, info.iterations, tr))
    e = relerr(out[1][0], out[0][0])
    print(f"cycle staged vs per-output: relerr {e:.2e}; iterations {out[0][1]} / {out[1][1]}")
    assert e < (1e-13 if prec == "f64" else 1e-5)
    assert out[0][1] == out[1][1] and out[1][2].max() < 1e-10



@pytest.mark.parametrize("N", [16, 32])
def test_mfma_transfers_same_cycle(smg, clean_table, N):
    """The FP32 restriction and prolongation on the matrix cores (a wave per 16 x 16 tile
    of an aggregate's chirality half; k = 32, so 16 block vectors per half) against the
    per-output kernels: the batched three-level cycle agrees to FP32 roundoff (the same
    products, summed in another order), and the mixed batched solve takes the same
    iterations to a 1e-10 FP64 true residual."""
    banner(f"matrix-core transfers: batched cycle N={N}")
    L, k = 32, 32
    V = L * L
    U = random_gauge(L, 0.4, 350)
    D = smg.DiracOp(L, U, 0.05, 1.0, 1.0, 0.0)
    dims = [2 * V, (L // 4) ** 2 * k]
    starts = np.concatenate([gaussian_vecs(k // 2, n, 351 + i).ravel() for i, n in enumerate(dims)])
    mg = smg.MGHierarchy(D, starts, n_levels=3, blocks=((4, 4), (2, 2)), k=(k, k), null_iters=6, w_cycle=False,
                         coarse_maxit=30, coarse_tol=0.1, coarse_half=True, chiral=True)
    B = gaussian_vecs(N, 2 * V, 352)
    shapes = [(L, L, 128 * 2 + k), (L // 4, L // 4, 128 * k + k)]
    out = []
    for choice in (0, 2):
        for nx, ny, w in shapes:
            for direction in (0, 1):
                smg.tune_set(smg.TUNE_TRANSFER, [nx, ny, w, N, direction], choice)
        X = mg.precondition_batched(B, smg.F32)
        Xs, info, tr = smg.gcr_solve_batched(D, B, mg, 200, 1e-10, trunc=10, prec=smg.F32)
        out.append((X, info.iterations, tr))
    e = relerr(out[1][0], out[0][0])
    print(f"cycle matrix-core vs per-output transfers: relerr {e:.2e}; iterations {out[0][1]} / {out[1][1]}")
    assert e < 1e-5
    assert abs(out[0][1] - out[1][1]) <= 1 and out[1][2].max() < 1e-10
