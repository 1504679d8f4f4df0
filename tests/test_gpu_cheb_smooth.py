"""The Chebyshev smoother of the batched cycles (set_cheb_smoother): the FP32 and FP64
cycles against a numpy cycle built from the hierarchy's own block vectors and Galerkin links
with the oracle's fine operator and the oracle's Chebyshev restatement; the one-launch
steps (fine D^dag D and FP16-link coarse apply with the updates in their epilogues)
against the apply plus separate vector passes; graph replays; mixed solves that reach
the 1e-10 true residual with it."""
import numpy as np
import pytest

import oracle as O
from helpers import banner, relerr
from oracle.cheb import cheb_smooth, links_to_dense
from schwinger_mg_amd.fields import gaussian_vecs, random_gauge

pytestmark = pytest.mark.gpu

MASS, R, CSW = 0.05, 1.0, 1.0
RATIO = 8.0


def _hierarchy(smg, L, levels, k, seed, coarse_half, coarse_solver=1, coarse_maxit=30):
    V = L * L
    U = random_gauge(L, 0.4, seed)
    D = smg.DiracOp(L, U, MASS, R, CSW, 0.0)
    dims = [2 * V, (L // 4) ** 2 * k]
    starts = np.concatenate([gaussian_vecs(k // 2, n, seed + 1 + i).ravel() for i, n in enumerate(dims[:levels - 1])])
    blocks = ((4, 4), (2, 2))[:levels - 1]
    mg = smg.MGHierarchy(D, starts, n_levels=levels, blocks=blocks, k=(k,) * (levels - 1), null_iters=5,
                         coarse_maxit=coarse_maxit, coarse_tol=0.1, coarse_solver=coarse_solver,
                         coarse_half=coarse_half, chiral=True, w_cycle=False)
    return U, D, mg


def _half(links, nx, k):
    """The FP16 links of a Hermitian coarse operator as the library stores them: each
    component rounded to FP16 (through FP32), then directions -x, -y, -x+y, -x-y replaced
    by the conjugate transposes of the neighbours' +x, +y, +x-y, +x+y links (so that the
    operator stays exactly Hermitian)."""
    h = (np.asarray(links.real, np.float32).astype(np.float16).astype(np.float64)
         + 1j * np.asarray(links.imag, np.float32).astype(np.float16).astype(np.float64))
    h = h.reshape(nx * nx, 9, k, k)
    for b in range(nx * nx):
        x, y = b % nx, b // nx
        # (direction, its reverse, neighbour offset): -x <- +x at x-1; -y <- +y at y-1;
        # -x+y <- +x-y at (x-1, y+1); -x-y <- +x+y at (x-1, y-1)
        for d, rd, dx, dy in ((2, 1, -1, 0), (4, 3, 0, -1), (7, 6, -1, 1), (8, 5, -1, -1)):
            n = (x + dx) % nx + ((y + dy) % nx) * nx
            h[b, d] = h[n, rd].conj().T
    return h.ravel()


def _numpy_cycle(mg, U, L, levels, k, hi, pre=3, post=3, half=False):
    """The V-cycle with Chebyshev smoothing on [hi_l / RATIO, hi_l] in FP64: the oracle's
    fine operator, the hierarchy's block vectors in the oracle's transfers, dense
    operators from its Galerkin links (rounded to FP16 at the smoothed coarse levels
    with half), the coarsest level solved exactly."""
    A0 = O.Dirac(L, U, MASS, R, CSW, 0.0).op()
    ops, Ps = [A0], []
    nx, nd = L, 2
    bl = ((4, 4), (2, 2))
    for l in range(levels - 1):
        bx, by = bl[l]
        P = O.Prolongator(nx, nx, nd, bx, by, k, chiral=True)
        P.set_block_vecs(mg.block_vecs(l))
        Ps.append(P)
        nx, nd = nx // bx, k
        links = mg.coarse_links(l)
        if half and l + 2 < levels:
            links = _half(links, nx, k)
        ops.append(links_to_dense(links, nx, nx, k))
    Minv = np.linalg.inv(ops[-1])
    apply = [ops[0]] + [(lambda M: (lambda v: M @ v))(M) for M in ops[1:-1]]

    def cycle(l, b):
        if l == levels - 1:
            return Minv @ b
        x, r = cheb_smooth(apply[l], b, hi[l] / RATIO, hi[l], pre)
        x = x + Ps[l].prolong(cycle(l + 1, Ps[l].restrict(r)))
        x, _ = cheb_smooth(apply[l], b, hi[l] / RATIO, hi[l], post, x0=x)
        return x

    return lambda b: cycle(0, b)


@pytest.mark.parametrize("levels,L,N,pre,post,k,mfma", [(2, 16, 16, 3, 3, 16, 0), (3, 32, 32, 3, 3, 16, 0),
                                                        (3, 32, 32, 1, 2, 16, 0), (2, 16, 16, 1, 1, 16, 0),
                                                        (2, 16, 16, 2, 1, 16, 0), (3, 32, 32, 1, 2, 32, 1),
                                                        (3, 32, 16, 3, 3, 32, 1)])
def test_cheb_smoothed_cycle_vs_numpy(smg, levels, L, N, pre, post, k, mfma):
    """FP32 links at level 1 (its smoothing through the apply and separate passes), the
    fine level's steps one fused D^dag D launch each (L a multiple of 16, N of 16; the
    first step of each smoothing also doing its initial pass, from x = 0 or from the
    residual), the dense coarsest inverse: one cycle on N right-hand sides equals the
    FP64 numpy cycle to FP32 accuracy, and differs clearly from the MR-smoothed cycle.
    With mfma, every restriction and prolongation runs on the matrix cores (k = 32:
    16 block vectors per chirality)."""
    banner(f"Chebyshev-smoothed cycle vs numpy levels={levels} L={L} N={N} pre={pre} post={post} k={k} "
           f"mfma={mfma}")
    U, D, mg = _hierarchy(smg, L, levels, k, 400 + L, coarse_half=0)
    smg.tune_clear()
    if mfma:   # the transfers' fine lattices: (L, L, 2 dof) and (L/4, L/4, k dof)
        for nx, w in ((L, 128 * 2 + k), (L // 4, 128 * k + k))[:levels - 1]:
            for direction in (0, 1):
                smg.tune_set(smg.TUNE_TRANSFER, [nx, nx, w, N, direction], 2)
    mg.set_cycle(pre, post, 30, 0.1)
    B = gaussian_vecs(N, 2 * L * L, 410 + L)
    x_mr = mg.precondition_batched(B, smg.F32)
    hi = mg.set_cheb_smoother(RATIO)
    assert len(hi) == levels - 1 and np.all(hi > 0)
    X = mg.precondition_batched(B, smg.F32)
    # hi_l = 1.1 x a 40-step power-iteration estimate of level l's largest eigenvalue:
    # against the dense operators' (the fine one from the oracle's operator on the unit
    # vectors, the coarse ones from the Galerkin links)
    A0 = O.Dirac(L, U, MASS, R, CSW, 0.0).op()
    n0 = 2 * L * L
    dense = [np.stack([A0(np.eye(1, n0, i, dtype=np.complex128)[0]) for i in range(n0)], axis=1)]
    for l in range(1, levels - 1):
        dense.append(links_to_dense(mg.coarse_links(l - 1), L // 4, L // 4, k))
    for l in range(levels - 1):
        lmax = np.linalg.eigvalsh(dense[l])[-1]
        print(f"level {l}: hi / 1.1 = {hi[l] / 1.1:.6g}, largest eigenvalue {lmax:.6g}")
        assert 0.9 * lmax <= hi[l] / 1.1 <= lmax * (1 + 1e-9)
    cycle = _numpy_cycle(mg, U, L, levels, k, hi, pre, post)
    ref = [cycle(B[j]) for j in range(N)]
    errs = [relerr(X[j], ref[j]) for j in range(N)]
    print(f"max relative deviation from the FP64 numpy cycle {max(errs):.2e}; from MR {relerr(X, x_mr):.2e}")
    assert max(errs) < 2e-5
    assert relerr(X, x_mr) > 1e-2
    # the FP64 batched cycle with the same smoother (FP64 links, dense coarsest inverse)
    X64 = mg.precondition_batched(B, smg.F64)
    e64 = max(relerr(X64[j], ref[j]) for j in range(N))
    print(f"FP64 cycle: max relative deviation from the numpy cycle {e64:.2e}")
    assert e64 < 1e-10
    mg.set_cheb_smoother(0.0)
    assert relerr(mg.precondition_batched(B, smg.F32), x_mr) < 1e-12   # MR restored
    smg.tune_clear()


@pytest.mark.parametrize("N,pre,post", [(16, 3, 3), (32, 3, 3), (32, 1, 2), (32, 1, 1)])
def test_fused_steps_equal_separate_passes(smg, N, pre, post):
    """FP16 links at level 1: both levels' smoothing steps are single launches (fine
    fused D^dag D, coarse matrix-core apply with the Chebyshev epilogue); the cycle
    equals the one whose steps are the apply plus separate vector passes, and both equal
    the FP64 numpy cycle with the same FP16-rounded level-1 links, to the accuracy of the
    level-1 products (split BF16: the FP32 sources keep about 16 bits)."""
    banner(f"fused vs separate Chebyshev smoothing N={N} pre={pre} post={post}")
    L, k = 32, 16
    U, D, mg = _hierarchy(smg, L, 3, k, 420, coarse_half=1)
    mg.set_cycle(pre, post, 30, 0.1)
    B = gaussian_vecs(N, 2 * L * L, 421)
    h1 = mg.set_cheb_smoother(RATIO, fused=True)
    Xf = mg.precondition_batched(B, smg.F32)
    h0 = mg.set_cheb_smoother(RATIO, fused=False)
    Xs = mg.precondition_batched(B, smg.F32)
    assert np.array_equal(h1, h0)
    cycle = _numpy_cycle(mg, U, L, 3, k, h1, pre, post, half=True)
    ref = np.array([cycle(B[j]) for j in range(N)])
    d, ef, es = relerr(Xf, Xs), relerr(Xf, ref), relerr(Xs, ref)
    print(f"fused vs separate passes {d:.2e}; vs the numpy cycle: fused {ef:.2e}, separate {es:.2e}")
    assert d < 1e-5 and ef < 1e-4 and es < 1e-4


def test_cheb_smoothed_cycle_graph_replays_equal_eager(smg):
    """The Chebyshev-smoothed cycle (no reductions in its smoothers) captured as a
    hipGraph: three replays with eager work between them equal the eager cycle bit for
    bit."""
    L, k, N = 32, 16, 32
    U, D, mg = _hierarchy(smg, L, 3, k, 430, coarse_half=2, coarse_solver=2, coarse_maxit=32)
    mg.set_cheb_smoother(RATIO)
    B = gaussian_vecs(N, 2 * L * L, 431)
    mg.set_cycle(3, 3, 32, 0.1, use_graph=False)
    eager = mg.precondition_batched(B, smg.F32)
    mg.set_cycle(3, 3, 32, 0.1, use_graph=True)
    for rep in range(3):
        assert np.array_equal(mg.precondition_batched(B, smg.F32), eager), rep
        D.apply(B[rep])
    mg.set_cycle(3, 3, 32, 0.1, use_graph=False)


@pytest.mark.parametrize("levels", [2, 3])
def test_mixed_solve_with_cheb_smoother(smg, levels):
    """The mixed batched solve (FP32 GCR and cycle inside FP64 refinement) with the
    Chebyshev smoother and the Chebyshev coarsest solve reaches the 1e-10 FP64 true
    residual (the oracle's operator), in no more iterations than with MR + 3, and the
    same solution."""
    banner(f"mixed solve, Chebyshev smoother, {levels} levels")
    L, k, N = 32, 32, 32
    U, D, mg = _hierarchy(smg, L, levels, k, 440, coarse_half=2, coarse_solver=2, coarse_maxit=40)
    A = O.Dirac(L, U, MASS, R, CSW, 0.0).op()
    B = gaussian_vecs(N, 2 * L * L, 441)
    out = {}
    for ratio in (0.0, RATIO):
        mg.set_cheb_smoother(ratio)
        X, info, tr = smg.gcr_solve_batched(D, B, mg, 200, 1e-10, 10, prec=smg.F32, inner_tol=1e-5)
        o = np.array([np.linalg.norm(B[j] - A(X[j])) / np.linalg.norm(B[j]) for j in range(N)])
        assert o.max() < 1e-10, (ratio, o.max())
        out[ratio] = (X, info.iterations)
    print(f"iterations: MR {out[0.0][1]}, Chebyshev {out[RATIO][1]}")
    assert out[RATIO][1] <= out[0.0][1] + 3
    assert relerr(out[RATIO][0], out[0.0][0]) < 1e-8
