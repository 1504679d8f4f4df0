"""Multigrid over many lattices of different sizes in one launch per kernel (no
reference counterpart; the reference runs one lattice at a time): every lattice's
V-cycle equals the single-lattice cycle on the same hierarchy to roundoff, the batched GCR
solves every lattice as fcg_solve_mg does (iteration counts within one, FP64 true
residuals below the tolerance), a lattice with a zero right-hand side stays at zero,
and the hipGraph replay of the cycle reproduces the eager cycle bit for bit."""
import numpy as np
import pytest

from helpers import banner, relerr
from schwinger_mg_amd.fields import gaussian_vec, random_gauge

pytestmark = pytest.mark.gpu

LS = [24, 32, 40, 48]          # coarse extents 6/8/10/12, coarsest 3/4/5/6
MASSES = [0.1, 0.05, 0.08, 0.1]
KW = dict(n_levels=3, blocks=((4, 4), (2, 2)), k=(8, 8), null_iters=10, pre_smooth=3, post_smooth=3,
          coarse_maxit=60, coarse_tol=1e-13, chiral=True)


def _setup(smg, use_graph=False, seed0=300):
    Us = [random_gauge(L, 0.4, seed0 + i) for i, L in enumerate(LS)]
    B = smg.Lattices(LS, Us, MASSES, 1.0, 1.0, 0.0)
    starts = []
    for g, L in enumerate(LS):
        sz = smg.mg_start_sizes(L, KW["n_levels"], KW["blocks"], KW["k"], KW["chiral"])
        starts.append(np.concatenate([smg.random_vec(nv * n, 40 + 7 * g + l) for l, (nv, n) in enumerate(sz)]))
    M = smg.LatticesMG(B, starts, use_graph=use_graph, **KW)
    return Us, B, M, starts


def _single(smg, Us, starts, g, **over):
    D = smg.DiracOp(LS[g], Us[g], MASSES[g], 1.0, 1.0, 0.0)
    kw = dict(KW)
    kw.update(over)
    return D, smg.MGHierarchy(D, starts[g], w_cycle=False, **kw)


def test_lattices_cycle_matches_single(smg):
    banner("many-lattice V-cycle vs the single-lattice cycle, 4 lattices L=24..48")
    Us, B, M, starts = _setup(smg)
    bs = [gaussian_vec(2 * L * L, 500 + g) for g, L in enumerate(LS)]
    got = M.precondition(bs)
    for g, L in enumerate(LS):
        D, mg = _single(smg, Us, starts, g)
        ref = mg.precondition(bs[g])
        e = relerr(got[g], ref)
        print(f"L={L}: cycle relerr {e:.2e}")
        # the many-lattice coarse operator sums a row in runs of 4 link entries per lane
        # (another order than the single-lattice kernel's), and the coarsest CG, run to
        # 1e-13, carries that roundoff through 60 iterations
        assert e < 1e-8, L


def test_lattices_gcr_matches_single(smg):
    banner("many-lattice GCR(7)+MG vs fcg_solve_mg per lattice to 1e-10")
    Us, B, M, starts = _setup(smg)
    bs = [gaussian_vec(2 * L * L, 600 + g) for g, L in enumerate(LS)]
    xs, its, tr, tot = M.gcr(bs, 200, 1e-10, trunc=7)
    for g, L in enumerate(LS):
        D, mg = _single(smg, Us, starts, g)
        r = smg.fcg_solve(D, bs[g], mg, 200, 1e-10, trunc=7)
        print(f"L={L}: {its[g]} vs {r.iterations} iterations, true residual {tr[g]:.2e} "
              f"(single {r.true_residual:.2e}), solution relerr {relerr(xs[g], r.solution):.2e}")
        assert abs(int(its[g]) - r.iterations) <= 1, L
        assert tr[g] < 1.5e-10, L
        assert relerr(xs[g], r.solution) < 1e-6, L   # both within 1e-10 in residual
    assert tot == int(its.max())


def test_lattices_gcr_zero_rhs_and_staggered_convergence(smg):
    banner("many-lattice GCR: a zero right-hand side, lattices converging at different iterations")
    Us, B, M, starts = _setup(smg)
    bs = [gaussian_vec(2 * L * L, 700 + g) for g, L in enumerate(LS)]
    bs[1] = np.zeros_like(bs[1])
    xs, its, tr, tot = M.gcr(bs, 200, 1e-10, trunc=5)
    print(f"iterations {list(its)}, true residuals {list(tr)}")
    assert its[1] == 0 and not np.any(xs[1])
    for g in (0, 2, 3):
        assert tr[g] < 1.5e-10
    # the same lattices alone (one at a time) take the same iterations
    for g in (0, 3):
        D, mg = _single(smg, Us, starts, g)
        r = smg.fcg_solve(D, bs[g], mg, 200, 1e-10, trunc=5)
        assert abs(int(its[g]) - r.iterations) <= 1


def test_lattices_gcr_fp16_links(smg):
    """The cycle on FP16 coarse links (FP64 vectors) inside the FP64 GCR: every lattice
    still reaches a 1e-10 FP64 true residual, within two iterations of the FP64-link solve."""
    banner("many-lattice GCR with FP16 coarse links")
    Us, B, M, starts = _setup(smg)
    M16 = smg.LatticesMG(B, starts, coarse_half=True, **KW)
    bs = [gaussian_vec(2 * L * L, 650 + g) for g, L in enumerate(LS)]
    _, its, tr, _ = M.gcr(bs, 200, 1e-10, trunc=7)
    x16, its16, tr16, _ = M16.gcr(bs, 200, 1e-10, trunc=7)
    print(f"FP64 links {list(its)}, FP16 links {list(its16)}; true residuals {list(tr16)}")
    assert (tr16 < 1.5e-10).all()
    assert (its16 <= its + 2).all()


def test_lattices_cycle_graph_replay(smg):
    banner("many-lattice V-cycle: hipGraph replays vs the eager cycle")
    Us, B, M, starts = _setup(smg, use_graph=False)
    _, _, Mg, _ = _setup(smg, use_graph=True)
    for rep in range(3):
        bs = [gaussian_vec(2 * L * L, 800 + 10 * rep + g) for g, L in enumerate(LS)]
        eager = M.precondition(bs)
        graph = Mg.precondition(bs)
        for g in range(len(LS)):
            assert np.array_equal(eager[g], graph[g]), (rep, g)


def test_config3_lattices_mg(smg):
    """The 40 lattices of configuration 3 (L in {64, 96, 128, 192, 256}, 8 each, theta ~
    N(0, 0.4^2), m = 0.1, c_sw = 1) under one three-level MG (4x4 then 2x2, chiral k=8)
    with GCR to 1e-10; the largest workgroup counts per lattice are capped, so every
    kernel strides.  Two lattices checked against the single-lattice solve."""
    banner("configuration-3 lattice batch: three-level MG + GCR over 40 lattices")
    Ls = [L for L in (64, 96, 128, 192, 256) for _ in range(8)]
    Us = [random_gauge(L, 0.4, 800 + i) for i, L in enumerate(Ls)]
    B = smg.Lattices(Ls, Us, 0.1, 1.0, 1.0, 0.0)
    kw = dict(n_levels=3, blocks=((4, 4), (2, 2)), k=(8, 8), null_iters=6, pre_smooth=3, post_smooth=3,
              coarse_maxit=30, coarse_tol=1e-1, chiral=True)
    starts = []
    for g, L in enumerate(Ls):
        sz = smg.mg_start_sizes(L, kw["n_levels"], kw["blocks"], kw["k"], kw["chiral"])
        starts.append(np.concatenate([smg.random_vec(nv * n, 90 + g + 50 * l) for l, (nv, n) in enumerate(sz)]))
    M = smg.LatticesMG(B, starts, use_graph=True, **kw)
    bs = [gaussian_vec(2 * L * L, 900 + i) for i, L in enumerate(Ls)]
    xs, its, tr, tot = M.gcr(bs, 500, 1e-10, trunc=7)
    print(f"iterations {int(its.min())}..{int(its.max())}, worst true residual {tr.max():.2e}")
    assert (tr < 1.5e-10).all()
    for g in (0, 39):
        D = smg.DiracOp(Ls[g], Us[g], 0.1, 1.0, 1.0, 0.0)
        mg = smg.MGHierarchy(D, starts[g], w_cycle=False, **kw)
        r = smg.fcg_solve(D, bs[g], mg, 500, 1e-10, trunc=7)
        print(f"L={Ls[g]}: {its[g]} vs {r.iterations} iterations; relerr {relerr(xs[g], r.solution):.2e}")
        assert abs(int(its[g]) - r.iterations) <= 1
