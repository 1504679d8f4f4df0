"""Many lattices of different sizes in one launch (no reference counterpart; the
reference runs one lattice at a time): every lattice's operator equals the
single-lattice operator and the oracle's on the same links, and the lockstep CG on
D^dag D solves every lattice as cg_solve does, at the configuration-3 batch of 40
lattices L in {64, 96, 128, 192, 256} (8 each)."""
import numpy as np
import pytest

import oracle as O
from helpers import banner, relerr
from schwinger_mg_amd.fields import gaussian_vec, random_gauge

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("csw,mu", [(0.0, 0.0), (1.0, 0.0), (1.0, 0.1)])
def test_lattice_batch_apply(smg, csw, mu):
    banner(f"lattice batch apply csw={csw} mu={mu}")
    Ls = [8, 6, 16, 10, 4, 32]
    Us = [random_gauge(L, 0.4, 500 + i) for i, L in enumerate(Ls)]
    masses = [0.05 + 0.01 * i for i in range(len(Ls))]
    B = smg.Lattices(Ls, Us, masses, 1.0, csw, mu)
    psis = [gaussian_vec(2 * L * L, 600 + i) for i, L in enumerate(Ls)]
    for kind in (smg.D_, smg.DDAG, smg.DDAGD, smg.G5D):
        got = B.apply(kind, psis)
        for g, L in enumerate(Ls):
            D = smg.DiracOp(L, Us[g], masses[g], 1.0, csw, mu)
            single = D._apply(kind, psis[g])
            oracle = O.Dirac(L, Us[g], masses[g], 1.0, csw, mu).apply(kind, psis[g])
            assert relerr(got[g], single) < 1e-15, (kind, L)
            assert relerr(got[g], oracle) < 1e-13, (kind, L)


def test_lattice_batch_cg_matches_single(smg):
    banner("lattice batch CG vs single-lattice CG")
    Ls = [16, 8, 32, 12]
    Us = [random_gauge(L, 0.4, 700 + i) for i, L in enumerate(Ls)]
    masses = [0.1, 0.2, 0.05, 0.1]
    B = smg.Lattices(Ls, Us, masses, 1.0, 1.0, 0.0)
    bs = [gaussian_vec(2 * L * L, 710 + i) for i, L in enumerate(Ls)]
    xs, its, tr = B.cg(bs, 2000, 1e-10)
    for g, L in enumerate(Ls):
        D = smg.DiracOp(L, Us[g], masses[g], 1.0, 1.0, 0.0)
        r = smg.cg_solve(D, bs[g], 2000, 1e-10)
        print(f"L={L}: {its[g]} vs {r.iterations} iterations, true residual {tr[g]:.2e}")
        assert abs(int(its[g]) - r.iterations) <= 1
        assert relerr(xs[g], r.solution) < 1e-8 and tr[g] < 2e-10


def test_config3_lattice_batch(smg):
    """The 40 lattices of configuration 3 (theta ~ N(0, 0.4^2), independent seeds, m=0.1,
    c_sw=1) solved together to 1e-10; two of them checked against the oracle's CG."""
    banner("configuration-3 lattice batch: 40 lattices")
    Ls = [L for L in (64, 96, 128, 192, 256) for _ in range(8)]
    Us = [random_gauge(L, 0.4, 800 + i) for i, L in enumerate(Ls)]
    B = smg.Lattices(Ls, Us, 0.1, 1.0, 1.0, 0.0)
    bs = [gaussian_vec(2 * L * L, 900 + i) for i, L in enumerate(Ls)]
    xs, its, tr = B.cg(bs, 5000, 1e-10)
    print(f"iterations {int(its.min())}..{int(its.max())}, worst true residual {tr.max():.2e}")
    assert (tr < 2e-10).all()
    for g in (0, 17):
        A = O.Dirac(Ls[g], Us[g], 0.1, 1.0, 1.0, 0.0).op()
        xo, ito, _ = O.cg(A, bs[g], 5000, 1e-10)
        assert abs(int(its[g]) - ito) <= 2 and relerr(xs[g], xo) < 1e-8
