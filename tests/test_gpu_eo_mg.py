"""Even-odd multigrid on the Schur operator M^dag M (EvenOddMG, src/eo_multigrid.h:
26-287) on the GPU: the Richardson weight, the dense Galerkin matrix and the
preconditioner against the reference's own output (tests/golden/eo_mg_vectors.npz,
written by tests/golden/make_eo_mg.py from the compiled reference) and against the
numpy restatement (oracle/eo_mg.py) on a larger lattice; the smoothing operator
passed as a host callable (the reference's A_MdagM argument) gives the device result."""
import os

import numpy as np
import pytest

import oracle as O
from helpers import banner, relerr
from oracle.eo_mg import EvenOddMG as OracleEOMG
from schwinger_mg_amd.fields import gaussian_vec, gaussian_vecs, random_gauge

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "eo_mg_vectors.npz")


@pytest.mark.parametrize("ci", range(3))
def test_eo_mg_vs_reference_fixture(smg, ci):
    g = np.load(GOLD)
    L, csw, mu, bx, by, k, ni, pre, post, seed = g[f"c{ci}_params"]
    L, bx, by, k, ni, pre, post, seed = (int(v) for v in (L, bx, by, k, ni, pre, post, seed))
    banner(f"even-odd MG vs the reference, L={L} csw={csw} mu={mu} blocks {bx}x{by} k={k}")
    V = L * L
    # the drop-in's draws equal the reference's: k x random_vec(V, rng) from mt19937(seed)
    assert np.array_equal(smg.random_vec(k * V, seed).reshape(k, V), g[f"c{ci}_starts"])
    D = smg.DiracOp(L, g[f"c{ci}_U"], 0.1, 1.0, csw, mu)
    M = smg.EvenOddMG(D, bx, by, k, pre, post, ni, g[f"c{ci}_starts"], g[f"c{ci}_lam"])
    assert M.coarse_dim == (L // bx) * (L // by) * k
    assert abs(M.richardson_omega / g[f"c{ci}_omega"][0] - 1) < 1e-12
    assert relerr(M.coarse_matrix(), g[f"c{ci}_mat"]) < 1e-12
    x = M.precondition(g[f"c{ci}_b"])
    e = relerr(x, g[f"c{ci}_x"])
    print(f"precondition relerr vs the reference {e:.2e}")
    assert e < 1e-10
    # the smoothing operator as a host callable, as the reference's callers pass it
    eo = smg.EvenOddDiracOp(D)
    assert relerr(M.precondition(g[f"c{ci}_b"], eo.apply_schur_dag_schur), x) < 1e-13


def test_eo_mg_vs_oracle_L32(smg):
    banner("even-odd MG vs the numpy oracle, L=32, 4x4 blocks, k=8")
    L, bx, by, k, ni, pre, post = 32, 4, 4, 8, 6, 3, 3
    V = L * L
    U = random_gauge(L, 0.4, 31)
    starts = gaussian_vecs(k, V, 32)
    lam = gaussian_vec(V, 33)
    b = gaussian_vec(V, 34)
    D = smg.DiracOp(L, U, 0.05, 1.0, 1.0, 0.02)
    M = smg.EvenOddMG(D, bx, by, k, pre, post, ni, starts, lam)
    Mo = OracleEOMG(O.Dirac(L, U, 0.05, 1.0, 1.0, 0.02), L, bx, by, k, ni, pre, post, starts, lam)
    assert abs(M.richardson_omega / Mo.omega - 1) < 1e-12
    assert relerr(M.coarse_matrix(), Mo.mat) < 1e-11
    e = relerr(M.precondition(b), Mo.precondition(b))
    print(f"precondition relerr vs the oracle {e:.2e}")
    assert e < 1e-10


def test_eo_mg_rejects_odd_blocks(smg):
    D = smg.DiracOp(16, random_gauge(16, 0.4, 1), 0.1, 1.0, 0.0, 0.0)
    with pytest.raises(RuntimeError, match="even bx"):
        smg.EvenOddMG(D, 3, 4, 4, 3, 3, 2, gaussian_vecs(4, 256, 2))
