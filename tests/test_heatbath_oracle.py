"""The U(1) heatbath's restatement (no reference counterpart; the reference
thermalises by HMC), pinned by what fixes it:
* Philox4x32-10 against the published known-answer vectors of its authors
  (Random123's kat_vectors: zero, all-ones and pi-digit counters and keys);
* the Best-Fisher sampler against the von Mises distribution (Kolmogorov-Smirnov);
* the heatbath's equilibrium against the exact 2D U(1) plaquette I1(beta)/I0(beta);
* the checkerboard's conditional: a link drawn with its staples fixed follows
  exp(beta Re(U A)), so <Re(U A)>/|A| = I1(kappa)/I0(kappa), kappa = beta |A|."""
import numpy as np
import pytest
from scipy import special, stats

import oracle as O
from schwinger_mg_amd.fields import random_gauge


@pytest.mark.parametrize("ctr,key,want", [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
])
def test_philox_known_answers(ctr, key, want):
    assert list(O.philox4x32(ctr, key)) == want


@pytest.mark.parametrize("kappa", [0.0, 0.3, 2.0, 12.0])
def test_von_mises_distribution(kappa):
    x = np.array([O.von_mises(kappa, i, 11, 2024) for i in range(20000)])
    assert np.all(np.abs(x) <= np.pi)
    if kappa == 0.0:
        p = stats.kstest((x + np.pi) / (2 * np.pi), "uniform").pvalue
    else:
        p = stats.kstest(x, stats.vonmises(kappa).cdf).pvalue
    assert p > 1e-3
    # and the first moment: <cos x> = I1/I0
    assert abs(np.cos(x).mean() - special.i1(kappa) / special.i0(kappa)) < 5 * np.sqrt(0.5 / len(x))


def _plaq(L, U):
    V = L * L
    U0, U1 = U[:V].reshape(L, L), U[V:].reshape(L, L)   # [y][x]
    return (U0 * np.roll(U1, -1, axis=1) * np.conj(np.roll(U0, -1, axis=0)) * np.conj(U1)).real.mean()


@pytest.mark.parametrize("beta", [1.0, 4.0])
def test_heatbath_equilibrium_plaquette(beta):
    """After thermalisation the average plaquette of the 2D U(1) Wilson action is
    I1(beta)/I0(beta) (up to exponentially small finite-volume terms); 400 sweeps at
    L = 32, blocked errors, 4 sigma."""
    L = 32
    U = O.heatbath(L, random_gauge(L, 0.35, 5), beta, 100, 99)
    P = []
    for sw in range(400):
        U = O.heatbath(L, U, beta, 1, 99, 100 + sw)
        P.append(_plaq(L, U))
    P = np.array(P)
    b = P.reshape(40, 10).mean(1)
    err = b.std(ddof=1) / np.sqrt(len(b))
    assert abs(P.mean() - special.i1(beta) / special.i0(beta)) < 4 * err


def test_heatbath_sweep_numbering_and_unit_links():
    """Sweeps are numbered: n then m sweeps equal n + m at once; links stay on U(1)."""
    L = 8
    U = random_gauge(L, 0.35, 6)
    a = O.heatbath(L, O.heatbath(L, U, 2.0, 3, 7), 2.0, 2, 7, 3)
    b = O.heatbath(L, U, 2.0, 5, 7)
    assert np.array_equal(a, b)
    assert np.abs(np.abs(b) - 1).max() < 1e-14
    assert not np.array_equal(O.heatbath(L, U, 2.0, 1, 8), O.heatbath(L, U, 2.0, 1, 7))
