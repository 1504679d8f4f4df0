"""The U(1) heatbath on the GPU against its restatement in the oracle (pinned by
Philox's known answers, the von Mises distribution and the exact plaquette): the same
Philox stream and the same sampler, so the fields agree to roundoff; and the
configuration-4 thermalisation (50 sweeps at beta = 4 from theta ~ N(0, 0.35^2)) at
L = 128 reaches the exact plaquette."""
import numpy as np
import pytest
from scipy import special

import oracle as O
from helpers import banner, relerr
from schwinger_mg_amd.fields import random_gauge

pytestmark = pytest.mark.gpu


def _plaq(L, U):
    V = L * L
    U0, U1 = U[:V].reshape(L, L), U[V:].reshape(L, L)
    return (U0 * np.roll(U1, -1, axis=1) * np.conj(np.roll(U0, -1, axis=0)) * np.conj(U1)).real.mean()


@pytest.mark.parametrize("L,beta,sweeps", [(8, 2.0, 5), (16, 4.0, 20), (32, 0.5, 10)])
def test_heatbath_vs_oracle(smg, L, beta, sweeps):
    banner(f"heatbath L={L} beta={beta} sweeps={sweeps}")
    U = random_gauge(L, 0.35, 40 + L)
    got = smg.heatbath(L, U, beta, sweeps, 1234)
    want = O.heatbath(L, U, beta, sweeps, 1234)
    print(f"max |dU| {np.abs(got - want).max():.2e}")
    assert np.abs(got - want).max() < 1e-10
    assert np.abs(np.abs(got) - 1).max() < 1e-14


def test_dirac_heatbath_thermalises_config4(smg):
    """L = 128, beta = 4: 50 sweeps on the operator's links, then the clover term and the
    FP32 links follow (D apply equals the oracle's on the new links); 200 more sweeps
    give the exact plaquette within 4 blocked standard errors."""
    banner("heatbath L=128 beta=4 (thermalisation)")
    L, beta = 128, 4.0
    U = random_gauge(L, 0.35, 77)
    D = smg.DiracOp(L, U, 0.1, 1.0, 1.0, 0.0)
    smg.dirac_heatbath(D, beta, 50, 4242)
    Ut = smg.gauge_of(D)
    # last-bit differences of the device and host transcendentals carry from sweep to
    # sweep: most links agree to 1e-13, the worst of 32768 after 50 sweeps to ~1e-10
    assert np.abs(Ut - O.heatbath(L, U, beta, 50, 4242)).max() < 1e-8
    psi = np.random.default_rng(78).standard_normal(2 * L * L) + 0j
    assert relerr(D.apply(psi), O.Dirac(L, Ut, 0.1, 1.0, 1.0, 0.0).apply(0, psi)) < 1e-12
    P = []
    for sw in range(200):
        smg.dirac_heatbath(D, beta, 1, 4242, 50 + sw)
        P.append(_plaq(L, smg.gauge_of(D)))
    P = np.array(P)
    b = P.reshape(20, 10).mean(1)
    err = b.std(ddof=1) / np.sqrt(len(b))
    exact = special.i1(beta) / special.i0(beta)
    print(f"<P> = {P.mean():.6f} +- {err:.6f}, exact {exact:.6f}")
    assert abs(P.mean() - exact) < 4 * err
