"""Parity of the HIP Wilson-clover stencil family with the oracle (FP64 at
<= 1e-12 relative, FP32 at <= 1e-5) and with the reference's golden vectors."""
import os

import numpy as np
import pytest

import oracle as O
from helpers import CONFIGS, banner, relerr
from schwinger_mg_amd.fields import antiperiodic_y, gaussian_vec, random_gauge

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_vectors.npz")


@pytest.mark.parametrize("L", [8, 16, 30])
@pytest.mark.parametrize("ci", range(4))
def test_dirac_family_fp64(smg, L, ci):
    csw, mu = CONFIGS[ci]
    banner(f"dirac fp64 L={L} csw={csw} mu={mu}")
    V = L * L
    U = random_gauge(L, 0.5, 100 + L)
    psi = gaussian_vec(2 * V, 200 + L)
    Dg, Do = smg.DiracOp(L, U, 0.1, 1.0, csw, mu), O.Dirac(L, U, 0.1, 1.0, csw, mu)
    assert relerr(Dg.clover_field, Do.clover()) < 1e-13 if csw else True
    assert relerr(Dg.apply(psi), Do.apply(0, psi)) < 1e-12
    assert relerr(Dg.apply_dag(psi), Do.apply(1, psi)) < 1e-12
    assert relerr(Dg.apply_DdagD(psi), Do.apply(2, psi)) < 1e-12
    assert relerr(Dg.apply_g5D(psi), Do.apply(3, psi)) < 1e-12
    pi = np.random.default_rng(L).normal(size=2 * V)
    assert relerr(Dg.apply_delta_D(psi, pi, 0.2), Do.delta(psi, pi, 0.2)) < 1e-12


@pytest.mark.parametrize("L", [8, 16])
@pytest.mark.parametrize("ci", range(4))
def test_even_odd_fp64(smg, L, ci):
    csw, mu = CONFIGS[ci]
    banner(f"even-odd fp64 L={L} csw={csw} mu={mu}")
    V = L * L
    U = random_gauge(L, 0.5, 300 + L)
    psi = gaussian_vec(2 * V, 400 + L)
    Dg, Do = smg.DiracOp(L, U, 0.1, 1.0, csw, mu), O.Dirac(L, U, 0.1, 1.0, csw, mu)
    eo = smg.EvenOddDiracOp(Dg)
    fns = [eo.apply_eo, eo.apply_oe, eo.apply_ee_inv, eo.apply_oo_inv, eo.apply_oo, eo.apply_schur,
           eo.apply_schur_dag, eo.apply_schur_dag_schur]
    for kind, fn in enumerate(fns):
        assert relerr(fn(psi[:V]), Do.eo(kind, psi[:V])) < 1e-12, kind
    assert relerr(eo.reconstruct(psi[:V], psi[::-1]), Do.reconstruct(psi[:V], psi[::-1])) < 1e-12
    assert relerr(eo.reconstruct_full(psi[:V]), Do.reconstruct(psi[:V])) < 1e-12
    assert abs(eo.log_det_ee() - Do.logdet()) < 1e-11 * abs(Do.logdet())


def test_golden_vectors(smg):
    banner("dirac golden L=8")
    g = np.load(GOLD)
    L, V = 8, 64
    for ci, (csw, mu) in enumerate(CONFIGS):
        D = smg.DiracOp(L, g["L8_U"], 0.1, 1.0, csw, mu)
        eo = smg.EvenOddDiracOp(D)
        for kind, fn in enumerate([D.apply, D.apply_dag, D.apply_DdagD, D.apply_g5D]):
            assert relerr(fn(g["L8_psi"]), g[f"L8_c{ci}_dirac{kind}"]) < 1e-12
        assert relerr(eo.apply_schur_dag_schur(g["L8_psi"][:V]), g[f"L8_c{ci}_eo7"]) < 1e-12
        assert relerr(smg.fermion_force(D, g["L8_psi"]), g[f"L8_c{ci}_force0"]) < 1e-12
    assert relerr(smg.gauge_force(L, g["L8_U"], 2.0), g["L8_gauge_force_b2"]) < 1e-12
    assert abs(smg.avg_plaq(L, g["L8_U"]) - g["L8_plaq"][0]) < 1e-13


@pytest.mark.parametrize("ci", range(4))
def test_dirac_fp32_dev_multi_rhs(smg, ci):
    import torch
    csw, mu = CONFIGS[ci]
    L, nrhs = 32, 4
    banner(f"dirac fp32/fp64 dev nrhs={nrhs} L={L} csw={csw} mu={mu}")
    V = L * L
    U = random_gauge(L, 0.5, 7)
    Dg, Do = smg.DiracOp(L, U, 0.1, 1.0, csw, mu), O.Dirac(L, U, 0.1, 1.0, csw, mu)
    src = np.stack([gaussian_vec(2 * V, 50 + i) for i in range(nrhs)])
    s64 = torch.from_numpy(src.copy()).cuda()
    s32 = s64.to(torch.complex64)
    for kind in range(4):
        want = np.stack([Do.apply(kind, src[i]) for i in range(nrhs)])
        d64 = torch.empty_like(s64)
        d32 = torch.empty_like(s32)
        st = torch.cuda.current_stream().cuda_stream
        Dg.apply_dev(kind, smg.F64, s64.data_ptr(), d64.data_ptr(), nrhs, st)
        Dg.apply_dev(kind, smg.F32, s32.data_ptr(), d32.data_ptr(), nrhs, st)
        torch.cuda.synchronize()
        assert relerr(d64.cpu().numpy(), want) < 1e-12, kind
        assert relerr(d32.cpu().numpy(), want) < 2e-6, kind


def test_free_field_and_antiperiodic(smg):
    banner("free field plane waves")
    L, m = 16, 0.3
    x = np.arange(L)
    s1 = np.array([[0, 1], [1, 0]]); s2 = np.array([[0, -1j], [1j, 0]])
    chi = np.array([0.4 - 0.1j, 0.2 + 0.9j])
    for anti in (False, True):
        U = np.ones(2 * L * L, np.complex128)
        if anti:
            U = antiperiodic_y(U, L)
        D = smg.DiracOp(L, U, m, 1.0, 1.0, 0.0)
        for nx, ny in [(0, 0), (3, 1), (5, 7)]:
            px = 2 * np.pi * nx / L
            py = 2 * np.pi * (ny + (0.5 if anti else 0.0)) / L
            ph = np.exp(1j * (px * x[None, :] + py * x[:, None])).ravel()
            Dp = (2 + m - np.cos(px) - np.cos(py)) * np.eye(2) + 1j * (s1 * np.sin(px) + s2 * np.sin(py))
            out = D.apply(np.stack([ph * chi[0], ph * chi[1]], 1).ravel())
            want = np.stack([ph * (Dp @ chi)[0], ph * (Dp @ chi)[1]], 1).ravel()
            assert relerr(out, want) < 1e-12


def test_blas_dev(smg):
    import torch
    banner("blas dot/norm dev")
    n = 100003
    a = gaussian_vec(n, 1); b = gaussian_vec(n, 2)
    ta, tb = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
    import ctypes
    out = np.zeros(2)
    smg._chk(smg.lib().smg_dot_dev(ta.data_ptr(), tb.data_ptr(), n, smg.F64, out.ctypes.data_as(ctypes.c_void_p), None))
    ref = O.dot(a, b)
    assert abs(complex(out[0], out[1]) - ref) < 1e-12 * abs(ref) + 1e-9
    nrm = ctypes.c_double()
    smg._chk(smg.lib().smg_norm_dev(ta.data_ptr(), n, smg.F64, ctypes.byref(nrm), None))
    assert abs(nrm.value - np.linalg.norm(a)) < 1e-12 * np.linalg.norm(a)


def test_error_path(smg):
    banner("error path")
    with pytest.raises(smg.SMGError):
        smg.Prolongator(10, 10, 2, 4, 4, 4)          # 4 does not tile 10
    D = smg.DiracOp(8, random_gauge(8, 0.3, 1), 0.1)
    with pytest.raises(smg.SMGError):
        D._apply(9, np.zeros(128, np.complex128))   # unknown kind
