"""The five workload configurations run as stated, on the GPU, against the oracle.

* C1: L = 64, theta ~ N(0, 0.5^2), m = 0.1, c_sw = 1, antiperiodic in y (the exact
  link flip U_y(x, L-1) -> -U_y): one Wilson-clover apply to 1e-12 of the oracle's,
  and CG on D^dag D to 1e-10 with the oracle's iteration count.
* C2: L = 256, theta ~ N(0, 0.3^2), m = 0.05, c_sw = 1: two levels, 4 x 4 aggregates of
  8 near-null vectors (chirally split: 16 coarse dof) from 200 FP32 MR sweeps, FP16
  coarse links in the FP32 cycle inside FP64 refinement around GCR(10), to a 1e-10
  true residual; the error against the FP64 solve of the same hierarchy.
* C3: L = 512, 32 right-hand sides, three levels (4 x 4 then 2 x 2, k = 32): every
  right-hand side's true residual recomputed with the oracle's operator, mixed and
  FP64, and the mixed-vs-FP64 solution error.
* C4: L = 128 after 50 quenched heatbath sweeps at beta = 4 from theta ~ N(0, 0.35^2):
  the adjoint force against central finite differences (eps = 1e-5, 64 random links,
  the cancellation-free identity S(U') - S(U) = -X'^dag (A' - A) X) to <= 1e-7, with X
  from the FP64 CG and from the mixed-precision MG solve, and against the oracle's
  fermion_force.
"""
import numpy as np
import pytest

import oracle as O
from helpers import banner, relerr
from schwinger_mg_amd.fields import antiperiodic_y, gaussian_vec, gaussian_vecs, random_gauge

pytestmark = pytest.mark.gpu


def _true_res(A, b, x):
    return np.linalg.norm(b - A(x)) / np.linalg.norm(b)


def test_config1_L64_antiperiodic_dslash_and_cg(smg):
    banner("config 1: L=64 antiperiodic, Dslash + CG on D^dag D")
    L, V, m = 64, 64 * 64, 0.1
    U = antiperiodic_y(random_gauge(L, 0.5, 1001), L)
    D, Do = smg.DiracOp(L, U, m, 1.0, 1.0, 0.0), O.Dirac(L, U, m, 1.0, 1.0, 0.0)
    psi = gaussian_vec(2 * V, 1002)
    assert relerr(D.apply(psi), Do.apply(0, psi)) < 1e-12
    # the flip leaves every plaquette, so the clover term, unchanged
    assert relerr(D.clover_field, O.clover(L, random_gauge(L, 0.5, 1001))) < 1e-13
    b = gaussian_vec(2 * V, 1003)
    r = smg.cg_solve(D, b, 5000, 1e-10)
    xo, ito, _ = O.cg(Do.op(), b, 5000, 1e-10)
    tres = _true_res(Do.op(), b, r.solution)
    print(f"C1 CG: {r.iterations} its (oracle {ito}), true residual {tres:.2e}, "
          f"vs oracle solution {relerr(r.solution, xo):.2e}")
    assert abs(r.iterations - ito) <= 1
    assert tres < 1e-10
    assert relerr(r.solution, xo) < 1e-8


def _config2_mg(smg, D, V):
    starts = gaussian_vecs(8, 2 * V, 2002)
    # 40 outer inverse iterations x 5 MR sweeps = 200 FP32 sweeps per vector
    return smg.MGHierarchy(D, starts, n_levels=2, blocks=((4, 4),), k=(16,), null_iters=40, chiral=True,
                           null_prec=smg.F32, coarse_half=True, coarse_maxit=200, coarse_tol=1e-2)


def test_config2_L256_two_level_fp16_coarse_gcr(smg):
    banner("config 2: L=256 two-level MG, FP16 coarse links, GCR(10)")
    L, V, m = 256, 256 * 256, 0.05
    U = random_gauge(L, 0.3, 2001)
    D, Do = smg.DiracOp(L, U, m, 1.0, 1.0, 0.0), O.Dirac(L, U, m, 1.0, 1.0, 0.0)
    A = Do.op()
    mg = _config2_mg(smg, D, V)
    # the Galerkin links of the device hierarchy equal the oracle's P^dag A P for the
    # same block vectors
    Po = O.Prolongator(L, L, 2, 4, 4, 16, chiral=True)
    Po.set_block_vecs(mg.block_vecs(0))
    assert relerr(mg.coarse_links(0), O.Coarse(A, Po, 1).links()) < 1e-12
    b = gaussian_vec(2 * V, 2003)
    r32 = smg.fcg_solve(D, b, mg, 500, 1e-10, trunc=10, prec=smg.F32)
    r64 = smg.fcg_solve(D, b, mg, 500, 1e-10, trunc=10, prec=smg.F64)
    t32, t64 = _true_res(A, b, r32.solution), _true_res(A, b, r64.solution)
    err = relerr(r32.solution, r64.solution)
    print(f"C2: mixed {r32.iterations} its ({r32.restarts} refinements) true res {t32:.2e}; "
          f"FP64 {r64.iterations} its true res {t64:.2e}; |x_mixed - x_64|/|x_64| = {err:.2e}")
    assert t32 < 1e-10 and t64 < 1e-10
    assert abs(r32.true_residual - t32) < 1e-2 * t32 + 1e-15   # the device's own FP64 residual
    assert err < 1e-8


@pytest.mark.parametrize("coarse_solver", [0, 2])
def test_config3_L512_32rhs_three_level(smg, coarse_solver):
    """coarse_solver 0: the coarsest solve is CG (BF16 vectors in the mixed cycle);
    2: the mixed cycle's coarsest solve is the Chebyshev iteration (40 steps), as the
    benchmark runs it, and the FP64 cycle's stays CG."""
    banner(f"config 3: L=512, 32 RHS, three-level MG, mixed and FP64, coarse_solver {coarse_solver}")
    L, V, m, N, k = 512, 512 * 512, 0.05, 32, 32
    U = random_gauge(L, 0.4, 2024)
    D = smg.DiracOp(L, U, m, 1.0, 1.0, 0.0)
    dims = [2 * V, (L // 4) ** 2 * k]
    starts = np.concatenate([gaussian_vecs(k // 2, n, 2034 + i).ravel() for i, n in enumerate(dims)])
    mg = smg.MGHierarchy(D, starts, n_levels=3, blocks=((4, 4), (2, 2)), k=(k, k), null_iters=10, w_cycle=False,
                         coarse_maxit=30, coarse_tol=0.1, coarse_half=2, use_graph=True, chiral=True,
                         coarse_solver=coarse_solver)
    B = gaussian_vecs(N, 2 * V, 2025)
    if coarse_solver == 2:
        mg.set_cycle(3, 3, 40, 0.1, use_graph=True)
    X32, i32, tr32 = smg.gcr_solve_batched(D, B, mg, 500, 1e-10, trunc=10, prec=smg.F32)
    mg.set_cycle(3, 3, 30, 0.1, use_graph=True)
    X64, i64, tr64 = smg.gcr_solve_batched(D, B, mg, 500, 1e-10, trunc=10, prec=smg.F64)
    A = O.Dirac(L, U, m, 1.0, 1.0, 0.0).op()
    o32 = np.array([_true_res(A, B[j], X32[j]) for j in range(N)])
    o64 = np.array([_true_res(A, B[j], X64[j]) for j in range(N)])
    err = max(np.linalg.norm(X32[j] - X64[j]) / np.linalg.norm(X64[j]) for j in range(N))
    print(f"C3: mixed {i32.iterations} its, max true res {o32.max():.2e} (device {tr32.max():.2e}); "
          f"FP64 {i64.iterations} its, max {o64.max():.2e}; max mixed-vs-FP64 error {err:.2e}")
    assert o32.max() < 1e-10 and o64.max() < 1e-10
    assert np.abs(o32 - tr32).max() < 1e-2 * o32.max()
    assert err < 1e-8


@pytest.fixture(scope="module")
def config4(smg):
    """L = 128 links after 50 heatbath sweeps, phi = D^dag eta, and the cancellation-free
    finite differences (FP64 CG to 1e-14) on 64 random links."""
    banner("config 4: L=128 heatbath links, finite differences on 64 links")
    L, V, m, beta, eps = 128, 128 * 128, 0.1, 4.0, 1e-5
    D = smg.DiracOp(L, random_gauge(L, 0.35, 4000), m, 1.0, 1.0, 0.0)
    smg.dirac_heatbath(D, beta, 50, 4001)
    U = smg.gauge_of(D)
    phi = D.apply_dag(gaussian_vec(2 * V, 4002))
    X = smg.cg_solve(D, phi, 20000, 1e-14).solution
    AX = D.apply_DdagD(X)
    links = np.random.default_rng(4003).choice(2 * V, size=64, replace=False)
    Dp = smg.DiracOp(L, U, m, 1.0, 1.0, 0.0)
    fd = np.zeros(len(links))
    for i, li in enumerate(links):
        dS = []
        for sgn in (1, -1):
            Up = U.copy()
            Up[li] *= np.exp(1j * sgn * eps)
            Dp.set_gauge(Up)
            Xp = smg.cg_solve(Dp, phi, 20000, 1e-14).solution
            dS.append(-np.vdot(Xp, Dp.apply_DdagD(X) - AX).real)
        fd[i] = -(dS[0] - dS[1]) / (2 * eps)
    return dict(L=L, V=V, m=m, U=U, D=D, phi=phi, X=X, links=links, fd=fd)


@pytest.mark.parametrize("form", ["fp64", "mixed"])
def test_config4_L128_force_vs_finite_differences(smg, config4, form):
    c = config4
    L, V, D, phi = c["L"], c["V"], c["D"], c["phi"]
    banner(f"config 4: adjoint force, X from the {form} solve")
    if form == "fp64":
        X = c["X"]
    else:
        starts = gaussian_vecs(8, 2 * V, 4004)
        mg = smg.MGHierarchy(D, starts, n_levels=2, blocks=((4, 4),), k=(16,), null_iters=20, chiral=True,
                             null_prec=smg.F32, coarse_half=True, coarse_maxit=200, coarse_tol=1e-2)
        r = smg.fcg_solve(D, phi, mg, 500, 1e-12, trunc=10, prec=smg.F32)
        assert r.true_residual < 1e-12
        X = r.solution
    F = smg.fermion_force(D, X)
    Fo = O.Dirac(L, c["U"], c["m"], 1.0, 1.0, 0.0).force(X)
    assert relerr(F, Fo) < 1e-12
    worst = np.abs(c["fd"] - F[c["links"]]).max() / np.abs(F).max()
    print(f"C4 {form}: max |FD - F| / max |F| = {worst:.2e}")
    assert worst < 1e-7


@pytest.fixture(scope="module")
def config5_small(smg):
    """Configuration 5's integrator at L = 32: beta = 6, m = 0.02, c_sw = 1, tau = 1,
    Omelyan outer steps (fermion force) n_outer = 32, Omelyan inner steps (gauge force)
    n_gauge = 8 per outer position update, links after 200 heatbath sweeps; the oracle's
    trajectory (FP64 CG to 1e-13) from the same momenta and pseudofermion."""
    L, V, m, beta, n_outer = 32, 32 * 32, 0.02, 6.0, 32
    D = smg.DiracOp(L, random_gauge(L, 0.3, 5001), m, 1.0, 1.0, 0.0)
    smg.dirac_heatbath(D, beta, 200, 5002)
    U = smg.gauge_of(D)
    rng = np.random.default_rng(5003)
    pi0 = rng.normal(size=2 * V)
    eta = rng.normal(size=2 * V) + 1j * rng.normal(size=2 * V)
    Do = O.Dirac(L, U, m, 1.0, 1.0, 0.0)
    phi = Do.apply(1, eta)
    h0, h1 = np.zeros(4), np.zeros(4)
    O.lib().orc_hmc_hamiltonian(L, O._ptr(U), m, 1.0, 1.0, 0.0, beta, 20000, 1e-13, O._ptr(pi0), O._ptr(phi),
                                O._ptr(h0))
    Uo, po, _ = O.hmc_nested_evolve(L, U, m, 1.0, 1.0, 0.0, beta, 1.0, n_outer, 8, 1, 1, 20000, 1e-13, pi0, phi)
    O.lib().orc_hmc_hamiltonian(L, O._ptr(Uo), m, 1.0, 1.0, 0.0, beta, 20000, 1e-13, O._ptr(po), O._ptr(phi),
                                O._ptr(h1))
    dHo = h1[:3].sum() - h0[:3].sum()
    return dict(L=L, V=V, m=m, beta=beta, n_outer=n_outer, U=U, pi0=pi0, eta=eta, Uo=Uo, dHo=dHo,
                hscale=max(1.0, np.abs(h0[:3]).max()))


@pytest.mark.parametrize("solver", ["fp64", "mixed"])
def test_config5_nested_omelyan_reduced_vs_oracle(smg, config5_small, solver):
    """The device trajectory equals the oracle's: with FP64 CG to dH 1e-8 (relative to
    the Hamiltonian's terms) and the links to 1e-10; with the mixed-precision MG solver,
    the chronological initial guess and the hierarchy refreshed every second outer
    position update, to the solver's 1e-12 tolerance.  At this step count |dH| < 1."""
    c = config5_small
    L, V = c["L"], c["V"]
    banner(f"config 5 reduced: L={L} beta=6 m=0.02 n_outer={c['n_outer']} n_gauge=8, {solver}")
    D = smg.DiracOp(L, c["U"], c["m"], 1.0, 1.0, 0.0)
    if solver == "fp64":
        p = smg.nested_params(c["beta"], 1.0, c["n_outer"], 8, True, True, smg.F64, 20000, 1e-13)
        mg = None
    else:
        starts = gaussian_vecs(4, 2 * V, 5004)
        mg = smg.MGHierarchy(D, starts, n_levels=2, blocks=((4, 4),), k=(8,), null_iters=20, chiral=True,
                             coarse_half=True, coarse_maxit=200, coarse_tol=1e-2)
        p = smg.nested_params(c["beta"], 1.0, c["n_outer"], 8, True, True, smg.F32, 500, 1e-12, mg_refresh=2,
                              mg_refresh_iters=2, chrono_guess=True)
    res = smg.hmc_trajectory_nested(D, p, c["pi0"], c["eta"], 0.05, mg=mg)
    dU = np.abs(smg.gauge_of(D) - c["Uo"]).max()
    print(f"C5 reduced {solver}: dH {res.dH:.6e} (oracle {c['dHo']:.6e}), accepted {res.accepted}, "
          f"solver iterations {res.total_cg_iters}, max true residual {res.max_true_residual:.2e}, |dU| {dU:.2e}")
    assert abs(c["dHo"]) < 1.0 and res.accepted == 1
    if solver == "fp64":
        assert abs(res.dH - c["dHo"]) < 1e-8 * c["hscale"]
        assert dU < 1e-10
        assert res.max_true_residual < 1e-12
    else:
        assert abs(res.dH - c["dHo"]) < 1e-6 * c["hscale"]
        assert dU < 1e-7
        assert res.max_true_residual < 1e-12


def test_nested_reject_with_refresh_restores_field_and_hierarchy(smg):
    """A rejected trajectory whose hierarchy followed the proposal: the links and the
    near-null vectors come back and the hierarchy is rebuilt on them, so the next solve on
    the kept field is the solve of the hierarchy the trajectory started from."""
    banner("nested reject with MG refresh L=32")
    L, V, m = 32, 32 * 32, 0.05
    U = random_gauge(L, 0.3, 5101)
    starts = gaussian_vecs(4, 2 * V, 5102)

    def hierarchy(D):
        return smg.MGHierarchy(D, starts, n_levels=2, blocks=((4, 4),), k=(8,), null_iters=20, chiral=True,
                               coarse_half=True, coarse_maxit=200, coarse_tol=1e-2)
    D = smg.DiracOp(L, U, m, 1.0, 1.0, 0.0)
    mg = hierarchy(D)
    rng = np.random.default_rng(5103)
    pi0 = rng.normal(size=2 * V)
    eta = rng.normal(size=2 * V) + 1j * rng.normal(size=2 * V)
    # two outer steps over tau = 1 at beta = 6: a violent proposal, rejected with u ~ 1
    p = smg.nested_params(6.0, 1.0, 2, 2, True, True, smg.F32, 500, 1e-10, mg_refresh=1, mg_refresh_iters=2)
    res = smg.hmc_trajectory_nested(D, p, pi0, eta, 1.0 - 1e-16, mg=mg)
    assert res.dH > 0 and res.accepted == 0
    assert res.max_true_residual < 1e-10
    assert np.array_equal(smg.gauge_of(D), U)
    b = gaussian_vec(2 * V, 5104)
    got = smg.fcg_solve(D, b, mg, 500, 1e-10, trunc=10, prec=smg.F32)
    D2 = smg.DiracOp(L, U, m, 1.0, 1.0, 0.0)
    mg2 = hierarchy(D2)
    ref = smg.fcg_solve(D2, b, mg2, 500, 1e-10, trunc=10, prec=smg.F32)
    print(f"after reject: {got.iterations} iterations (fresh hierarchy {ref.iterations}), "
          f"solution difference {relerr(got.solution, ref.solution):.2e}")
    assert got.true_residual < 1e-10
    assert got.iterations == ref.iterations
    assert relerr(got.solution, ref.solution) < 1e-13
