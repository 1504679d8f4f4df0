"""The coarsest-level Chebyshev iteration's restatement (oracle/cheb.py), pinned by
what the mathematics fixes: its residual is the shifted and scaled Chebyshev
polynomial of the operator applied to the right-hand side, and the 9-point
operator it is checked on equals the oracle's dense Galerkin product."""
import numpy as np

import oracle as O
from oracle.cheb import cheb_smooth, cheb_solve, links_to_dense, residual_factor
from schwinger_mg_amd.fields import gaussian_vec, gaussian_vecs, random_gauge


def _bound(lo, hi, steps):
    """max |p| over [lo, hi]: 1 / T_n(theta / delta)."""
    sigma = (hi + lo) / (hi - lo)
    return 1.0 / np.polynomial.chebyshev.Chebyshev.basis(steps + 1)(sigma)


def test_cheb_residual_is_the_chebyshev_polynomial():
    """On a diagonal operator, b - A x = T_n((theta - A)/delta) / T_n(sigma) b exactly
    (n = steps + 1), for eigenvalues inside the interval and below it; T_n from numpy's
    Chebyshev basis, not from the recurrence.  Inside the interval |p| <= 1/T_n(sigma)."""
    rng = np.random.default_rng(5)
    lo, hi = 0.02, 7.5
    lam = np.concatenate([rng.uniform(lo, hi, 200), rng.uniform(1e-4, lo, 20), [lo, hi]])
    b = rng.normal(size=lam.size) + 1j * rng.normal(size=lam.size)
    for steps in (0, 1, 2, 7, 30):
        x = cheb_solve(lambda v: lam * v, b, lo, hi, steps)
        p = residual_factor(lam, lo, hi, steps)
        assert np.abs((b - lam * x) - p * b).max() < 1e-10 * np.abs(b).max(), steps
        inside = (lam >= lo) & (lam <= hi)
        assert np.abs(p[inside]).max() <= _bound(lo, hi, steps) * (1 + 1e-12)
        assert np.all(np.abs(p[~inside]) < 1.0)   # below the interval: damped, never amplified
    # steps = 0 is x = b / theta
    x0 = cheb_solve(lambda v: lam * v, b, lo, hi, 0)
    assert np.allclose(x0, b / (0.5 * (lo + hi)), rtol=1e-15, atol=0)


def test_cheb_smoother_residual_polynomials():
    """The smoother on a diagonal operator: from x = 0 its residual (returned, and
    b - A x) is T_n((theta - A)/delta)/T_n(sigma) b with n = niter; from x0 the new
    residual is the same with n = niter + 1 times b - A x0.  T_n from numpy's basis."""
    rng = np.random.default_rng(7)
    lo, hi = 0.3, 6.0
    lam = np.concatenate([rng.uniform(0.01, hi, 100), [lo, hi]])
    b = rng.normal(size=lam.size) + 1j * rng.normal(size=lam.size)
    x0 = rng.normal(size=lam.size) + 1j * rng.normal(size=lam.size)
    A = lambda v: lam * v
    for niter in (1, 2, 3, 8):
        x, r = cheb_smooth(A, b, lo, hi, niter)
        want = residual_factor(lam, lo, hi, niter - 1) * b
        assert np.abs(r - want).max() < 1e-10 * np.abs(b).max(), niter
        assert np.abs((b - lam * x) - want).max() < 1e-10 * np.abs(b).max(), niter
        x1, r1 = cheb_smooth(A, b, lo, hi, niter, x0)
        assert r1 is None
        want1 = residual_factor(lam, lo, hi, niter) * (b - lam * x0)
        assert np.abs((b - lam * x1) - want1).max() < 1e-10 * np.abs(b - lam * x0).max(), niter
    x, r = cheb_smooth(A, b, lo, hi, 0)
    assert not np.any(x) and np.array_equal(r, b)


def test_cheb_on_a_dense_hermitian_operator():
    """A Hermitian positive-definite matrix with its spectrum inside [lo, hi]: the error
    falls at least as the polynomial's bound 1/T_n(sigma) predicts, and the residual is
    p(A) b computed by eigendecomposition."""
    rng = np.random.default_rng(6)
    n = 60
    Q, _ = np.linalg.qr(rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))
    lam = np.linspace(0.05, 4.0, n)
    A = (Q * lam) @ Q.conj().T
    b = rng.normal(size=n) + 1j * rng.normal(size=n)
    steps = 25
    x = cheb_solve(lambda v: A @ v, b, 0.05, 4.0, steps)
    want = Q @ (residual_factor(lam, 0.05, 4.0, steps) * (Q.conj().T @ b))
    assert np.abs((b - A @ x) - want).max() < 1e-10 * np.abs(b).max()
    assert np.linalg.norm(b - A @ x) <= _bound(0.05, 4.0, steps) * np.linalg.norm(b) * (1 + 1e-10)


def test_links_to_dense_equals_the_dense_galerkin_product():
    """links_to_dense of the oracle's 9-point Galerkin links equals the oracle's dense
    P^dag A P (coarse extents 4 x 4 and 8 x 8)."""
    for L, b, k, seed in ((16, 4, 4, 11), (32, 4, 8, 12)):
        U = random_gauge(L, 0.4, seed)
        P = O.Prolongator(L, L, 2, b, b, k).build(gaussian_vecs(k, 2 * L * L, seed + 1).ravel())
        A = O.Dirac(L, U, 0.1, 1.0, 1.0, 0.0).op()
        Co = O.Coarse(A, P, 1)
        dense = O.dense_galerkin(A, P).T   # dense_galerkin returns the columns as rows
        M = links_to_dense(Co.links(), L // b, L // b, k)
        assert np.abs(M - dense).max() < 1e-12 * np.abs(dense).max()
        v = gaussian_vec(M.shape[0], seed + 2)
        assert np.abs(M @ v - Co.op()(v)).max() < 1e-12 * np.abs(M @ v).max()
