"""The coarsest-level Chebyshev iteration restated in numpy.

TEST INFRASTRUCTURE ONLY: imported by tests/ and nowhere else.

No reference counterpart: the reference's coarsest solve is CG
(SparseCoarseOp::solve, src/coarse_op.h:492-547) and its smoother MR
(src/smoother.cpp:9-40).  This is the algorithm of the library's coarse_solver 2
option and of its optional Chebyshev smoother, Chebyshev acceleration for a Hermitian
positive-definite operator whose spectrum lies in [lo, hi] (Saad, Iterative Methods
for Sparse Linear Systems, 2nd ed., Algorithm 12.1), from x = 0.
"""
import numpy as np


def cheb_solve(A, b, lo, hi, steps):
    """x after `steps` applications of A: x_1 = b / theta, then per step
    r_k = r_{k-1} - A d_{k-1}, rho_k = 1/(2 sigma - rho_{k-1}),
    d_k = rho_k rho_{k-1} d_{k-1} + (2 rho_k / delta) r_k, x_{k+1} = x_k + d_k."""
    theta, delta = 0.5 * (hi + lo), 0.5 * (hi - lo)
    sigma = theta / delta
    rho = 1.0 / sigma
    r = np.array(b, dtype=np.complex128)
    d = r / theta
    x = d.copy()
    for _ in range(steps):
        r = r - A(d)
        rho_k = 1.0 / (2.0 * sigma - rho)
        d = (rho_k * rho) * d + (2.0 * rho_k / delta) * r
        x = x + d
        rho = rho_k
    return x


def cheb_smooth(A, b, lo, hi, niter, x0=None):
    """The batched FP32 cycle's Chebyshev smoother (niter applications of A).  From
    x = 0: x = cheb_solve(A, b, lo, hi, niter - 1), returned with its residual
    r = b - A x (pre-smoothing).  From x0: x0 + cheb_solve(A, b - A x0, lo, hi, niter),
    the post-smoothing's niter steps after its residual; r is then not formed (None).
    Residual polynomials: T_n((theta - A) / delta) / T_n(sigma) with n = niter from
    x = 0 and n = niter + 1 (times the residual of x0) from x0."""
    if x0 is None:
        if niter == 0:
            return np.zeros_like(np.asarray(b, dtype=np.complex128)), np.array(b, dtype=np.complex128)
        x = cheb_solve(A, b, lo, hi, niter - 1)
        return x, np.asarray(b) - A(x)
    x0 = np.asarray(x0, dtype=np.complex128)
    if niter == 0:
        return x0.copy(), None
    return x0 + cheb_solve(A, np.asarray(b) - A(x0), lo, hi, niter), None


def residual_factor(lam, lo, hi, steps):
    """The residual polynomial of cheb_solve: b - A x = p(A) b with
    p(lam) = T_n((theta - lam) / delta) / T_n(theta / delta), n = steps + 1, T_n the
    Chebyshev polynomial of the first kind, here from numpy's Chebyshev basis."""
    theta, delta = 0.5 * (hi + lo), 0.5 * (hi - lo)
    T = np.polynomial.chebyshev.Chebyshev.basis(steps + 1)
    return T((theta - np.asarray(lam)) / delta) / T(theta / delta)


_DIRS = ((0, 0), (1, 0), (-1, 0), (0, 1), (0, -1), (1, 1), (1, -1), (-1, 1), (-1, -1))


def links_to_dense(links, nx, ny, k):
    """The 9-point coarse operator of links[(b*9 + d)*k*k + ki*k + kj] (directions
    {0, +x, -x, +y, -y, +x+y, +x-y, -x+y, -x-y}, periodic) as a dense matrix; extents
    >= 3, where the nine neighbours of a site are distinct."""
    assert nx >= 3 and ny >= 3
    L = np.asarray(links).reshape(nx * ny, 9, k, k)
    n = nx * ny * k
    M = np.zeros((n, n), np.complex128)
    for b in range(nx * ny):
        x, y = b % nx, b // nx
        for d, (dx, dy) in enumerate(_DIRS):
            nb = (x + dx) % nx + ((y + dy) % ny) * nx
            M[b * k:(b + 1) * k, nb * k:(nb + 1) * k] += L[b, d]
    return M
