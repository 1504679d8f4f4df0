"""ctypes bindings for oracle/_ref/libref.so: the reference's own code, compiled
from its sources by oracle/Makefile (see ref_harness.cpp).

TEST INFRASTRUCTURE ONLY: used to pin the oracle and to write tests/golden/
fixtures.  ``available()`` is False wherever the library was not built.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PATH = os.path.join(HERE, "_ref", "libref.so")
_P, _I, _D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
_lib = None


def available():
    return os.path.exists(PATH)


def _ptr(a):
    return a.ctypes.data_as(_P) if a is not None else None


def _c(a, dt=np.complex128):
    return np.ascontiguousarray(a, dtype=dt)


def lib():
    global _lib
    if _lib is None:
        L = ctypes.CDLL(PATH)
        L.ref_random_vec.argtypes = [_I, ctypes.c_uint, _P]
        L.ref_eo_maps.argtypes = [_I, _P, _P, _P]
        L.ref_clover.argtypes = [_I, _P, _P]
        L.ref_avg_plaq.restype = _D
        L.ref_avg_plaq.argtypes = [_I, _P]
        L.ref_dirac_apply.argtypes = [_I, _P, _D, _D, _D, _D, _I, _P, _P]
        L.ref_dirac_delta.argtypes = [_I, _P, _D, _D, _D, _D, _P, _P, _P, _D]
        L.ref_eo_apply.argtypes = [_I, _P, _D, _D, _D, _D, _I, _P, _P]
        L.ref_eo_reconstruct.argtypes = [_I, _P, _D, _D, _D, _D, _P, _P, _P]
        L.ref_eo_logdet.restype = _D
        L.ref_eo_logdet.argtypes = [_I, _P, _D, _D, _D, _D]
        L.ref_prolong_build.argtypes = [_I, _I, _I, _P, _P]
        L.ref_prolong_apply.argtypes = [_I, _I, _I, _P, _I, _P, _P]
        L.ref_coarse_prolong_build.argtypes = [_I, _I, _I, _P, _P]
        L.ref_galerkin.argtypes = [_I, _P, _D, _D, _D, _D, _I, _I, _P, _P, _P]
        L.ref_dense_solve.argtypes = [_I, _P, _P, _P]
        L.ref_mr_smooth.argtypes = [_I, _P, _D, _D, _D, _D, _P, _P, _I]
        L.ref_lambda_max.restype = _D
        L.ref_lambda_max.argtypes = [_I, _P, _D, _D, _D, _D, _I]
        L.ref_krylov.restype = _I
        L.ref_krylov.argtypes = [_I, _P, _D, _D, _D, _D, _I, _P, _I, _D, _I, _P, _P]
        L.ref_fgmres_ritz.restype = _I
        L.ref_fgmres_ritz.argtypes = [_I, _P, _D, _D, _D, _D, _P, _I, _D, _I, _I, _P, _P, _P, _P, _P]
        L.ref_mg.restype = _I
        L.ref_mg.argtypes = [_I, _P, _D, _D, _D, _D, _I, _I, _I, _I, _I, _P, _I, _I, _I, _I,
                             _P, _P, _I, _D, _I, _P, _P]
        L.ref_gauge_force.argtypes = [_I, _P, _D, _P]
        L.ref_eo_mg.argtypes = [_I, _P, _D, _D, _D, _D, _I, _I, _I, _I, _I, _I, ctypes.c_uint, _P, _P, _P, _P, _P]
        L.ref_gauge_action.restype = _D
        L.ref_gauge_action.argtypes = [_I, _P, _D]
        L.ref_fermion_force.argtypes = [_I, _P, _D, _D, _D, _D, _I, _P, _P, _P]
        L.ref_fermion_action.restype = _D
        L.ref_fermion_action.argtypes = [_I, _P, _D, _D, _D, _D, _P, _I, _D]
        for name, args in (("ref_gauge_save_raw", [_I, _P, ctypes.c_char_p]),
                           ("ref_gauge_load_raw", [_I, _P, ctypes.c_char_p]),
                           ("ref_gauge_save", [_I, _P, _D, _D, ctypes.c_char_p]),
                           ("ref_gauge_load", [_I, _P, _P, ctypes.c_char_p])):
            getattr(L, name).restype = _I
            getattr(L, name).argtypes = args
        L.ref_hmc_draws.argtypes = [_I, ctypes.c_uint, _P, _I, _P]
        L.ref_hmc_trajectory.argtypes = [_I, _P, _D, _D, _D, _D, _D, _I, _I, _I, _I, _D, ctypes.c_uint, _P]
        L.ref_lanczos_eigen.argtypes = [_I, _P, _P, _P]
        L.ref_trlm_dirac.restype = _I
        L.ref_trlm_dirac.argtypes = [_I, _P, _D, _D, _D, _D, _I, _I, _I, _D, _P, _P]
        L.ref_lobpcg_dirac.restype = _I
        L.ref_lobpcg_dirac.argtypes = [_I, _P, _D, _D, _D, _D, _I, _P, _I, _D, _P, _P]
        L.ref_hmc_ms_trajectory.argtypes = [_I, _P, _D, _D, _D, _D, _D, _I, _I, _I, _I, _I, _I, _D, _I, _I, _P, _I,
                                            ctypes.c_uint, _P, _P, _P, _P, _P]
        L.ref_cg_deflated.restype = _I
        L.ref_cg_deflated.argtypes = [_I, _P, _D, _D, _D, _D, _I, _P, _P, _I, _D, _P, _P]
        L.ref_coarse_deflated_solve.restype = _I
        L.ref_coarse_deflated_solve.argtypes = [_I, _P, _D, _D, _D, _D, _I, _I, _P, _I, _I, _I, _D, _P, _I, _D,
                                                _P, _P]
        L.ref_leapfrog_evolve.restype = _I
        L.ref_leapfrog_evolve.argtypes = [_I, _P, _D, _D, _D, _D, _D, _I, _I, _I, _D, _P, _P]
        _lib = L
    return _lib


def random_vec(n, seed):
    """random_vec(n, std::mt19937(seed)) -- src/linalg.cpp:44-49."""
    out = np.zeros(n, np.complex128)
    lib().ref_random_vec(n, seed, _ptr(out))
    return out


def eo_maps(L):
    V = L * L
    e = np.zeros(V // 2, np.int32); o = np.zeros(V // 2, np.int32); i = np.zeros(V, np.int32)
    lib().ref_eo_maps(L, _ptr(e), _ptr(o), _ptr(i))
    return e, o, i


def clover(L, U):
    U = _c(U); F = np.zeros(L * L)
    lib().ref_clover(L, _ptr(U), _ptr(F))
    return F


def avg_plaq(L, U):
    return lib().ref_avg_plaq(L, _ptr(_c(U)))


class Dirac:
    def __init__(self, L, U, mass, r=1.0, csw=0.0, mu_t=0.0):
        self.L, self.V, self.U = L, L * L, _c(U)
        self.args = (mass, r, csw, mu_t)

    def apply(self, kind, src):
        src = _c(src); dst = np.zeros(2 * self.V, np.complex128)
        lib().ref_dirac_apply(self.L, _ptr(self.U), *self.args, kind, _ptr(src), _ptr(dst))
        return dst

    def delta(self, src, pi, dt):
        src = _c(src); pi = _c(pi, np.float64); dst = np.zeros(2 * self.V, np.complex128)
        lib().ref_dirac_delta(self.L, _ptr(self.U), *self.args, _ptr(src), _ptr(dst), _ptr(pi), dt)
        return dst

    def eo(self, kind, src):
        src = _c(src); dst = np.zeros(self.V, np.complex128)
        lib().ref_eo_apply(self.L, _ptr(self.U), *self.args, kind, _ptr(src), _ptr(dst))
        return dst

    def reconstruct(self, x_o, phi=None):
        x_o = _c(x_o); out = np.zeros(2 * self.V, np.complex128)
        lib().ref_eo_reconstruct(self.L, _ptr(self.U), *self.args, _ptr(x_o),
                                 _ptr(_c(phi)) if phi is not None else None, _ptr(out))
        return out

    def logdet(self):
        return lib().ref_eo_logdet(self.L, _ptr(self.U), *self.args)

    def galerkin(self, b, k, nulls, dense=False):
        nulls = _c(nulls)
        nb = (self.L // b) ** 2
        links = np.zeros(nb * 9 * k * k, np.complex128)
        dim = nb * k
        cols = np.zeros(dim * dim, np.complex128) if dense else None
        lib().ref_galerkin(self.L, _ptr(self.U), *self.args, b, k, _ptr(nulls), _ptr(links), _ptr(cols))
        return links, (cols.reshape(dim, dim) if dense else None)

    def mr_smooth(self, x, b, niter):
        x = _c(x).copy(); b = _c(b)
        lib().ref_mr_smooth(self.L, _ptr(self.U), *self.args, _ptr(x), _ptr(b), niter)
        return x

    def lambda_max(self, niter):
        return lib().ref_lambda_max(self.L, _ptr(self.U), *self.args, niter)

    def krylov(self, solver, rhs, maxit, tol, trunc=30):
        rhs = _c(rhs); x = np.zeros(2 * self.V, np.complex128); res = np.zeros(1)
        it = lib().ref_krylov(self.L, _ptr(self.U), *self.args, solver, _ptr(rhs), maxit, tol, trunc,
                              _ptr(x), _ptr(res))
        return x, it, res[0]

    def fgmres_ritz(self, rhs, maxit, tol, restart, n_ritz):
        rhs = _c(rhs); n = 2 * self.V
        x = np.zeros(n, np.complex128); res = np.zeros(1); nr = np.zeros(1, np.int32)
        cap = 2 * n_ritz + restart
        vals = np.zeros(cap); vecs = np.zeros((cap, n), np.complex128)
        it = lib().ref_fgmres_ritz(self.L, _ptr(self.U), *self.args, _ptr(rhs), maxit, tol, restart, n_ritz,
                                   _ptr(x), _ptr(res), _ptr(nr), _ptr(vals), _ptr(vecs))
        k = int(nr[0])
        return x, it, res[0], vals[:k].copy(), vecs[:k].copy()

    def mg(self, n_levels, block, k, coarse_block, null_iters, starts, pre, post, w_cycle, mode, b,
           maxit=100, tol=1e-10, trunc=20, sym_starts=None):
        starts = _c(starts); b = _c(b); x = np.zeros(2 * self.V, np.complex128); res = np.zeros(1)
        it = lib().ref_mg(self.L, _ptr(self.U), *self.args, n_levels, block, k, coarse_block, null_iters,
                          _ptr(starts), pre, post, int(w_cycle), mode, _ptr(b), _ptr(x), maxit, tol, trunc,
                          _ptr(_c(sym_starts)) if sym_starts is not None else None, _ptr(res))
        return x, it, res[0]

    def force(self, kind, X=None, chi=None):
        F = np.zeros(2 * self.V)
        lib().ref_fermion_force(self.L, _ptr(self.U), *self.args, kind,
                                _ptr(_c(X)) if X is not None else None,
                                _ptr(_c(chi)) if chi is not None else None, _ptr(F))
        return F

    def fermion_action(self, phi, maxit, tol):
        return lib().ref_fermion_action(self.L, _ptr(self.U), *self.args, _ptr(_c(phi)), maxit, tol)


def prolong_build(L, b, k, nulls):
    nulls = _c(nulls); bv = np.zeros((L // b) ** 2 * k * 2 * b * b, np.complex128)
    lib().ref_prolong_build(L, b, k, _ptr(nulls), _ptr(bv))
    return bv


def prolong_apply(L, b, k, nulls, direction, vin):
    nulls = _c(nulls); vin = _c(vin)
    n = 2 * L * L if direction == 0 else (L // b) ** 2 * k
    out = np.zeros(n, np.complex128)
    lib().ref_prolong_apply(L, b, k, _ptr(nulls), direction, _ptr(vin), _ptr(out))
    return out


def coarse_prolong_build(fine_dim, block_k, k, nulls):
    nulls = _c(nulls); bv = np.zeros(fine_dim // block_k * k * block_k, np.complex128)
    lib().ref_coarse_prolong_build(fine_dim, block_k, k, _ptr(nulls), _ptr(bv))
    return bv


def dense_solve(cols, b):
    cols = _c(cols); b = _c(b); x = np.zeros(b.size, np.complex128)
    lib().ref_dense_solve(b.size, _ptr(cols), _ptr(b), _ptr(x))
    return x


def gauge_force(L, U, beta):
    F = np.zeros(2 * L * L)
    lib().ref_gauge_force(L, _ptr(_c(U)), beta, _ptr(F))
    return F


def gauge_action(L, U, beta):
    return lib().ref_gauge_action(L, _ptr(_c(U)), beta)


def leapfrog_evolve(L, U, mass, r, csw, beta, tau, n_steps, omelyan, cg_maxit, cg_tol, pi, phi):
    """plain_leapfrog_evolve of the compiled reference: returns (U, pi, cg iterations)."""
    U = _c(U).copy()
    pi = np.ascontiguousarray(pi, dtype=np.float64).copy()
    phi = _c(phi)
    it = lib().ref_leapfrog_evolve(L, _ptr(U), mass, r, csw, beta, tau, n_steps, int(omelyan), cg_maxit, cg_tol,
                                   _ptr(pi), _ptr(phi))
    return U, pi, it


def gauge_save_raw(L, U, path):
    return lib().ref_gauge_save_raw(L, _ptr(_c(U)), path.encode())


def gauge_load_raw(L, path):
    U = np.zeros(2 * L * L, np.complex128)
    rc = lib().ref_gauge_load_raw(L, _ptr(U), path.encode())
    return U if rc == 0 else None


def gauge_save(L, U, beta, mass, path):
    return lib().ref_gauge_save(L, _ptr(_c(U)), beta, mass, path.encode())


def gauge_load(L, path):
    """(U, (beta, mass, avg_plaq)) or None when the reference refuses the file."""
    U = np.zeros(2 * L * L, np.complex128)
    h = np.zeros(3)
    rc = lib().ref_gauge_load(L, _ptr(U), _ptr(h), path.encode())
    return (U, tuple(h)) if rc == 0 else None


def hmc_draws(L, seed, n_eta):
    """The momenta and pseudofermion Gaussians hmc_trajectory draws from mt19937(seed)."""
    pi = np.zeros(2 * L * L)
    eta = np.zeros(n_eta, np.complex128)
    lib().ref_hmc_draws(L, seed, _ptr(pi), n_eta, _ptr(eta))
    return pi, eta


def hmc_trajectory(L, U, mass, r, csw, beta, tau, n_steps, omelyan, use_eo, cg_maxit, cg_tol, seed):
    """hmc_trajectory with force_accept: (final U, [dH, dKE, dSG, dSF, dLD, cg])."""
    U = _c(U).copy()
    out = np.zeros(6)
    lib().ref_hmc_trajectory(L, _ptr(U), mass, r, csw, beta, tau, n_steps, int(omelyan), int(use_eo), cg_maxit,
                             cg_tol, seed, _ptr(out))
    return U, out


def lanczos_eigen(A):
    """lanczos_eigen of a Hermitian matrix: (ascending eigenvalues, eigenvector columns)."""
    n = A.shape[0]
    cols = np.ascontiguousarray(np.asarray(A, np.complex128).T)
    ev = np.zeros(n)
    vecs = np.zeros((n, n), np.complex128)
    lib().ref_lanczos_eigen(n, _ptr(cols), _ptr(ev), _ptr(vecs))
    return ev, vecs


def trlm_dirac(L, U, m, r, csw, mu, n_ev, n_kr, max_restarts, tol):
    """trlm_eigensolver on D^dag D: (restarts or -1, eigenvalues, eigenvectors [n_ev][2V])."""
    ev = np.zeros(n_ev)
    vecs = np.zeros((n_ev, 2 * L * L), np.complex128)
    rs = lib().ref_trlm_dirac(L, _ptr(_c(U)), m, r, csw, mu, n_ev, n_kr, max_restarts, tol, _ptr(ev), _ptr(vecs))
    return rs, ev, vecs


def coarse_deflated_solve(L, U, m, r, csw, mu, b, k, nulls, n_ev, n_kr, max_restarts, tol, rhs, maxit, cg_tol):
    """TRLM deflation of the Galerkin coarse operator, then its deflated CG solve:
    (eigenvalues, x)."""
    ev = np.zeros(n_ev)
    x = np.zeros(len(rhs), np.complex128)
    nd = lib().ref_coarse_deflated_solve(L, _ptr(_c(U)), m, r, csw, mu, b, k, _ptr(_c(nulls)), n_ev, n_kr,
                                         max_restarts, tol, _ptr(_c(rhs)), maxit, cg_tol, _ptr(ev), _ptr(x))
    return ev[:nd], x


def cg_deflated(L, U, m, r, csw, mu, vecs, rhs, maxit, tol):
    """cg_solve_deflated on D^dag D, identity preconditioner: (x, iterations, final residual)."""
    vecs = np.ascontiguousarray(vecs, np.complex128)
    x = np.zeros(L * L * 2, np.complex128)
    res = np.zeros(1)
    it = lib().ref_cg_deflated(L, _ptr(_c(U)), m, r, csw, mu, vecs.shape[0], _ptr(vecs), _ptr(_c(rhs)), maxit, tol,
                               _ptr(x), _ptr(res))
    return x, it, res[0]


def lobpcg_dirac(L, U, m, r, csw, mu, X0, max_iter, tol):
    """lobpcg_update on D^dag D, identity preconditioner: (eigenvalues, X [k][2V], iterations)."""
    X0 = np.ascontiguousarray(X0, np.complex128)
    k = X0.shape[0]
    ev = np.zeros(k)
    X = np.zeros_like(X0)
    it = lib().ref_lobpcg_dirac(L, _ptr(_c(U)), m, r, csw, mu, k, _ptr(X0), max_iter, tol, _ptr(ev), _ptr(X))
    return ev, X, it


def hmc_ms_trajectory(L, U, mass, r, csw, beta, tau, n_outer, n_inner, omelyan, inner_smooth, defl_refresh,
                      cg_maxit, cg_tol, b, k, nulls, n_defl, seed):
    """hmc_trajectory_mg_multiscale (identity mg_precond) from mt19937(seed): (U after,
    out = [accepted, dH, dKE, dSG, dSF, cg, lowmode evals], (vals0, vecs0), (vals1, vecs1))."""
    U = _c(U).copy()
    nc = (L // b) ** 2 * k
    out = np.zeros(7)
    v0 = np.zeros(n_defl); V0 = np.zeros((n_defl, nc), np.complex128)
    v1 = np.zeros(n_defl); V1 = np.zeros((n_defl, nc), np.complex128)
    lib().ref_hmc_ms_trajectory(L, _ptr(U), mass, r, csw, beta, tau, n_outer, n_inner, omelyan, inner_smooth,
                                defl_refresh, cg_maxit, cg_tol, b, k, _ptr(_c(nulls)), n_defl, seed, _ptr(out),
                                _ptr(v0), _ptr(V0), _ptr(v1), _ptr(V1))
    return U, out, (v0, V0), (v1, V1)


def eo_mg(L, U, mass, r, csw, mu, bx, by, k, null_iters, pre, post, seed, b):
    """EvenOddMG built with mt19937(seed), then x = precondition(b): (x, omega,
    the dense coarse matrix M[i][j])."""
    V = L * L
    dim = (L // bx) * (L // by) * k
    x = np.zeros(V, np.complex128)
    om = np.zeros(1)
    cd = np.zeros(1, np.int32)
    cols = np.zeros((dim, dim), np.complex128)
    lib().ref_eo_mg(L, _ptr(_c(U)), mass, r, csw, mu, bx, by, k, null_iters, pre, post, seed, _ptr(_c(b)), _ptr(x),
                    _ptr(om), _ptr(cd), _ptr(cols))
    assert int(cd[0]) == dim
    return x, float(om[0]), cols.T.copy()

