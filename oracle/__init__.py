"""ctypes bindings for the oracle (liboracle.so) and the compiled reference (oracle/_ref/libref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, and nowhere else.  The product package
``schwinger_mg_amd`` never imports this module.

Arrays are numpy complex128 in the reference layouts: spinor ``psi[2*s+spin]``,
gauge ``U[mu*V + s]``, coarse vector ``c[b*k+j]``, forces ``F[mu*V + s]``.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double


def _ptr(a):
    return a.ctypes.data_as(_P) if a is not None else None


def _c(a, dt=np.complex128):
    return np.ascontiguousarray(a, dtype=dt)


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        L = ctypes.CDLL(path)
        L.orc_set_threads.argtypes = [_I]
        L.orc_get_threads.restype = _I
        L.orc_dirac_new.restype = _P
        L.orc_dirac_new.argtypes = [_I, _P, _D, _D, _D, _D]
        L.orc_dirac_free.argtypes = [_P]
        L.orc_dirac_clover.restype = ctypes.POINTER(_D)
        L.orc_dirac_clover.argtypes = [_P]
        L.orc_dirac_apply.argtypes = [_P, _I, _P, _P]
        L.orc_dirac_delta.argtypes = [_P, _P, _P, _P, _D]
        L.orc_eo_apply.argtypes = [_P, _I, _P, _P]
        L.orc_eo_reconstruct.argtypes = [_P, _P, _P, _P]
        L.orc_eo_reconstruct_full.argtypes = [_P, _P, _P]
        L.orc_eo_logdet.restype = _D
        L.orc_eo_logdet.argtypes = [_P]
        L.orc_clover.argtypes = [_I, _P, _P]
        L.orc_avg_plaq.restype = _D
        L.orc_avg_plaq.argtypes = [_I, _P]
        L.orc_gauge_action.restype = _D
        L.orc_gauge_action.argtypes = [_I, _P, _D]
        L.orc_kinetic_energy.restype = _D
        L.orc_kinetic_energy.argtypes = [_I, _P]
        L.orc_hmc_evolve.restype = _I
        L.orc_hmc_evolve.argtypes = [_I, _P, _D, _D, _D, _D, _D, _D, _I, _I, _I, _D, _P, _P]
        L.orc_hmc_nested_evolve.restype = _I
        L.orc_hmc_nested_evolve.argtypes = [_I, _P, _D, _D, _D, _D, _D, _D, _I, _I, _I, _I, _I, _D, _P, _P]
        L.orc_hmc_hamiltonian.restype = _I
        L.orc_hmc_hamiltonian.argtypes = [_I, _P, _D, _D, _D, _D, _D, _I, _D, _P, _P, _P]
        L.orc_lanczos_eigen.argtypes = [_I, _P, _P, _P]
        L.orc_rr_evolve.argtypes = [_P, _I, _P, _P]
        L.orc_coarse_lowmode_force.argtypes = [_P, _P, _I, _P, _P, _P, _I, _P]
        L.orc_hmc_ms_evolve.restype = _I
        L.orc_hmc_ms_evolve.argtypes = [_I, _P, _D, _D, _D, _D, _D, _D, _I, _I, _I, _I, _I, _I, _D, _P, _I, _P, _P,
                                        _P, _P, _P]
        L.orc_philox4x32.argtypes = [_P, _P, _P]
        L.orc_von_mises.restype = _D
        L.orc_von_mises.argtypes = [_D, ctypes.c_uint32, ctypes.c_uint32, _P]
        L.orc_heatbath.argtypes = [_I, _P, _D, _I, ctypes.c_uint64, ctypes.c_uint32]
        L.orc_lobpcg_ex.restype = _I
        L.orc_lobpcg_ex.argtypes = [_P, _P, _I, _P, _I, _D, _I, _P, _P]
        L.orc_lobpcg.restype = _I
        L.orc_lobpcg.argtypes = [_P, _P, _I, _P, _I, _D, _P, _P]
        L.orc_cg_deflated.restype = _I
        L.orc_cg_deflated.argtypes = [_P, _P, _P, _I, _P, _I, _D, _P, _P]
        L.orc_trlm.restype = _I
        L.orc_trlm.argtypes = [_P, _I, _I, _I, _D, _P, _P, _P, _P, _P]
        L.orc_deflated_coarse_cg.restype = _I
        L.orc_deflated_coarse_cg.argtypes = [_P, _I, _P, _P, _P, _I, _D, _P]
        L.orc_hmc_evolve_eo.restype = _I
        L.orc_hmc_evolve_eo.argtypes = [_I, _P, _D, _D, _D, _D, _D, _D, _I, _I, _D, _P, _P]
        L.orc_hmc_hamiltonian_eo.restype = _I
        L.orc_hmc_hamiltonian_eo.argtypes = [_I, _P, _D, _D, _D, _D, _D, _I, _D, _P, _P, _P]
        L.orc_gauge_force.argtypes = [_I, _P, _D, _P]
        L.orc_eo_maps.argtypes = [_I, _P, _P, _P]
        L.orc_dot2.argtypes = [_I, _P, _P, _P]
        L.orc_norm.restype = _D
        L.orc_norm.argtypes = [_I, _P]
        L.orc_op_dirac.restype = _P
        L.orc_op_dirac.argtypes = [_P, _I]
        L.orc_op_coarse.restype = _P
        L.orc_op_coarse.argtypes = [_P]
        L.orc_op_apply.argtypes = [_P, _P, _P]
        L.orc_op_dim.restype = _I
        L.orc_op_dim.argtypes = [_P]
        L.orc_op_free.argtypes = [_P]
        L.orc_prolong_new.restype = _P
        L.orc_prolong_new.argtypes = [_I, _I, _I, _I, _I, _I, _I]
        L.orc_prolong_free.argtypes = [_P]
        L.orc_prolong_build.argtypes = [_P, _P]
        L.orc_prolong_get.argtypes = [_P, _P]
        L.orc_prolong_set.argtypes = [_P, _P]
        L.orc_prolong_apply.argtypes = [_P, _P, _P]
        L.orc_restrict_apply.argtypes = [_P, _P, _P]
        L.orc_prolong_coarse_dim.restype = _I
        L.orc_prolong_coarse_dim.argtypes = [_P]
        L.orc_coarse_build.restype = _P
        L.orc_coarse_build.argtypes = [_P, _P, _I]
        L.orc_coarse_free.argtypes = [_P]
        L.orc_coarse_get.argtypes = [_P, _P]
        L.orc_coarse_set.argtypes = [_P, _P]
        L.orc_dense_build.restype = _P
        L.orc_dense_build.argtypes = [_P, _P]
        L.orc_dense_free.argtypes = [_P]
        L.orc_dense_get.argtypes = [_P, _P]
        L.orc_dense_solve.argtypes = [_P, _P, _P]
        L.orc_dense_apply.argtypes = [_P, _P, _P]
        L.orc_mr_smooth.argtypes = [_P, _P, _P, _I]
        L.orc_richardson.argtypes = [_P, _P, _P, _I, _D]
        L.orc_lambda_max.restype = _D
        L.orc_lambda_max.argtypes = [_P, _P, _I]
        L.orc_null_space.argtypes = [_P, _I, _I, _P, _P]
        L.orc_mg_new.restype = _P
        L.orc_mg_new.argtypes = [_P, _I, _I, _I, _P, _P]
        L.orc_mg_assemble.restype = _P
        L.orc_mg_assemble.argtypes = [_P, _I, _I, _I, _P, _P, _P]
        L.orc_mg_free.argtypes = [_P]
        L.orc_mg_set_symmetric.argtypes = [_P, _D, _P]
        L.orc_mg_precondition.argtypes = [_P, _P, _P]
        L.orc_mg_prolong.restype = _P
        L.orc_mg_prolong.argtypes = [_P, _I]
        L.orc_mg_coarse.restype = _P
        L.orc_mg_coarse.argtypes = [_P, _I]
        L.orc_mg_omega.restype = _D
        L.orc_mg_omega.argtypes = [_P, _I]
        L.orc_mg_null_vecs.argtypes = [_P, _I, _P]
        L.orc_cg.restype = _I
        L.orc_cg.argtypes = [_P, _P, _P, _P, _I, _D, _P, _P]
        L.orc_fcg.restype = _I
        L.orc_fcg.argtypes = [_P, _P, _P, _I, _D, _I, _I, _P, _P]
        L.orc_fgmres.restype = _I
        L.orc_fgmres.argtypes = [_P, _P, _P, _I, _I, _D, _P, _P]
        L.orc_fgmres_ritz.restype = _I
        L.orc_fgmres_ritz.argtypes = [_P, _P, _P, _I, _I, _D, _I, _P, _P, _P, _P, _P]
        L.orc_coarse_cg.restype = _I
        L.orc_coarse_cg.argtypes = [_P, _P, _I, _D, _P]
        L.orc_fermion_force.argtypes = [_P, _P, _P]
        L.orc_fermion_force_bilinear.argtypes = [_P, _P, _P, _P]
        L.orc_hopping_force.argtypes = [_P, _P, _P, _P]
        L.orc_clover_deriv_insert.argtypes = [_P, _P, _D, _P]
        L.orc_logdet_ee_force.argtypes = [_P, _P]
        L.orc_eo_fermion_force.argtypes = [_P, _P, _P, _P]
        _lib = L
    return _lib


class MGParams(ctypes.Structure):
    _fields_ = [("n_levels", _I), ("bx", _I * 4), ("by", _I * 4), ("k", _I * 4),
                ("null_iters", _I), ("pre_smooth", _I), ("post_smooth", _I), ("w_cycle", _I),
                ("coarse_solver", _I), ("coarse_maxit", _I), ("coarse_tol", _D),
                ("colored_build", _I), ("chiral", _I)]


D_, DDAG, DDAGD, G5D = 0, 1, 2, 3
SCHUR_NORMAL = 100


def set_threads(n):
    """The OpenMP thread count of the oracle (and of anything sharing its runtime);
    returns the count now in effect."""
    lib().orc_set_threads(n)
    return lib().orc_get_threads()


def eo_maps(L):
    V = L * L
    e = np.zeros(V // 2, np.int32); o = np.zeros(V // 2, np.int32); i = np.zeros(V, np.int32)
    lib().orc_eo_maps(L, _ptr(e), _ptr(o), _ptr(i))
    return e, o, i


def clover(L, U):
    U = _c(U); F = np.zeros(L * L)
    lib().orc_clover(L, _ptr(U), _ptr(F))
    return F


def avg_plaq(L, U):
    U = _c(U)
    return lib().orc_avg_plaq(L, _ptr(U))


def gauge_action(L, U, beta):
    U = _c(U)
    return lib().orc_gauge_action(L, _ptr(U), beta)


def gauge_force(L, U, beta):
    U = _c(U); F = np.zeros(2 * L * L)
    lib().orc_gauge_force(L, _ptr(U), beta, _ptr(F))
    return F


def dot(a, b):
    a = _c(a); b = _c(b); out = np.zeros(2)
    lib().orc_dot2(a.size, _ptr(a), _ptr(b), _ptr(out))
    return complex(out[0], out[1])


class Dirac:
    """The reference's DiracOp (src/dirac.h:8-183) and EvenOddDiracOp (:190-553)."""

    def __init__(self, L, U, mass, r=1.0, csw=0.0, mu_t=0.0):
        self.L, self.V = L, L * L
        self.U = _c(U)
        self.h = lib().orc_dirac_new(L, _ptr(self.U), mass, r, csw, mu_t)
        self.mass, self.r, self.csw, self.mu_t = mass, r, csw, mu_t

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_dirac_free(self.h)
            self.h = None

    def clover(self):
        p = lib().orc_dirac_clover(self.h)
        return np.ctypeslib.as_array(p, shape=(self.V,)).copy()

    def apply(self, kind, src):
        src = _c(src); dst = np.zeros(2 * self.V, np.complex128)
        lib().orc_dirac_apply(self.h, kind, _ptr(src), _ptr(dst))
        return dst

    def delta(self, src, pi, dt):
        src = _c(src); pi = _c(pi, np.float64); dst = np.zeros(2 * self.V, np.complex128)
        lib().orc_dirac_delta(self.h, _ptr(src), _ptr(dst), _ptr(pi), dt)
        return dst

    def eo(self, kind, src):
        src = _c(src); dst = np.zeros(self.V, np.complex128)
        lib().orc_eo_apply(self.h, kind, _ptr(src), _ptr(dst))
        return dst

    def reconstruct(self, x_o, phi=None):
        x_o = _c(x_o); out = np.zeros(2 * self.V, np.complex128)
        if phi is None:
            lib().orc_eo_reconstruct_full(self.h, _ptr(x_o), _ptr(out))
        else:
            phi = _c(phi)
            lib().orc_eo_reconstruct(self.h, _ptr(x_o), _ptr(phi), _ptr(out))
        return out

    def logdet(self):
        return lib().orc_eo_logdet(self.h)

    def force(self, X):
        X = _c(X); F = np.zeros(2 * self.V)
        lib().orc_fermion_force(self.h, _ptr(X), _ptr(F))
        return F

    def force_bilinear(self, chi, X):
        chi = _c(chi); X = _c(X); F = np.zeros(2 * self.V)
        lib().orc_fermion_force_bilinear(self.h, _ptr(chi), _ptr(X), _ptr(F))
        return F

    def hopping_force(self, chi, X):
        chi = _c(chi); X = _c(X); F = np.zeros(2 * self.V)
        lib().orc_hopping_force(self.h, _ptr(chi), _ptr(X), _ptr(F))
        return F

    def logdet_force(self):
        F = np.zeros(2 * self.V)
        lib().orc_logdet_ee_force(self.h, _ptr(F))
        return F

    def eo_force(self, x_o, y_o):
        x_o = _c(x_o); y_o = _c(y_o); F = np.zeros(2 * self.V)
        lib().orc_eo_fermion_force(self.h, _ptr(x_o), _ptr(y_o), _ptr(F))
        return F

    def op(self, kind=DDAGD):
        return Op(lib().orc_op_dirac(self.h, kind), keep=self)


class Op:
    def __init__(self, h, keep=None, owned=True):
        self.h, self.keep, self.owned = h, keep, owned
        self.n = lib().orc_op_dim(h)

    def __del__(self):
        if self.owned and getattr(self, "h", None):
            lib().orc_op_free(self.h)
            self.h = None

    def __call__(self, src):
        src = _c(src); dst = np.zeros(self.n, np.complex128)
        lib().orc_op_apply(self.h, _ptr(src), _ptr(dst))
        return dst


class Prolongator:
    def __init__(self, nx, ny, nd, bx, by, k, chiral=False):
        self.h = lib().orc_prolong_new(nx, ny, nd, bx, by, k, int(chiral))
        self.chiral = bool(chiral)
        self.nx, self.ny, self.nd, self.bx, self.by, self.k = nx, ny, nd, bx, by, k
        self.nblocks = (nx // bx) * (ny // by)
        self.coarse_dim = self.nblocks * k
        self.fine_dim = nx * ny * nd

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_prolong_free(self.h)
            self.h = None

    def build(self, null_vecs):
        nv = _c(null_vecs)
        lib().orc_prolong_build(self.h, _ptr(nv))
        return self

    def block_vecs(self):
        out = np.zeros(self.nblocks * self.k * self.bx * self.by * self.nd, np.complex128)
        lib().orc_prolong_get(self.h, _ptr(out))
        return out

    def set_block_vecs(self, bv):
        bv = _c(bv)
        lib().orc_prolong_set(self.h, _ptr(bv))

    def prolong(self, c):
        c = _c(c); f = np.zeros(self.fine_dim, np.complex128)
        lib().orc_prolong_apply(self.h, _ptr(c), _ptr(f))
        return f

    def restrict(self, f):
        f = _c(f); c = np.zeros(self.coarse_dim, np.complex128)
        lib().orc_restrict_apply(self.h, _ptr(f), _ptr(c))
        return c


class Coarse:
    def __init__(self, op, P, colored=1):
        self.h = lib().orc_coarse_build(op.h, P.h, colored)
        self.k, self.nb = P.k, P.nblocks
        self.keep = (op, P)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_coarse_free(self.h)
            self.h = None

    def links(self):
        out = np.zeros(self.nb * 9 * self.k * self.k, np.complex128)
        lib().orc_coarse_get(self.h, _ptr(out))
        return out

    def op(self):
        return Op(lib().orc_op_coarse(self.h), keep=self)


def dense_galerkin(op, P=None):
    h = lib().orc_dense_build(op.h, P.h if P is not None else None)
    dim = P.coarse_dim if P is not None else op.n
    cols = np.zeros(dim * dim, np.complex128)
    lib().orc_dense_get(h, _ptr(cols))
    lib().orc_dense_free(h)
    return cols.reshape(dim, dim)   # cols[j] = column j


def dense_solve_galerkin(op, P, b):
    """CoarseOp::build + CoarseOp::solve (dense LU with partial pivoting)."""
    h = lib().orc_dense_build(op.h, P.h)
    x = np.zeros(P.coarse_dim, np.complex128)
    lib().orc_dense_solve(h, _ptr(_c(b)), _ptr(x))
    lib().orc_dense_free(h)
    return x


def mr_smooth(op, x, b, niter):
    x = _c(x).copy(); b = _c(b)
    lib().orc_mr_smooth(op.h, _ptr(x), _ptr(b), niter)
    return x


def richardson(op, x, b, niter, omega):
    x = _c(x).copy(); b = _c(b)
    lib().orc_richardson(op.h, _ptr(x), _ptr(b), niter, omega)
    return x


def lambda_max(op, v0, niter):
    v0 = _c(v0)
    return lib().orc_lambda_max(op.h, _ptr(v0), niter)


def null_space(op, k, outer_iters, starts):
    starts = _c(starts); out = np.zeros(k * op.n, np.complex128)
    lib().orc_null_space(op.h, k, outer_iters, _ptr(starts), _ptr(out))
    return out.reshape(k, op.n)


class MG:
    def __init__(self, fine_op, nx, ny, nd, starts, n_levels=2, block=((4, 4),), k=(4,),
                 null_iters=20, pre=3, post=3, w_cycle=True, coarse_solver=0,
                 coarse_maxit=200, coarse_tol=1e-12, colored=1, chiral=False, parts=None):
        """parts = (block_vecs, links): the per-level block vectors and coarse links,
        each concatenated over levels, to assemble the hierarchy from instead of
        running the setup (starts is then unused)."""
        p = MGParams()
        p.n_levels = n_levels
        for i, (bx, by) in enumerate(block):
            p.bx[i], p.by[i] = bx, by
        for i, kk in enumerate(k):
            p.k[i] = kk
        p.null_iters, p.pre_smooth, p.post_smooth = null_iters, pre, post
        p.w_cycle = int(w_cycle)
        p.coarse_solver, p.coarse_maxit, p.coarse_tol = coarse_solver, coarse_maxit, coarse_tol
        p.colored_build = colored
        p.chiral = int(chiral)
        self.prm = p
        self.fine = fine_op
        self.nx, self.ny = nx, ny
        if parts is None:
            self.starts = _c(starts)
            self.h = lib().orc_mg_new(fine_op.h, nx, ny, nd, ctypes.byref(p), _ptr(self.starts))
        else:
            self.parts = (_c(parts[0]), _c(parts[1]))
            self.h = lib().orc_mg_assemble(fine_op.h, nx, ny, nd, ctypes.byref(p), _ptr(self.parts[0]),
                                           _ptr(self.parts[1]))
        self.n = fine_op.n

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_mg_free(self.h)
            self.h = None

    def set_symmetric(self, damping, starts):
        s = _c(starts)
        lib().orc_mg_set_symmetric(self.h, damping, _ptr(s))

    def __call__(self, b):
        b = _c(b); x = np.zeros(self.n, np.complex128)
        lib().orc_mg_precondition(self.h, _ptr(b), _ptr(x))
        return x

    def omega(self, level):
        return lib().orc_mg_omega(self.h, level)

    def _shape(self, level):
        """(nblocks, k, block size) of transfer `level`."""
        nx, ny, nd = self.nx, self.ny, 2
        for l in range(level + 1):
            bx, by, k = self.prm.bx[l], self.prm.by[l], self.prm.k[l]
            if l < level:
                nx, ny, nd = nx // bx, ny // by, k
        return (nx // bx) * (ny // by), k, bx * by * nd

    def block_vecs(self, level):
        nb, k, bs = self._shape(level)
        out = np.zeros(nb * k * bs, np.complex128)
        lib().orc_prolong_get(lib().orc_mg_prolong(self.h, level), _ptr(out))
        return out

    def coarse_links(self, level):
        nb, k, _ = self._shape(level)
        out = np.zeros(nb * 9 * k * k, np.complex128)
        lib().orc_coarse_get(lib().orc_mg_coarse(self.h, level), _ptr(out))
        return out


def cg(op, rhs, maxit, tol, x0=None, mg=None):
    rhs = _c(rhs); x = np.zeros(op.n, np.complex128); res = np.zeros(1)
    x0p = _ptr(_c(x0)) if x0 is not None else None
    it = lib().orc_cg(op.h, _ptr(rhs), x0p, mg.h if mg else None, maxit, tol, _ptr(x), _ptr(res))
    return x, it, res[0]


def gcr(op, rhs, maxit, tol, trunc=30, mg=None, ref_projection=True):
    """fcg_solve (src/solvers.cpp:223-300); ref_projection=False uses the textbook
    GCR coefficient <AP_i, Az> instead of the reference's conjugate."""
    rhs = _c(rhs); x = np.zeros(op.n, np.complex128); res = np.zeros(1)
    it = lib().orc_fcg(op.h, _ptr(rhs), mg.h if mg else None, maxit, tol, trunc, int(ref_projection),
                       _ptr(x), _ptr(res))
    return x, it, res[0]


def fgmres(op, rhs, restart, maxit, tol, mg=None):
    rhs = _c(rhs); x = np.zeros(op.n, np.complex128); res = np.zeros(1)
    it = lib().orc_fgmres(op.h, _ptr(rhs), mg.h if mg else None, restart, maxit, tol, _ptr(x), _ptr(res))
    return x, it, res[0]


def fgmres_ritz(op, rhs, restart, maxit, tol, n_ritz, mg=None):
    """fgmres_solve_generic (src/solvers.cpp:555-772) with n_ritz_harvest = n_ritz:
    (x, iterations, residual, ritz values, ritz vectors [k][n])."""
    rhs = _c(rhs); x = np.zeros(op.n, np.complex128); res = np.zeros(1); nr = np.zeros(1, np.int32)
    cap = 2 * max(n_ritz, 1) + restart
    vals = np.zeros(cap); vecs = np.zeros((cap, op.n), np.complex128)
    it = lib().orc_fgmres_ritz(op.h, _ptr(rhs), mg.h if mg else None, restart, maxit, tol, n_ritz,
                               _ptr(x), _ptr(res), _ptr(nr), _ptr(vals), _ptr(vecs))
    k = int(nr[0])
    return x, it, res[0], vals[:k].copy(), vecs[:k].copy()


def coarse_cg(op, b, maxit, tol):
    b = _c(b); x = np.zeros(op.n, np.complex128)
    it = lib().orc_coarse_cg(op.h, _ptr(b), maxit, tol, _ptr(x))
    return x, it


def clover_deriv_insert(D, w, factor, F):
    w = _c(w, np.float64); F = _c(F, np.float64).copy()
    lib().orc_clover_deriv_insert(D.h, _ptr(w), factor, _ptr(F))
    return F


def hmc_nested_evolve(L, U, mass, r, csw, mu_t, beta, tau, n_outer, n_gauge, omelyan_outer, omelyan_inner,
                      cg_maxit, cg_tol, pi, phi):
    """The nested integrator with the gauge force on the fine scale: (U, pi, CG iterations)."""
    U = _c(U).copy()
    pi = np.ascontiguousarray(pi, dtype=np.float64).copy()
    phi = _c(phi)
    it = lib().orc_hmc_nested_evolve(L, _ptr(U), mass, r, csw, mu_t, beta, tau, n_outer, n_gauge, int(omelyan_outer),
                                     int(omelyan_inner), cg_maxit, cg_tol, _ptr(pi), _ptr(phi))
    return U, pi, it


def hmc_evolve(L, U, mass, r, csw, mu_t, beta, tau, n_steps, omelyan, cg_maxit, cg_tol, pi, phi):
    """plain_leapfrog_evolve restated: returns (U, pi, cg iterations) after the MD trajectory."""
    U = _c(U).copy()
    pi = np.ascontiguousarray(pi, dtype=np.float64).copy()
    phi = _c(phi)
    it = lib().orc_hmc_evolve(L, _ptr(U), mass, r, csw, mu_t, beta, tau, n_steps, int(omelyan), cg_maxit, cg_tol,
                              _ptr(pi), _ptr(phi))
    return U, pi, it


def hmc_hamiltonian(L, U, mass, r, csw, mu_t, beta, cg_maxit, cg_tol, pi, phi):
    """(KE, S_g, S_f) of the trajectory's Hamiltonian."""
    U, phi = _c(U), _c(phi)
    pi = np.ascontiguousarray(pi, dtype=np.float64)
    out = np.zeros(3)
    lib().orc_hmc_hamiltonian(L, _ptr(U), mass, r, csw, mu_t, beta, cg_maxit, cg_tol, _ptr(pi), _ptr(phi), _ptr(out))
    return out


def hmc_evolve_eo(L, U, mass, r, csw, mu_t, beta, tau, n_steps, cg_maxit, cg_tol, pi, phi_o):
    """The even-odd trajectory's leapfrog: returns (U, pi, cg iterations)."""
    U = _c(U).copy()
    pi = np.ascontiguousarray(pi, dtype=np.float64).copy()
    it = lib().orc_hmc_evolve_eo(L, _ptr(U), mass, r, csw, mu_t, beta, tau, n_steps, cg_maxit, cg_tol, _ptr(pi),
                                 _ptr(_c(phi_o)))
    return U, pi, it


def hmc_hamiltonian_eo(L, U, mass, r, csw, mu_t, beta, cg_maxit, cg_tol, pi, phi_o):
    """(KE, S_g, S_f, LD) of the even-odd trajectory."""
    out = np.zeros(4)
    lib().orc_hmc_hamiltonian_eo(L, _ptr(_c(U)), mass, r, csw, mu_t, beta, cg_maxit, cg_tol,
                                 _ptr(np.ascontiguousarray(pi, dtype=np.float64)), _ptr(_c(phi_o)), _ptr(out))
    return out


def lanczos_eigen(A):
    """Dense Hermitian eigensolver of lanczos_eigen: (ascending eigenvalues, eigenvector columns)."""
    n = A.shape[0]
    cols = np.ascontiguousarray(np.asarray(A, np.complex128).T)
    ev = np.zeros(n)
    vecs = np.zeros((n, n), np.complex128)
    lib().orc_lanczos_eigen(n, _ptr(cols), _ptr(ev), _ptr(vecs))
    return ev, vecs


def trlm(op, n_ev, n_kr, max_restarts, tol, start):
    """Thick-restart Lanczos: (converged, restarts, eigenvalues, eigenvectors [n_ev][n],
    operator applications)."""
    ev = np.zeros(n_ev)
    vecs = np.zeros((n_ev, op.n), np.complex128)
    rs, its = ctypes.c_int(), ctypes.c_int()
    conv = lib().orc_trlm(op.h, n_ev, n_kr, max_restarts, tol, _ptr(_c(start)), _ptr(ev), _ptr(vecs),
                          ctypes.byref(rs), ctypes.byref(its))
    return bool(conv), rs.value, ev, vecs, its.value


def deflated_coarse_cg(op, vecs, vals, b, maxit, tol):
    """SparseCoarseOp::solve with deflation vectors: (x, iterations)."""
    x = np.zeros(op.n, np.complex128)
    vecs = np.ascontiguousarray(vecs, np.complex128)
    vals = np.ascontiguousarray(vals, np.float64)
    it = lib().orc_deflated_coarse_cg(op.h, len(vals), _ptr(vecs), _ptr(vals), _ptr(_c(b)), maxit, tol, _ptr(x))
    return x, it


def cg_deflated(op, b, U, maxit, tol, mg=None):
    """cg_solve_deflated with the deflation vectors U [ndefl][n] and the MG cycle (or
    none) as preconditioner: (x, iterations, final residual)."""
    x = np.zeros(op.n, np.complex128)
    U = np.ascontiguousarray(U, np.complex128).reshape(-1, op.n)
    res = np.zeros(1)
    it = lib().orc_cg_deflated(op.h, _ptr(_c(b)), mg.h if mg is not None else None, U.shape[0], _ptr(U), maxit,
                               tol, _ptr(x), _ptr(res))
    return x, it, res[0]


def lobpcg(op, X0, max_iter, tol, mg=None, exact_ax=False):
    """lobpcg_update from the k start vectors X0 [k][n]: (eigenvalues, X [k][n], iterations);
    exact_ax keeps AX = A X through the re-orthonormalisation (the reference does not)."""
    X0 = np.ascontiguousarray(X0, np.complex128).reshape(-1, op.n)
    k = X0.shape[0]
    X = np.zeros_like(X0)
    ev = np.zeros(k)
    it = lib().orc_lobpcg_ex(op.h, mg.h if mg is not None else None, k, _ptr(X0), max_iter, tol, int(exact_ax),
                             _ptr(X), _ptr(ev))
    return ev, X, it


def rr_evolve(op, V):
    """rr_evolve: (rotated vectors [k][n], ascending Ritz values)."""
    V = np.ascontiguousarray(V, np.complex128).copy()
    vals = np.zeros(V.shape[0])
    lib().orc_rr_evolve(op.h, V.shape[0], _ptr(V), _ptr(vals))
    return V, vals


def coarse_lowmode_force(D, P, V, vals, phi, smooth):
    """coarse_lowmode_force: F[2V] (F[mu*V + s])."""
    V = np.ascontiguousarray(V, np.complex128)
    F = np.zeros(2 * D.L * D.L)
    lib().orc_coarse_lowmode_force(D.h, P.h, V.shape[0], _ptr(V), _ptr(np.ascontiguousarray(vals, np.float64)),
                                   _ptr(_c(phi)), smooth, _ptr(F))
    return F


def hmc_ms_evolve(L, U, mass, r, csw, mu_t, beta, tau, n_outer, n_inner, omelyan, inner_smooth, defl_refresh,
                  cg_maxit, cg_tol, P, V, vals, pi, phi):
    """The multi-timescale MD: (U, pi, V, vals, outer CG iterations, low-mode evaluations)."""
    U = _c(U).copy()
    pi = np.ascontiguousarray(pi, np.float64).copy()
    V = np.ascontiguousarray(V, np.complex128).copy()
    vals = np.ascontiguousarray(vals, np.float64).copy()
    nl = ctypes.c_int()
    it = lib().orc_hmc_ms_evolve(L, _ptr(U), mass, r, csw, mu_t, beta, tau, n_outer, n_inner, omelyan, inner_smooth,
                                 defl_refresh, cg_maxit, cg_tol, P.h, V.shape[0], _ptr(V), _ptr(vals), _ptr(pi),
                                 _ptr(_c(phi)), ctypes.byref(nl))
    return U, pi, V, vals, it, nl.value


def philox4x32(ctr, key):
    """Philox4x32-10 of the 4-word counter and 2-word key."""
    c = np.ascontiguousarray(ctr, np.uint32); k = np.ascontiguousarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    lib().orc_philox4x32(_ptr(c), _ptr(k), _ptr(out))
    return out


def von_mises(kappa, c0, c1, seed):
    k = np.array([seed & 0xFFFFFFFF, seed >> 32], np.uint32)
    return lib().orc_von_mises(kappa, c0, c1, _ptr(k))


def heatbath(L, U, beta, n_sweeps, seed, first_sweep=0):
    """The checkerboard U(1) heatbath: the new links."""
    U = _c(U).copy()
    lib().orc_heatbath(L, _ptr(U), beta, n_sweeps, seed, first_sweep)
    return U
