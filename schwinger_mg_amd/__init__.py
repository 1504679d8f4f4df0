"""schwinger_mg_amd -- MI355X (gfx950) Wilson-clover operator, adaptive multigrid and
HMC fermion force for the 2D Schwinger model.

The compute path is libsmg.so (hand-written HIP kernels behind the C-ABI in
include/smg.h).  This module is the Python host mirror of the reference's C++
operator API: ``DiracOp`` and ``EvenOddDiracOp`` (src/dirac.h), ``Prolongator``
(src/prolongator.h), ``SparseCoarseOp`` (src/coarse_op.h), ``MGHierarchy`` /
``build_mg_hierarchy`` (src/multigrid.h), the smoothers (src/smoother.h), the
Krylov entry points (src/solvers.h) and the force routines (src/hmc.h), with the
same names, argument meaning and layouts.  Vectors are numpy complex128 arrays
in the reference layouts; every call runs on the GPU.  There is no CPU fallback:
if libsmg.so is missing or no GPU is present the first call raises.
"""
import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsmg.so")

F64, F32, F16, BF16V = 0, 1, 2, 3
D_, DDAG, DDAGD, G5D = 0, 1, 2, 3
OP_DDAGD, OP_SCHUR = 0, 1
SMOOTH_MR, SMOOTH_RICHARDSON = 0, 1
EO_HOP_EO, EO_HOP_OE, EO_EE_INV, EO_OO_INV, EO_OO, EO_SCHUR, EO_SCHUR_DAG, EO_SCHUR_DAG_SCHUR = range(8)

_P, _I, _D, _SZ = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_size_t


class SMGError(RuntimeError):
    pass


class MGParams(ctypes.Structure):
    _fields_ = [("n_levels", _I), ("bx", _I * 4), ("by", _I * 4), ("k", _I * 4), ("null_iters", _I),
                ("pre_smooth", _I), ("post_smooth", _I), ("w_cycle", _I), ("coarse_maxit", _I),
                ("coarse_tol", _D), ("coarse_half", _I), ("coarse_solver", _I), ("use_graph", _I),
                ("chiral", _I), ("null_prec", _I)]


class SolveInfo(ctypes.Structure):
    _fields_ = [("iterations", _I), ("final_residual", _D), ("true_residual", _D), ("restarts", _I)]


class GaugeHeader(ctypes.Structure):
    """GaugeHeader (src/hmc.h:19-25)."""
    _fields_ = [("magic", ctypes.c_uint32), ("L", ctypes.c_uint32), ("beta", _D), ("mass", _D), ("avg_plaq", _D)]


class HMCParams(ctypes.Structure):
    """HMCParams (src/hmc.h:107-119), full lattice without preconditioner."""
    _fields_ = [("beta", _D), ("tau", _D), ("n_steps", _I), ("omelyan", _I), ("cg_maxiter", _I), ("cg_tol", _D),
                ("force_accept", _I), ("use_eo", _I)]


class HMCResult(ctypes.Structure):
    """HMCResult (src/hmc.h:120-128)."""
    _fields_ = [("accepted", _I), ("dH", _D), ("dKE", _D), ("dSG", _D), ("dSF", _D), ("total_cg_iters", _I),
                ("dLD", _D), ("max_true_residual", _D)]


# (name, restype, argtypes) of every entry point declared in include/smg.h
SIGNATURES = [
    ("smg_init", _I, [_I]), ("smg_last_error", ctypes.c_char_p, []), ("smg_version", _I, []),
    ("smg_synchronize", _I, []), ("smg_get_stream", _I, [_P]),
    ("smg_eo_maps", _I, [_I, _P, _P, _P]), ("smg_avg_plaq", _I, [_I, _P, _P]),
    ("smg_gauge_action", _I, [_I, _P, _D, _P]), ("smg_gauge_force", _I, [_I, _P, _D, _P]),
    ("smg_dirac_create", _I, [_P, _I, _P, _D, _D, _D, _D]), ("smg_dirac_destroy", _I, [_P]),
    ("smg_dirac_set_gauge", _I, [_P, _P]), ("smg_dirac_clover", _I, [_P, _P]),
    ("smg_dirac_apply", _I, [_P, _I, _P, _P]), ("smg_dirac_apply_dev", _I, [_P, _I, _I, _P, _P, _I, _P]),
    ("smg_dirac_delta", _I, [_P, _P, _P, _P, _D]),
    ("smg_eo_apply", _I, [_P, _I, _P, _P]), ("smg_eo_apply_dev", _I, [_P, _I, _I, _P, _P, _P]),
    ("smg_eo_reconstruct", _I, [_P, _P, _P, _P]), ("smg_eo_logdet", _I, [_P, _P]),
    ("smg_dot_dev", _I, [_P, _P, _SZ, _I, _P, _P]), ("smg_norm_dev", _I, [_P, _SZ, _I, _P, _P]),
    ("smg_cg_solve_op", _I, [_P, _P, _P, _P, _P, _I, _D, _P]),
    ("smg_gcr_solve_op", _I, [_P, _P, _P, _P, _I, _D, _I, _I, _P]),
    ("smg_fgmres_solve_op", _I, [_P, _P, _P, _P, _I, _I, _D, _I, _P, _P, _P, _P]),
    ("smg_smooth_op", _I, [_P, _I, _P, _P, _I, _D]), ("smg_lambda_max_op", _I, [_P, _P, _I, _P]),
    ("smg_coarse_create_galerkin_op", _I, [_P, _P, _P]),
    ("smg_mg_coarse_apply_dev", _I, [_P, _I, _I, _I, _P, _P, _I, _P]),
    ("smg_tune_mg", _I, [_P, _I, _P]), ("smg_tune_save", _I, [ctypes.c_char_p]),
    ("smg_tune_load", _I, [ctypes.c_char_p]), ("smg_tune_clear", _I, []), ("smg_tune_get", _I, [_I, _P, _P]),
    ("smg_tune_set", _I, [_I, _P, _I]),
    ("smg_dot_dev_noalloc", _I, [_P, _P, _SZ, _I, _P, _P]),
    ("smg_prolongator_create", _I, [_P, _I, _I, _I, _I, _I, _I, _I]), ("smg_prolongator_destroy", _I, [_P]),
    ("smg_prolongator_build", _I, [_P, _P]), ("smg_prolongator_get", _I, [_P, _P]),
    ("smg_prolongator_set", _I, [_P, _P]), ("smg_prolong", _I, [_P, _P, _P]), ("smg_restrict", _I, [_P, _P, _P]),
    ("smg_coarse_create_galerkin", _I, [_P, _P, _P]), ("smg_coarse_create_from_coarse", _I, [_P, _P, _P]),
    ("smg_coarse_create_links", _I, [_P, _I, _I, _I, _P]), ("smg_coarse_destroy", _I, [_P]),
    ("smg_coarse_get_links", _I, [_P, _P]), ("smg_coarse_apply", _I, [_P, _I, _P, _P]),
    ("smg_coarse_solve", _I, [_P, _P, _P, _I, _D, _P]),
    ("smg_smooth_dirac", _I, [_P, _I, _P, _P, _I, _D]), ("smg_smooth_coarse", _I, [_P, _I, _P, _P, _I, _D]),
    ("smg_lambda_max_dirac", _I, [_P, _P, _I, _P]),
    ("smg_mg_setup", _I, [_P, _P, _P, _P]), ("smg_mg_destroy", _I, [_P]),
    ("smg_mg_set_symmetric", _I, [_P, _D, _P]), ("smg_mg_precondition", _I, [_P, _I, _P, _P]),
    ("smg_mg_precondition_dev", _I, [_P, _I, _P, _P, _P]), ("smg_mg_level_dim", _I, [_P, _I, _P]),
    ("smg_mg_null_vecs", _I, [_P, _I, _P]), ("smg_mg_block_vecs", _I, [_P, _I, _P]),
    ("smg_mg_coarse_links", _I, [_P, _I, _P]), ("smg_mg_omega", _I, [_P, _I, _P]),
    ("smg_mg_set_cycle", _I, [_P, _I, _I, _I, _D, _I]),
    ("smg_cg_solve", _I, [_P, _I, _P, _P, _P, _P, _I, _D, _P]),
    ("smg_gcr_solve", _I, [_P, _P, _I, _P, _P, _I, _D, _I, _I, _D, _P]),
    ("smg_gcr_solve_dev", _I, [_P, _P, _I, _P, _P, _I, _D, _I, _D, _P, _P]),
    ("smg_fgmres_solve", _I, [_P, _P, _P, _P, _I, _I, _D, _P]),
    ("smg_fgmres_solve_ritz", _I, [_P, _P, _P, _P, _I, _I, _D, _I, _P, _P, _P, _P]),
    ("smg_dirac_apply_batched_dev", _I, [_P, _I, _I, _P, _P, _I, _P]),
    ("smg_coarse_apply_batched", _I, [_P, _I, _I, _P, _P]),
    ("smg_coarse_dense", _I, [_P, _P]), ("smg_coarse_dense_solve", _I, [_P, _P, _P]),
    ("smg_coarse_apply_batched_dev", _I, [_P, _I, _I, _P, _P, _P]),
    ("smg_mg_precondition_batched", _I, [_P, _I, _I, _P, _P]),
    ("smg_mg_coarsest_batched", _I, [_P, _I, _P, _P]), ("smg_mg_cheb_interval", _I, [_P, _P, _P]),
    ("smg_mg_set_cheb_interval", _I, [_P, _D, _D]),
    ("smg_mg_set_cheb_smoother", _I, [_P, _D, _I, _P]),
    ("smg_gcr_solve_batched", _I, [_P, _P, _I, _I, _P, _P, _I, _D, _I, _D, _P, _P]),
    ("smg_gcr_solve_batched_dev", _I, [_P, _P, _I, _I, _P, _P, _I, _D, _I, _D, _P, _P, _P]),
    ("smg_force_forward", _I, [_P, _P, _P]), ("smg_force_backward", _I, [_P, _P, _P, _P]),
    ("smg_fermion_force", _I, [_P, _P, _P]), ("smg_fermion_force_dev", _I, [_P, _P, _P, _P]),
    ("smg_hopping_force", _I, [_P, _P, _P, _P]), ("smg_logdet_force", _I, [_P, _P]),
    ("smg_eo_fermion_force", _I, [_P, _P, _P, _P]), ("smg_fermion_action", _I, [_P, _P, _I, _D, _P, _P, _P]),
    ("smg_dirac_get_gauge", _I, [_P, _P]), ("smg_hmc_evolve", _I, [_P, _P, _P, _P, _P]),
    ("smg_hmc_hamiltonian", _I, [_P, _I, _D, _P, _P, _I, _D, _P, _P]),
    ("smg_hmc_trajectory", _I, [_P, _P, _P, _P, _D, _P]),
    ("smg_gauge_write_raw", _I, [ctypes.c_char_p, _I, _P]), ("smg_gauge_read_raw", _I, [ctypes.c_char_p, _I, _P]),
    ("smg_gauge_write", _I, [ctypes.c_char_p, _I, _P, _D, _D]),
    ("smg_gauge_read", _I, [ctypes.c_char_p, _I, _P, _P]),
    ("smg_therm_cfg_path", _I, [_I, _D, _I, ctypes.c_char_p, ctypes.c_size_t]),
    ("smg_random_vec", _I, [_I, ctypes.c_uint, _P]), ("smg_dense_eigh", _I, [_I, _P, _P, _P]),
    ("smg_dirac_trlm", _I, [_P, _I, _I, _I, _I, _D, _P, _P, _P, _P]),
    ("smg_coarse_trlm", _I, [_P, _I, _I, _I, _D, _P, _P, _P, _P]),
    ("smg_coarse_setup_deflation", _I, [_P, _I, _I, _I, _D, _P, _P]),
    ("smg_coarse_deflation", _I, [_P, _P, _P, _P]), ("smg_coarse_set_deflation", _I, [_P, _I, _P, _P]),
    ("smg_cg_solve_deflated", _I, [_P, _I, _P, _P, _I, _P, _P, _I, _D, _P]),
    ("smg_dirac_lobpcg_ex", _I, [_P, _P, _I, _P, _I, _D, _I, _P, _P, _P]),
    ("smg_dirac_lobpcg", _I, [_P, _P, _I, _P, _I, _D, _P, _P, _P]),
    ("smg_coarse_lowmode_force", _I, [_P, _P, _P, _P, _I, _P]),
    ("smg_coarse_evolve_deflation", _I, [_P, _P]),
    ("smg_hmc_ms_evolve", _I, [_P, _P, _P, _P, _P, _P, _P, _P, _P]),
    ("smg_hmc_ms_trajectory", _I, [_P, _P, _P, _P, _P, _P, _P, _D, _P]),
    ("smg_hmc_nested_evolve", _I, [_P, _P, _P, _P, _P, _P]), ("smg_mg_refresh", _I, [_P, _I]),
    ("smg_hmc_nested_trajectory", _I, [_P, _P, _P, _P, _P, _D, _P]),
    ("smg_dirac_heatbath", _I, [_P, _D, _I, ctypes.c_ulonglong, ctypes.c_uint]),
    ("smg_lattices_create", _I, [_P, _I, _P, _P, _P, _D, _D, _D]), ("smg_lattices_destroy", _I, [_P]),
    ("smg_lattices_size", _I, [_P, _P]), ("smg_lattices_apply", _I, [_P, _I, _P, _P]),
    ("smg_lattices_apply_dev", _I, [_P, _I, _P, _P, _P]),
    ("smg_lattices_cg", _I, [_P, _P, _P, _I, _D, _P, _P]),
    ("smg_lattices_mg_setup", _I, [_P, _P, _P, _P]), ("smg_lattices_mg_destroy", _I, [_P]),
    ("smg_lattices_mg_precondition", _I, [_P, _P, _P]),
    ("smg_lattices_mg_precondition_dev", _I, [_P, _P, _P, _P]),
    ("smg_lattices_gcr", _I, [_P, _P, _P, _I, _D, _I, _P, _P, _P]),
    ("smg_lattices_gcr_dev", _I, [_P, _P, _P, _I, _D, _I, _P, _P, _P, _P]),
    ("smg_eo_mg_setup", _I, [_P, _P, _I, _I, _I, _I, _I, _I, _P, _P]), ("smg_eo_mg_destroy", _I, [_P]),
    ("smg_eo_mg_info", _I, [_P, _P, _P, _P]), ("smg_eo_mg_coarse_matrix", _I, [_P, _P]),
    ("smg_eo_mg_precondition", _I, [_P, _P, _P, _P]), ("smg_eo_mg_precondition_dev", _I, [_P, _P, _P, _P]),
    ("smg_heatbath", _I, [_I, _P, _D, _I, ctypes.c_ulonglong, ctypes.c_uint]),
    ("smg_coarse_lobpcg", _I, [_P, _I, _P, _I, _D, _P, _P, _P]),
    ("smg_mg_update_coarsest_deflation", _I, [_P, _I, _I]),
    ("smg_mg_coarsest_deflation", _I, [_P, _P, _P, _P]),
    ("smg_mg_setup_sparse_coarse", _I, [_P, _I, _D, _I, _P]), ("smg_mg_rebuild", _I, [_P]),
]

_lib = None


def load_library():
    """Load libsmg.so and bind every C-ABI entry point (no GPU call)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SMGError(f"{LIB_PATH} is not built; run `python -m schwinger_mg_amd.build`")
        # One HIP runtime per process: PyTorch ships its own libamdhip64 with the
        # same soname (libamdhip64.so.7) as /opt/rocm's.  Loading torch first makes
        # libsmg.so bind to the runtime torch already mapped, so device pointers
        # and streams are shared instead of two runtimes competing for the GPU.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        lib = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


_initialised = False


def lib():
    global _initialised
    L = load_library()
    if not _initialised:
        rc = L.smg_init(0)
        if rc != 0:
            raise SMGError("smg_init failed: " + L.smg_last_error().decode())
        _initialised = True
    return L


def _chk(rc):
    if rc != 0:
        raise SMGError(lib().smg_last_error().decode())


def _c(a, dt=np.complex128):
    return np.ascontiguousarray(a, dtype=dt)


def _p(a):
    return a.ctypes.data_as(_P) if a is not None else None


def synchronize():
    _chk(lib().smg_synchronize())


def library_stream():
    """The hipStream_t (as int) the library launches on."""
    s = _P()
    _chk(lib().smg_get_stream(ctypes.byref(s)))
    return s.value or 0


# ------------------------------------------------------------------ gauge
def avg_plaq(L, U):
    out = _D()
    _chk(lib().smg_avg_plaq(L, _p(_c(U)), ctypes.byref(out)))
    return out.value


def gauge_action(L, U, beta):
    out = _D()
    _chk(lib().smg_gauge_action(L, _p(_c(U)), beta, ctypes.byref(out)))
    return out.value


def gauge_force(L, U, beta):
    F = np.zeros(2 * L * L)
    _chk(lib().smg_gauge_force(L, _p(_c(U)), beta, _p(F)))
    return F


def eo_maps(L):
    V = L * L
    e = np.zeros(V // 2, np.int32); o = np.zeros(V // 2, np.int32); i = np.zeros(V, np.int32)
    _chk(lib().smg_eo_maps(L, _p(e), _p(o), _p(i)))
    return e, o, i


# ------------------------------------------------------------------ Dirac
class DiracOp:
    """DiracOp(lat, gauge, m, r, csw, mu_t) -- src/dirac.h:8-183."""

    def __init__(self, L, U, mass, r=1.0, csw=0.0, mu_t=0.0):
        self.L, self.V = L, L * L
        self.ndof = 2 * self.V
        self.mass, self.r, self.c_sw, self.mu_t = mass, r, csw, mu_t
        h = _P()
        _chk(lib().smg_dirac_create(ctypes.byref(h), L, _p(_c(U)), mass, r, csw, mu_t))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.smg_dirac_destroy(self.h)
            self.h = None

    def set_gauge(self, U):
        _chk(lib().smg_dirac_set_gauge(self.h, _p(_c(U))))

    @property
    def clover_field(self):
        F = np.zeros(self.V)
        _chk(lib().smg_dirac_clover(self.h, _p(F)))
        return F

    def _apply(self, kind, src):
        dst = np.zeros(self.ndof, np.complex128)
        _chk(lib().smg_dirac_apply(self.h, kind, _p(_c(src)), _p(dst)))
        return dst

    def apply(self, src):
        return self._apply(D_, src)

    def apply_dag(self, src):
        return self._apply(DDAG, src)

    def apply_DdagD(self, src):
        return self._apply(DDAGD, src)

    def apply_g5D(self, src):
        return self._apply(G5D, src)

    def apply_delta_D(self, src, pi, dt):
        dst = np.zeros(self.ndof, np.complex128)
        _chk(lib().smg_dirac_delta(self.h, _p(_c(src)), _p(dst), _p(_c(pi, np.float64)), dt))
        return dst

    def apply_dev(self, kind, prec, src_ptr, dst_ptr, nrhs=1, stream=None):
        _chk(lib().smg_dirac_apply_dev(self.h, kind, prec, src_ptr, dst_ptr, nrhs, stream))

    def as_DdagD_op(self):
        return self.apply_DdagD


class EvenOddDiracOp:
    """EvenOddDiracOp(D) -- src/dirac.h:190-553 (half-lattice vectors of V complex)."""

    def __init__(self, D):
        self.D = D
        self.nh = D.V

    def _eo(self, kind, src):
        dst = np.zeros(self.nh, np.complex128)
        _chk(lib().smg_eo_apply(self.D.h, kind, _p(_c(src)), _p(dst)))
        return dst

    def apply_eo(self, src_e):
        return self._eo(EO_HOP_EO, src_e)

    def apply_oe(self, src_o):
        return self._eo(EO_HOP_OE, src_o)

    def apply_ee_inv(self, src):
        return self._eo(EO_EE_INV, src)

    def apply_oo_inv(self, src):
        return self._eo(EO_OO_INV, src)

    def apply_oo(self, src):
        return self._eo(EO_OO, src)

    def apply_schur(self, src):
        return self._eo(EO_SCHUR, src)

    def apply_schur_dag(self, src):
        return self._eo(EO_SCHUR_DAG, src)

    def apply_schur_dag_schur(self, src):
        return self._eo(EO_SCHUR_DAG_SCHUR, src)

    def reconstruct(self, x_o, phi):
        out = np.zeros(self.D.ndof, np.complex128)
        _chk(lib().smg_eo_reconstruct(self.D.h, _p(_c(x_o)), _p(_c(phi)), _p(out)))
        return out

    def reconstruct_full(self, x_o):
        out = np.zeros(self.D.ndof, np.complex128)
        _chk(lib().smg_eo_reconstruct(self.D.h, _p(_c(x_o)), None, _p(out)))
        return out

    def log_det_ee(self):
        out = _D()
        _chk(lib().smg_eo_logdet(self.D.h, ctypes.byref(out)))
        return out.value


# ------------------------------------------------------------------ prolongator / coarse op
class Prolongator:
    """Prolongator(lat, bx, by, k) -- src/prolongator.h:14-142, with nd dof per site
    (nd = 2 on the fine lattice; nd = k of the level above on a coarse lattice,
    which also covers CoarseProlongator, src/prolongator.h:149-228).  k is the
    coarse dof per aggregate; chiral=True splits each of k/2 null vectors into its
    two chiralities (no reference counterpart)."""

    def __init__(self, nx, ny, nd, bx, by, k, chiral=False):
        self.nx, self.ny, self.nd, self.bx, self.by, self.k_vec = nx, ny, nd, bx, by, k
        self.chiral = bool(chiral)
        self.nvec = k // 2 if chiral else k
        self.nbx, self.nby = nx // bx, ny // by
        self.nblocks = self.nbx * self.nby
        self.coarse_dim = self.nblocks * k
        self.fine_dim = nx * ny * nd
        h = _P()
        _chk(lib().smg_prolongator_create(ctypes.byref(h), nx, ny, nd, bx, by, k, int(chiral)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.smg_prolongator_destroy(self.h)
            self.h = None

    def build_from_vectors(self, null_vecs):
        nv = _c(np.asarray(null_vecs)[: self.nvec]).ravel()
        _chk(lib().smg_prolongator_build(self.h, _p(nv)))
        return self

    @property
    def block_vecs(self):
        out = np.zeros(self.nblocks * self.k_vec * self.bx * self.by * self.nd, np.complex128)
        _chk(lib().smg_prolongator_get(self.h, _p(out)))
        return out

    def set_block_vecs(self, bv):
        _chk(lib().smg_prolongator_set(self.h, _p(_c(bv))))

    def prolong(self, coarse):
        out = np.zeros(self.fine_dim, np.complex128)
        _chk(lib().smg_prolong(self.h, _p(_c(coarse)), _p(out)))
        return out

    def restrict_vec(self, fine):
        out = np.zeros(self.coarse_dim, np.complex128)
        _chk(lib().smg_restrict(self.h, _p(_c(fine)), _p(out)))
        return out


class SparseCoarseOp:
    """9-point coarse operator -- src/coarse_op.h:154-548 (built by Galerkin projection)."""

    def __init__(self, handle, nx, ny, k, keep=()):
        self.h, self.nbx, self.nby, self.k_vec = handle, nx, ny, k
        self.nblocks = nx * ny
        self.dim = self.nblocks * k
        self._keep = keep

    @classmethod
    def build(cls, P, D):
        """A_c = P^dag (D^dag D) P (src/coarse_op.h:241-279 with A = D.as_DdagD_op())."""
        h = _P()
        _chk(lib().smg_coarse_create_galerkin(ctypes.byref(h), D.h, P.h))
        return cls(h, P.nbx, P.nby, P.k_vec)

    @classmethod
    def build_from_coarse(cls, P, A):
        h = _P()
        _chk(lib().smg_coarse_create_from_coarse(ctypes.byref(h), A.h, P.h))
        return cls(h, P.nbx, P.nby, P.k_vec)

    @classmethod
    def from_links(cls, nx, ny, k, links):
        h = _P()
        _chk(lib().smg_coarse_create_links(ctypes.byref(h), nx, ny, k, _p(_c(links))))
        return cls(h, nx, ny, k)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.smg_coarse_destroy(self.h)
            self.h = None

    @property
    def link(self):
        out = np.zeros(self.nblocks * 9 * self.k_vec ** 2, np.complex128)
        _chk(lib().smg_coarse_get_links(self.h, _p(out)))
        return out

    def apply_to(self, src, prec=F64):
        out = np.zeros(self.dim, np.complex128)
        _chk(lib().smg_coarse_apply(self.h, prec, _p(_c(src)), _p(out)))
        return out

    def apply_batched(self, S, prec=F64):
        """apply_to on N right-hand sides at once; S is (N, dim)."""
        S = _c(S)
        out = np.zeros_like(S)
        _chk(lib().smg_coarse_apply_batched(self.h, prec, S.shape[0], _p(S), _p(out)))
        return out

    def dense(self):
        """CoarseOp's dense matrix, returned as columns: out[j] = A_c e_j (src/coarse_op.h:18-37)."""
        out = np.zeros(self.dim * self.dim, np.complex128)
        _chk(lib().smg_coarse_dense(self.h, _p(out)))
        return out.reshape(self.dim, self.dim)

    def dense_solve(self, b):
        """CoarseOp::solve (src/coarse_op.h:101-137) through the dense inverse."""
        x = np.zeros(self.dim, np.complex128)
        _chk(lib().smg_coarse_dense_solve(self.h, _p(_c(b)), _p(x)))
        return x

    def solve(self, b, max_iter=200, cg_tol=1e-12):
        x = np.zeros(self.dim, np.complex128)
        it = _I()
        _chk(lib().smg_coarse_solve(self.h, _p(_c(b)), _p(x), max_iter, cg_tol, ctypes.byref(it)))
        self.last_iterations = it.value
        return x

    def as_op(self):
        return self.apply_to

    def setup_deflation(self, n_ev, n_kr=0, max_restarts=100, tol=1e-10, solver="trlm", start_vec=None):
        """SparseCoarseOp::setup_deflation (src/coarse_op.cpp:10-54), TRLM branch: the
        n_ev smallest eigenpairs of A_c become the deflation space of solve().
        start_vec defaults to the reference's random_vec(dim, mt19937(42))."""
        if solver != "trlm":
            raise SMGError(f"setup_deflation: solver {solver!r} is not provided (only 'trlm')")
        start = random_vec(self.dim, 42) if start_vec is None else _c(start_vec)
        info = TRLMInfo()
        _chk(lib().smg_coarse_setup_deflation(self.h, n_ev, n_kr, max_restarts, tol, _p(start),
                                              ctypes.byref(info)))
        self.last_trlm = info
        return info

    def _deflation(self):
        nd = _I()
        _chk(lib().smg_coarse_deflation(self.h, ctypes.byref(nd), None, None))
        vals = np.zeros(nd.value)
        vecs = np.zeros((nd.value, self.dim), np.complex128)
        if nd.value:
            _chk(lib().smg_coarse_deflation(self.h, ctypes.byref(nd), _p(vals), _p(vecs)))
        return vals, vecs

    @property
    def defl_vals(self):
        """SparseCoarseOp::defl_vals (src/coarse_op.h:170)."""
        return self._deflation()[0]

    @property
    def defl_vecs(self):
        """SparseCoarseOp::defl_vecs (src/coarse_op.h:169), one row per vector."""
        return self._deflation()[1]

    def set_deflation(self, vals, vecs):
        """Replace the deflation space (an empty one switches deflation off)."""
        vals = np.ascontiguousarray(vals, np.float64)
        vecs = _c(vecs).reshape(len(vals), -1) if len(vals) else np.zeros((0, self.dim), np.complex128)
        _chk(lib().smg_coarse_set_deflation(self.h, len(vals), _p(vals) if len(vals) else None,
                                            _p(vecs) if len(vals) else None))


# ------------------------------------------------------------------ smoothers
def mr_smooth(D, x, b, niter):
    """mr_smooth(D, x, b, niter) -- src/smoother.cpp:9-23; returns the updated x."""
    x = _c(x).copy()
    _chk(lib().smg_smooth_dirac(D.h, SMOOTH_MR, _p(x), _p(_c(b)), niter, 0.0))
    return x


def richardson_smooth(D, x, b, niter, omega):
    x = _c(x).copy()
    _chk(lib().smg_smooth_dirac(D.h, SMOOTH_RICHARDSON, _p(x), _p(_c(b)), niter, omega))
    return x


def mr_smooth_op(C, x, b, niter):
    """mr_smooth_op on a SparseCoarseOp -- src/smoother.cpp:26-40."""
    x = _c(x).copy()
    _chk(lib().smg_smooth_coarse(C.h, SMOOTH_MR, _p(x), _p(_c(b)), niter, 0.0))
    return x


def richardson_smooth_op(C, x, b, niter, omega):
    x = _c(x).copy()
    _chk(lib().smg_smooth_coarse(C.h, SMOOTH_RICHARDSON, _p(x), _p(_c(b)), niter, omega))
    return x


def estimate_lambda_max(D, v0, niter):
    """estimate_lambda_max of D^dag D (src/smoother.cpp:43-59) from start vector v0."""
    out = _D()
    _chk(lib().smg_lambda_max_dirac(D.h, _p(_c(v0)), niter, ctypes.byref(out)))
    return out.value


# ------------------------------------------------------------------ multigrid
def _mg_params(n_levels, blocks, k, null_iters, pre_smooth, post_smooth, w_cycle, coarse_maxit, coarse_tol,
               coarse_half, coarse_solver, use_graph, chiral, null_prec):
    p = MGParams()
    p.n_levels = n_levels
    blocks = list(blocks) + [blocks[-1]] * (4 - len(blocks))
    k = list(k) + [k[-1]] * (4 - len(k))
    for i in range(4):
        p.bx[i], p.by[i] = blocks[i]
        p.k[i] = k[i]
    p.null_iters, p.pre_smooth, p.post_smooth = null_iters, pre_smooth, post_smooth
    p.w_cycle, p.coarse_maxit, p.coarse_tol, p.coarse_half = int(w_cycle), coarse_maxit, coarse_tol, int(coarse_half)
    p.coarse_solver, p.use_graph, p.chiral = coarse_solver, int(use_graph), int(chiral)
    p.null_prec = 0 if null_prec == F64 else 1
    return p


def mg_start_sizes(L, n_levels, blocks, k, chiral):
    """The near-null start vectors an L x L lattice's hierarchy takes: one
    (count, length) pair per non-coarsest level (k[l] vectors, k[l] / 2 with chiral
    aggregation, of that level's dimension)."""
    blocks = list(blocks) + [blocks[-1]] * (4 - len(blocks))
    k = list(k) + [k[-1]] * (4 - len(k))
    out, nx, ny, nd = [], L, L, 2
    for l in range(n_levels - 1):
        out.append((k[l] // 2 if chiral else k[l], nx * ny * nd))
        nx, ny, nd = nx // blocks[l][0], ny // blocks[l][1], k[l]
    return out


class MGHierarchy:
    """MGHierarchy (src/multigrid.h:41-107) built by build_mg_hierarchy."""

    def __init__(self, D, starts, n_levels=2, blocks=((4, 4),), k=(4,), null_iters=20, pre_smooth=3,
                 post_smooth=3, w_cycle=True, coarse_maxit=200, coarse_tol=1e-12, coarse_half=False,
                 coarse_solver=0, use_graph=False, chiral=False, null_prec=F64):
        p = _mg_params(n_levels, blocks, k, null_iters, pre_smooth, post_smooth, w_cycle, coarse_maxit, coarse_tol,
                       coarse_half, coarse_solver, use_graph, chiral, null_prec)
        self.params, self.D, self.n_levels = p, D, n_levels
        h = _P()
        _chk(lib().smg_mg_setup(ctypes.byref(h), D.h, ctypes.byref(p), _p(_c(starts))))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.smg_mg_destroy(self.h)
            self.h = None

    def level_dim(self, level):
        n = _SZ()
        _chk(lib().smg_mg_level_dim(self.h, level, ctypes.byref(n)))
        return n.value

    def update_coarsest_deflation(self, n_defl, lobpcg_iters=3):
        """MGHierarchy::update_coarsest_deflation (src/multigrid.cpp:111-141)."""
        _chk(lib().smg_mg_update_coarsest_deflation(self.h, n_defl, lobpcg_iters))

    def setup_sparse_coarse(self, n_defl, cg_tol=1e-12, max_cg_iter=200, start_vec=None):
        """MGHierarchy::setup_sparse_coarse (src/multigrid.cpp:754-794), TRLM: the FP64
        cycle's coarsest solve becomes deflated CG (two-level hierarchies)."""
        dim = self.level_dim(1)
        start = random_vec(dim, 42) if start_vec is None else _c(start_vec)
        _chk(lib().smg_mg_setup_sparse_coarse(self.h, n_defl, cg_tol, max_cg_iter, _p(start)))

    def refresh(self, null_iters=5):
        """build_full_mg with warm_start (src/mg_builder.cpp:3-31): null_iters more
        near-null iterations from the current vectors on the current gauge field at
        every level, then the aggregations and Galerkin operators rebuilt."""
        _chk(lib().smg_mg_refresh(self.h, null_iters))

    def rebuild(self):
        """Re-project every level's Galerkin operator for the operator's current gauge
        field (Ac.build + rebuild_deeper_levels, src/multigrid.cpp:163-194)."""
        _chk(lib().smg_mg_rebuild(self.h))

    def coarsest_deflation(self):
        """(defl_vals, defl_vecs) of the coarsest level (src/multigrid.h:28-29)."""
        nd = _I()
        _chk(lib().smg_mg_coarsest_deflation(self.h, ctypes.byref(nd), None, None))
        vals = np.zeros(nd.value)
        vecs = np.zeros((nd.value, self.level_dim(self.n_levels - 1)), np.complex128)
        if nd.value:
            _chk(lib().smg_mg_coarsest_deflation(self.h, ctypes.byref(nd), _p(vals), _p(vecs)))
        return vals, vecs

    def set_symmetric(self, damping, starts):
        """set_symmetric (src/multigrid.cpp:146-152) with explicit power-iteration starts."""
        _chk(lib().smg_mg_set_symmetric(self.h, damping, _p(_c(starts))))

    def precondition(self, b, prec=F64):
        x = np.zeros(self.D.ndof, np.complex128)
        _chk(lib().smg_mg_precondition(self.h, prec, _p(_c(b)), _p(x)))
        return x

    def null_vecs(self, level=0):
        n = self.level_dim(level)
        nv = self.params.k[level] // 2 if self.params.chiral else self.params.k[level]
        out = np.zeros(nv * n, np.complex128)
        _chk(lib().smg_mg_null_vecs(self.h, level, _p(out)))
        return out.reshape(nv, n)

    def block_vecs(self, level=0):
        nc = self.level_dim(level + 1)
        nf = self.level_dim(level)
        out = np.zeros(nc // self.params.k[level] * self.params.k[level] * (nf * self.params.k[level] // nc),
                       np.complex128)
        _chk(lib().smg_mg_block_vecs(self.h, level, _p(out)))
        return out

    def coarse_links(self, level=0):
        k = self.params.k[level]
        nb = self.level_dim(level + 1) // k
        out = np.zeros(nb * 9 * k * k, np.complex128)
        _chk(lib().smg_mg_coarse_links(self.h, level, _p(out)))
        return out

    def set_cycle(self, pre_smooth, post_smooth, coarse_maxit, coarse_tol, use_graph=False):
        _chk(lib().smg_mg_set_cycle(self.h, pre_smooth, post_smooth, coarse_maxit, coarse_tol, int(use_graph)))

    def precondition_batched(self, B, prec=F64):
        """One cycle on N right-hand sides at once; B is (N, ndof)."""
        B = _c(B)
        N = B.shape[0]
        X = np.zeros_like(B)
        _chk(lib().smg_mg_precondition_batched(self.h, prec, N, _p(B), _p(X)))
        return X

    def coarsest_batched(self, B):
        """The batched FP32 cycle's coarsest-level solve on N right-hand sides at once;
        B is (N, coarsest dimension)."""
        B = _c(B)
        X = np.zeros_like(B)
        _chk(lib().smg_mg_coarsest_batched(self.h, B.shape[0], _p(B), _p(X)))
        return X

    def set_cheb_smoother(self, ratio, fused=True):
        """Chebyshev smoothing on [hi / ratio, hi] at every non-coarsest level of the batched
        cycles, FP32 and FP64 (ratio > 0); 0 restores MR.  fused: each step one launch of the level's
        operator where it has that form, else the apply and separate vector passes.
        Returns each non-coarsest level's hi."""
        hi = np.zeros(self.n_levels - 1)
        _chk(lib().smg_mg_set_cheb_smoother(self.h, float(ratio), int(bool(fused)), _p(hi)))
        return hi

    def cheb_interval(self):
        """(lo, hi): the spectral interval of the Chebyshev coarsest solve (coarse_solver=2)."""
        lo, hi = _D(), _D()
        _chk(lib().smg_mg_cheb_interval(self.h, ctypes.byref(lo), ctypes.byref(hi)))
        return lo.value, hi.value

    def set_cheb_interval(self, lo, hi):
        """Replace the coarsest Chebyshev interval until the next rebuild or refresh."""
        _chk(lib().smg_mg_set_cheb_interval(self.h, float(lo), float(hi)))

    def omega(self, level):
        out = _D()
        _chk(lib().smg_mg_omega(self.h, level, ctypes.byref(out)))
        return out.value


def build_mg_hierarchy(D, n_levels, block_size, k_null, null_iters, starts, w_cycle=True, pre_smooth=3,
                       post_smooth=3, coarse_blocks=None, **kw):
    """build_mg_hierarchy (src/multigrid.cpp:595-687) with explicit null-space starts."""
    blocks = [(block_size, block_size)] + list(coarse_blocks or [(2, 2)] * 3)
    return MGHierarchy(D, starts, n_levels=n_levels, blocks=blocks[:4], k=(k_null,) * 4, null_iters=null_iters,
                       pre_smooth=pre_smooth, post_smooth=post_smooth, w_cycle=w_cycle, **kw)


# ------------------------------------------------------------------ solvers
@dataclass
class SolveResult:
    solution: np.ndarray
    iterations: int
    final_residual: float
    true_residual: float
    restarts: int = 0


def _res(x, info):
    return SolveResult(x, info.iterations, info.final_residual, info.true_residual, info.restarts)


def cg_solve(D, rhs, max_iter, tol, x0=None, mg=None, op=OP_DDAGD):
    """cg_solve / cg_solve_x0 / cg_solve_precond (src/solvers.cpp:16-96)."""
    n = D.ndof if op == OP_DDAGD else D.V
    x = np.zeros(n, np.complex128)
    info = SolveInfo()
    _chk(lib().smg_cg_solve(D.h, op, mg.h if mg else None, _p(_c(rhs)), _p(_c(x0)) if x0 is not None else None,
                            _p(x), max_iter, tol, ctypes.byref(info)))
    return _res(x, info)


def fcg_solve(D, rhs, mg, max_iter, tol, trunc=30, prec=F64, ref_projection=False, inner_tol=1e-5):
    """fcg_solve / fcg_solve_mg (truncated GCR, src/solvers.cpp:223-300, 803-811)."""
    x = np.zeros(D.ndof, np.complex128)
    info = SolveInfo()
    _chk(lib().smg_gcr_solve(D.h, mg.h if mg else None, prec, _p(_c(rhs)), _p(x), max_iter, tol, trunc,
                             int(ref_projection), inner_tol, ctypes.byref(info)))
    return _res(x, info)


gcr_solve = fcg_solve


def gcr_solve_batched(D, B, mg, max_iter, tol, trunc=10, prec=F64, inner_tol=1e-5):
    """GCR + MG on N right-hand sides in lockstep; B is (N, ndof).  Returns
    (X, info, per-RHS FP64 true residuals)."""
    B = _c(B)
    N = B.shape[0]
    X = np.zeros_like(B)
    info = SolveInfo()
    tr = np.zeros(N)
    _chk(lib().smg_gcr_solve_batched(D.h, mg.h if mg else None, prec, N, _p(B), _p(X), max_iter, tol, trunc,
                                     inner_tol, _p(tr), ctypes.byref(info)))
    return X, info, tr


@dataclass
class RitzPair:
    """RitzPair (src/solvers.h:13-16)."""
    value: float
    vector: np.ndarray


@dataclass
class FGMRESResult(SolveResult):
    """FGMRESResult (src/solvers.h:18-23): the solve and the harvested Ritz pairs."""
    ritz_pairs: list = field(default_factory=list)


def fgmres_solve(D, rhs, mg, restart, max_iter, tol, n_ritz_harvest=4):
    """fgmres_solve_generic / fgmres_solve_mg (src/solvers.cpp:555-800), with the
    reference's default n_ritz_harvest = 4: after every restart cycle of at least
    n_ritz_harvest steps the lowest Ritz pairs of the cycle's Hessenberg are kept
    (at most 2 n_ritz_harvest) and deflate the later preconditioned vectors.
    mg = None: the identity preconditioner."""
    x = np.zeros(D.ndof, np.complex128)
    info = SolveInfo()
    k = max(int(n_ritz_harvest), 0)
    vals = np.zeros(max(2 * k, 1))
    vecs = np.zeros((max(2 * k, 1), D.ndof), np.complex128)
    npairs = ctypes.c_int(0)
    _chk(lib().smg_fgmres_solve_ritz(D.h, mg.h if mg else None, _p(_c(rhs)), _p(x), restart, max_iter, tol, k,
                                     ctypes.byref(info), ctypes.byref(npairs), _p(vals), _p(vecs)))
    pairs = [RitzPair(float(vals[q]), vecs[q].copy()) for q in range(npairs.value)]
    return FGMRESResult(x, info.iterations, info.final_residual, info.true_residual, info.restarts, pairs)


def cg_solve_deflated(D, rhs, mg, defl_vecs, defl_vals, max_iter, tol, op=OP_DDAGD):
    """cg_solve_deflated (src/solvers.cpp:111-213): CG started in the deflation space
    U (rows of defl_vecs) with the projected preconditioned residual; mg = the FP64
    cycle as preconditioner or None.  defl_vals is accepted as in the reference and,
    as there, not read."""
    n = D.ndof if op == OP_DDAGD else D.V
    U = _c(defl_vecs).reshape(-1, n) if len(defl_vecs) else np.zeros((0, n), np.complex128)
    x = np.zeros(n, np.complex128)
    info = SolveInfo()
    _chk(lib().smg_cg_solve_deflated(D.h, op, mg.h if mg else None, _p(_c(rhs)), U.shape[0],
                                     _p(U) if U.shape[0] else None, _p(x), max_iter, tol, ctypes.byref(info)))
    return _res(x, info)


# ------------------------------------------------------------------ eigensolvers
class TRLMInfo(ctypes.Structure):
    _fields_ = [("iterations", _I), ("restarts", _I), ("converged", _I)]


@dataclass
class TRLMResult:
    """TRLMResult (src/eigensolver.h:19-25)."""
    eigvecs: np.ndarray   # (n_ev, n): eigvecs[i] is eigenvector i
    eigvals: np.ndarray
    iterations: int
    num_restarts: int
    converged: bool


def random_vec(n, seed):
    """random_vec(n, std::mt19937(seed)) (src/linalg.cpp:44-49), drawn as the
    reference's g++ build draws it."""
    out = np.zeros(n, np.complex128)
    _chk(load_library().smg_random_vec(n, seed, _p(out)))
    return out


def lanczos_eigen(A_cols):
    """lanczos_eigen (src/eigensolver_core.cpp:103-265) on the Hermitian matrix given
    as its columns, A_cols[j][i] = A_ij: (ascending eigenvalues, evecs) with
    evecs[row][col] component row of eigenvector col, the reference's layout."""
    cols = _c(A_cols)
    n = cols.shape[0]
    ev = np.zeros(n)
    out = np.zeros((n, n), np.complex128)
    if n:
        _chk(load_library().smg_dense_eigh(n, _p(cols), _p(ev), _p(out)))
    return ev, out


def trlm_eigensolver(A, n_ev, n_kr=0, max_restarts=100, tol=1e-10, start_vec=None, op=OP_DDAGD):
    """trlm_eigensolver (src/eigensolver_core.cpp:691-1057) without the Chebyshev
    filter.  A = a DiracOp (the operator D^dag D for op OP_DDAGD, the reference's
    D.as_DdagD_op(), or M^dag M for OP_SCHUR) or a SparseCoarseOp (its as_op()); start_vec
    defaults to random_vec(n, 42) as in the reference."""
    coarse = isinstance(A, SparseCoarseOp)
    n = A.dim if coarse else (A.ndof if op == OP_DDAGD else A.V)
    start = random_vec(n, 42) if start_vec is None else _c(start_vec)
    ev = np.zeros(n_ev)
    vecs = np.zeros((n_ev, n), np.complex128)
    info = TRLMInfo()
    if coarse:
        _chk(lib().smg_coarse_trlm(A.h, n_ev, n_kr, max_restarts, tol, _p(start), _p(ev), _p(vecs),
                                   ctypes.byref(info)))
    else:
        _chk(lib().smg_dirac_trlm(A.h, op, n_ev, n_kr, max_restarts, tol, _p(start), _p(ev), _p(vecs),
                                  ctypes.byref(info)))
    return TRLMResult(vecs, ev, info.iterations, info.restarts, bool(info.converged))


@dataclass
class LOBPCGResult:
    """LOBPCGResult (src/eigensolver.h:6-10)."""
    eigvecs: np.ndarray   # (k, n)
    eigvals: np.ndarray
    iterations: int


def lobpcg_update(A, X0, precond=None, max_iter=5, tol=1e-8, exact_ax=False):
    """lobpcg_update (src/eigensolver_core.cpp:279-533) from the k start vectors X0
    (rows).  A = a DiracOp (the operator D^dag D) or a SparseCoarseOp; precond = an
    MGHierarchy (its FP64 cycle; fine operator only) or None for the identity.
    exact_ax (DiracOp only) keeps AX = A X through the re-orthonormalisation, where
    the reference's update does not."""
    X0 = _c(X0)
    k, n = X0.shape
    ev = np.zeros(k)
    X = np.zeros_like(X0)
    it = _I()
    if isinstance(A, SparseCoarseOp):
        if exact_ax:
            raise SMGError("lobpcg_update: exact_ax is offered on the Dirac operator")
        if precond is not None:
            raise SMGError("lobpcg_update: the coarse operator takes the identity preconditioner")
        _chk(lib().smg_coarse_lobpcg(A.h, k, _p(X0), max_iter, tol, _p(ev), _p(X), ctypes.byref(it)))
    elif exact_ax:
        _chk(lib().smg_dirac_lobpcg_ex(A.h, precond.h if precond else None, k, _p(X0), max_iter, tol, 1, _p(ev),
                                       _p(X), ctypes.byref(it)))
    else:
        _chk(lib().smg_dirac_lobpcg(A.h, precond.h if precond else None, k, _p(X0), max_iter, tol, _p(ev), _p(X),
                                    ctypes.byref(it)))
    return LOBPCGResult(X, ev, it.value)


# ------------------------------------------------------------------ multi-timescale MG HMC
class MSParams(ctypes.Structure):
    """MGMultiScaleParams (src/hmc.h:346-365), the fields the device trajectory uses."""
    _fields_ = [("beta", _D), ("tau", _D), ("n_outer", _I), ("n_inner", _I), ("omelyan", _I),
                ("inner_smooth", _I), ("defl_refresh", _I), ("cg_maxiter", _I), ("cg_tol", _D),
                ("force_accept", _I)]


class MSResult(ctypes.Structure):
    """MGMultiScaleResult (src/hmc.h:367-376) without the timings."""
    _fields_ = [("accepted", _I), ("dH", _D), ("dKE", _D), ("dSG", _D), ("dSF", _D),
                ("highmode_cg_iters", _I), ("lowmode_force_evals", _I)]


def ms_params(beta, tau=1.0, n_outer=4, n_inner=5, omelyan=False, inner_smooth=3, defl_refresh=0, cg_maxiter=500,
              cg_tol=1e-10, force_accept=False):
    return MSParams(beta, tau, n_outer, n_inner, int(omelyan), inner_smooth, defl_refresh, cg_maxiter, cg_tol,
                    int(force_accept))


def coarse_lowmode_force(D, P, defl, phi, smooth_iters=0):
    """coarse_lowmode_force (src/hmc_multiscale.cpp:237-276) with the deflation space of
    the SparseCoarseOp defl on P's coarse space: F[mu*V + s]."""
    F = np.zeros(2 * D.V)
    _chk(lib().smg_coarse_lowmode_force(D.h, P.h, defl.h, _p(_c(phi)), smooth_iters, _p(F)))
    return F


def evolve_coarse_deflation(defl, A_new):
    """evolve_coarse_deflation without forecast (src/hmc_multiscale.cpp:278-338)."""
    _chk(lib().smg_coarse_evolve_deflation(defl.h, A_new.h))


def hmc_ms_evolve(D, P, defl, params, pi, phi, mg=None):
    """The multi-timescale molecular dynamics: (pi, outer CG iterations, low-mode force evaluations)."""
    pi = np.ascontiguousarray(pi, np.float64).copy()
    it, nl = _I(), _I()
    _chk(lib().smg_hmc_ms_evolve(D.h, mg.h if mg else None, P.h, defl.h, ctypes.byref(params), _p(pi),
                                 _p(_c(phi)), ctypes.byref(it), ctypes.byref(nl)))
    return pi, it.value, nl.value


def hmc_trajectory_mg_multiscale(D, P, defl, params, pi0, eta, u, mg=None):
    """hmc_trajectory_mg_multiscale (src/hmc_multiscale.cpp:340-888) from the momenta pi0
    and the pseudofermion Gaussian eta (phi = D^dag eta), u the Metropolis uniform."""
    res = MSResult()
    _chk(lib().smg_hmc_ms_trajectory(D.h, mg.h if mg else None, P.h, defl.h, ctypes.byref(params),
                                     _p(np.ascontiguousarray(pi0, np.float64)), _p(_c(eta)), u, ctypes.byref(res)))
    return res


# ------------------------------------------------------------------ nested integrator, gauge force on the fine scale
class NestedParams(ctypes.Structure):
    """The nested integrator (no reference counterpart): the fermion force on the
    outer scale, n_gauge inner gauge-force steps per outer position update."""
    _fields_ = [("beta", _D), ("tau", _D), ("n_outer", _I), ("n_gauge", _I), ("omelyan_outer", _I),
                ("omelyan_inner", _I), ("solver", _I), ("cg_maxiter", _I), ("cg_tol", _D), ("inner_tol", _D),
                ("force_accept", _I), ("mg_refresh", _I), ("mg_refresh_iters", _I), ("chrono_guess", _I)]


def nested_params(beta, tau=1.0, n_outer=1, n_gauge=8, omelyan_outer=True, omelyan_inner=True, solver=F64,
                  cg_maxiter=1000, cg_tol=1e-10, inner_tol=1e-5, force_accept=False, mg_refresh=0,
                  mg_refresh_iters=2, chrono_guess=False):
    return NestedParams(beta, tau, n_outer, n_gauge, int(omelyan_outer), int(omelyan_inner), solver, cg_maxiter,
                        cg_tol, inner_tol, int(force_accept), mg_refresh, mg_refresh_iters, int(chrono_guess))


def hmc_nested_evolve(D, params, pi, phi, mg=None):
    """The nested molecular dynamics: D's gauge field evolves in place; (pi, solver iterations)."""
    pi = np.ascontiguousarray(pi, np.float64).copy()
    it = _I()
    _chk(lib().smg_hmc_nested_evolve(D.h, mg.h if mg else None, ctypes.byref(params), _p(pi), _p(_c(phi)),
                                     ctypes.byref(it)))
    return pi, it.value


def hmc_trajectory_nested(D, params, pi0, eta, u, mg=None):
    """One trajectory of the nested integrator from the momenta pi0 and the Gaussian eta
    of phi = D^dag eta, u the Metropolis uniform: an HMCResult (dLD = 0)."""
    res = HMCResult()
    _chk(lib().smg_hmc_nested_trajectory(D.h, mg.h if mg else None, ctypes.byref(params),
                                         _p(np.ascontiguousarray(pi0, np.float64)), _p(_c(eta)), u,
                                         ctypes.byref(res)))
    return res


# ------------------------------------------------------------------ many lattices in one launch
class Lattices:
    """G Wilson-clover lattices of different sizes whose operators and CG solves run as
    one launch per kernel (no reference counterpart).  Us: the lattices' link arrays
    (each U[mu*V + s]); vectors are lists of per-lattice vectors."""

    def __init__(self, Ls, Us, masses, r=1.0, csw=0.0, mu_t=0.0):
        self.Ls = [int(L) for L in Ls]
        self.G = len(self.Ls)
        Ucat = np.concatenate([_c(U) for U in Us])
        Larr = np.ascontiguousarray(self.Ls, np.int32)
        m = np.ascontiguousarray(np.broadcast_to(masses, (self.G,)), np.float64)
        h = _P()
        _chk(lib().smg_lattices_create(ctypes.byref(h), self.G, _p(Larr), _p(Ucat), _p(m), r, csw, mu_t))
        self.h = h
        self.sizes = [2 * L * L for L in self.Ls]
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)])

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.smg_lattices_destroy(self.h)
            self.h = None

    def _cat(self, vs):
        return np.concatenate([_c(v) for v in vs])

    def _split(self, v):
        return [v[self.offsets[g]:self.offsets[g + 1]] for g in range(self.G)]

    def apply(self, kind, vs):
        src = self._cat(vs)
        dst = np.zeros_like(src)
        _chk(lib().smg_lattices_apply(self.h, kind, _p(src), _p(dst)))
        return self._split(dst)

    def cg(self, bs, max_iter, tol):
        """CG on D^dag D for every lattice at once: (solutions, iterations, true residuals)."""
        b = self._cat(bs)
        x = np.zeros_like(b)
        it = np.zeros(self.G, np.int32)
        tr = np.zeros(self.G)
        _chk(lib().smg_lattices_cg(self.h, _p(b), _p(x), max_iter, tol, _p(it), _p(tr)))
        return self._split(x), it, tr


# a host callable as an smg_op (the boundary's std::function operator)
_OPFN = ctypes.CFUNCTYPE(_I, _P, _P, _P)


class _SmgOp(ctypes.Structure):
    _fields_ = [("n", ctypes.c_size_t), ("fn", _OPFN), ("ctx", _P), ("dirac", _P), ("coarse", _P)]


def _callback_op(f, n):
    """An smg_op calling f(vector) -> vector on n complex numbers; keep the returned
    pair alive for the duration of the call."""
    def cb(ctx, src, dst):
        try:
            v = np.ctypeslib.as_array(ctypes.cast(src, ctypes.POINTER(ctypes.c_double)), (2 * n,)).view(np.complex128)
            out = np.ascontiguousarray(f(v.copy()), np.complex128)
            if out.shape != (n,):
                return 1
            ctypes.memmove(dst, out.ctypes.data, 16 * n)
            return 0
        except Exception:   # reported by the library as a failed call
            return 1
    fn = _OPFN(cb)
    return _SmgOp(n, fn, None, None, None), fn


class EvenOddMG:
    """EvenOddMG (src/eo_multigrid.h:229-287): two-level multigrid on the Schur operator
    M^dag M of D's odd half-lattice.  starts: the k near-null start vectors (k x V), as
    the reference draws them from its rng (random_vec(V, rng) k times); lambda_start:
    the power iteration's start (the reference's random_vec(V, mt19937(12345)))."""

    def __init__(self, D, bx, by, k, pre_smooth=3, post_smooth=3, null_iters=20, starts=None, lambda_start=None):
        V = D.V
        starts = np.ascontiguousarray(starts, np.complex128).reshape(k, V)
        lam = random_vec(V, 12345) if lambda_start is None else _c(lambda_start)
        h = _P()
        _chk(lib().smg_eo_mg_setup(ctypes.byref(h), D.h, bx, by, k, pre_smooth, post_smooth, null_iters, _p(starts),
                                   _p(lam)))
        self.h, self.D, self.n = h, D, V
        om, lm, cd = _D(), _D(), _I()
        _chk(lib().smg_eo_mg_info(h, ctypes.byref(om), ctypes.byref(lm), ctypes.byref(cd)))
        self.richardson_omega, self.lambda_max, self.coarse_dim = om.value, lm.value, cd.value
        self.fine_half_dim = V

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.smg_eo_mg_destroy(self.h)
            self.h = None

    def coarse_matrix(self):
        """The dense P^dag M^dag M P, M[i][j]."""
        out = np.zeros((self.coarse_dim, self.coarse_dim), np.complex128)
        _chk(lib().smg_eo_mg_coarse_matrix(self.h, _p(out)))
        return out

    def precondition(self, b, A_MdagM=None):
        """One cycle x = M b; A_MdagM: the smoothing operator (None: the setup's M^dag M on
        the device; else a callable on half-lattice vectors, called on the host)."""
        x = np.zeros(self.n, np.complex128)
        if A_MdagM is None:
            _chk(lib().smg_eo_mg_precondition(self.h, None, _p(_c(b)), _p(x)))
        else:
            op, keep = _callback_op(A_MdagM, self.n)
            _chk(lib().smg_eo_mg_precondition(self.h, ctypes.byref(op), _p(_c(b)), _p(x)))
            del keep
        return x


class LatticesMG:
    """Multigrid over the lattices of a Lattices batch (no reference counterpart): every
    lattice's hierarchy is the one MGHierarchy builds for it alone (starts: one start
    array per lattice, as MGHierarchy takes it), held in one storage per level, and
    every kernel of the cycle and of GCR runs once for all lattices (FP64 vectors and
    arithmetic; coarse_half: the coarse links read in FP16)."""

    def __init__(self, B, starts, n_levels=3, blocks=((4, 4), (2, 2)), k=(8, 8), null_iters=20, pre_smooth=3,
                 post_smooth=3, w_cycle=False, coarse_maxit=50, coarse_tol=1e-2, use_graph=False, chiral=True,
                 null_prec=F64, coarse_half=False):
        p = _mg_params(n_levels, blocks, k, null_iters, pre_smooth, post_smooth, w_cycle, coarse_maxit, coarse_tol,
                       coarse_half, 0, use_graph, chiral, null_prec)
        self.B, self.params = B, p
        cat = np.concatenate([_c(s) for s in starts])
        h = _P()
        _chk(lib().smg_lattices_mg_setup(ctypes.byref(h), B.h, ctypes.byref(p), _p(cat)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.smg_lattices_mg_destroy(self.h)
            self.h = None

    def precondition(self, bs):
        """One V-cycle on every lattice: the list of x = M b."""
        b = self.B._cat(bs)
        x = np.zeros_like(b)
        _chk(lib().smg_lattices_mg_precondition(self.h, _p(b), _p(x)))
        return self.B._split(x)

    def gcr(self, bs, max_iter, tol, trunc=7):
        """GCR(trunc) with the cycle on every lattice at once from x = 0:
        (solutions, iterations per lattice, FP64 true residuals, iterations run)."""
        b = self.B._cat(bs)
        x = np.zeros_like(b)
        it = np.zeros(self.B.G, np.int32)
        tr = np.zeros(self.B.G)
        tot = _I()
        _chk(lib().smg_lattices_gcr(self.h, _p(b), _p(x), max_iter, tol, trunc, _p(it), _p(tr), ctypes.byref(tot)))
        return self.B._split(x), it, tr, tot.value


# ------------------------------------------------------------------ U(1) heatbath
def heatbath(L, U, beta, n_sweeps, seed, first_sweep=0):
    """n_sweeps checkerboard U(1) heatbath sweeps of the Wilson action on host links
    U[mu*V + s] (no reference counterpart): returns the new links."""
    U = _c(U).copy()
    _chk(lib().smg_heatbath(L, _p(U), beta, n_sweeps, seed, first_sweep))
    return U


def dirac_heatbath(D, beta, n_sweeps, seed, first_sweep=0):
    """The heatbath on the links of a DiracOp, in place on the device."""
    _chk(lib().smg_dirac_heatbath(D.h, beta, n_sweeps, seed, first_sweep))


# ------------------------------------------------------------------ forces
def fermion_force(D, X):
    """fermion_force (src/hmc_forces.cpp:134-139): F = -dS_f/dtheta, F[mu*V + s]."""
    F = np.zeros(2 * D.V)
    _chk(lib().smg_fermion_force(D.h, _p(_c(X)), _p(F)))
    return F


def force_forward(D, X):
    chi = np.zeros(D.ndof, np.complex128)
    _chk(lib().smg_force_forward(D.h, _p(_c(X)), _p(chi)))
    return chi


def fermion_force_bilinear(D, chi, X):
    """fermion_force_bilinear (src/hmc_forces.cpp:141-194) -- the backward half."""
    F = np.zeros(2 * D.V)
    _chk(lib().smg_force_backward(D.h, _p(_c(chi)), _p(_c(X)), _p(F)))
    return F


force_backward = fermion_force_bilinear


def hopping_force_bilinear(D, chi, X):
    F = np.zeros(2 * D.V)
    _chk(lib().smg_hopping_force(D.h, _p(_c(chi)), _p(_c(X)), _p(F)))
    return F


def logdet_ee_force(D):
    F = np.zeros(2 * D.V)
    _chk(lib().smg_logdet_force(D.h, _p(F)))
    return F


def eo_fermion_force(D, x_o, y_o):
    """eo_fermion_force (src/hmc_forces.cpp:830-867)."""
    F = np.zeros(2 * D.V)
    _chk(lib().smg_eo_fermion_force(D.h, _p(_c(x_o)), _p(_c(y_o)), _p(F)))
    return F


def fermion_action(D, phi, max_iter, tol):
    """fermion_action (src/hmc_forces.cpp:113-119): returns (S, X)."""
    S = _D()
    X = np.zeros(D.ndof, np.complex128)
    _chk(lib().smg_fermion_action(D.h, _p(_c(phi)), max_iter, tol, ctypes.byref(S), _p(X), None))
    return S.value, X


def hmc_params(beta, tau, n_steps, omelyan=False, cg_maxiter=1000, cg_tol=1e-12, force_accept=False, use_eo=False):
    return HMCParams(beta, tau, n_steps, int(omelyan), cg_maxiter, cg_tol, int(force_accept), int(use_eo))


def gauge_of(D):
    """The operator's current (device) gauge field, U[mu*V + s]."""
    U = np.zeros(2 * D.V, np.complex128)
    _chk(lib().smg_dirac_get_gauge(D.h, _p(U)))
    return U


def hmc_evolve(D, params, pi, phi):
    """plain_leapfrog_evolve (src/hmc_trajectory.cpp:243-409) on the device: D's gauge
    field evolves in place; returns (pi, cg iterations)."""
    pi = np.ascontiguousarray(pi, dtype=np.float64).copy()
    it = _I()
    _chk(lib().smg_hmc_evolve(D.h, ctypes.byref(params), _p(pi), _p(_c(phi)), ctypes.byref(it)))
    return pi, it.value


def hmc_hamiltonian(D, beta, pi, phi, cg_maxiter=1000, cg_tol=1e-12, use_eo=False):
    """(KE, S_g, S_f, LD) on D's current gauge field (LD: even-odd with clover)."""
    out = np.zeros(4)
    _chk(lib().smg_hmc_hamiltonian(D.h, int(use_eo), beta, _p(np.ascontiguousarray(pi, dtype=np.float64)),
                                   _p(_c(phi)), cg_maxiter, cg_tol, _p(out), None))
    return out


def hmc_trajectory(D, params, pi0, eta, u):
    """hmc_trajectory (src/hmc_trajectory.cpp:38-236, full lattice): momenta pi0 and
    the Gaussian eta of phi = D^dag eta given; u the Metropolis uniform."""
    res = HMCResult()
    _chk(lib().smg_hmc_trajectory(D.h, ctypes.byref(params), _p(np.ascontiguousarray(pi0, dtype=np.float64)),
                                  _p(_c(eta)), u, ctypes.byref(res)))
    return res


def _hchk(rc):
    """Status of a host-only entry point (no device initialisation)."""
    if rc != 0:
        raise SMGError(load_library().smg_last_error().decode())


def gauge_save_raw(L, U, path):
    """GaugeField::save (src/gauge.h:38-46)."""
    _hchk(load_library().smg_gauge_write_raw(os.fsencode(path), L, _p(_c(U))))


def gauge_load_raw(L, path):
    """GaugeField::load (src/gauge.h:48-57)."""
    U = np.zeros(2 * L * L, np.complex128)
    _hchk(load_library().smg_gauge_read_raw(os.fsencode(path), L, _p(U)))
    return U


def save_gauge(L, U, beta, mass, path):
    """save_gauge (src/hmc_forces.cpp:18-36): 'SWNG' header, then the links."""
    _hchk(load_library().smg_gauge_write(os.fsencode(path), L, _p(_c(U)), beta, mass))


def load_gauge(L, path):
    """load_gauge (src/hmc_forces.cpp:38-62): returns (U, header)."""
    U = np.zeros(2 * L * L, np.complex128)
    h = GaugeHeader()
    _hchk(load_library().smg_gauge_read(os.fsencode(path), L, _p(U), ctypes.byref(h)))
    return U, h


def therm_cfg_path(L, beta, n_therm):
    """therm_cfg_path (src/hmc_utils.cpp:7-11)."""
    buf = ctypes.create_string_buffer(256)
    _hchk(load_library().smg_therm_cfg_path(L, beta, n_therm, buf, 256))
    return buf.value.decode()


# ------------------------------------------------------------------ launch-shape autotuner
TUNE_COARSE_BF16, TUNE_DSLASH_DOTS, TUNE_DSLASH_STORE, TUNE_CG_UPDATE, TUNE_TRANSFER, TUNE_DDAGD = 0, 1, 2, 3, 4, 5


def tune_mg(mg, nrhs):
    """Time the launch shapes of mg's batched cycle for nrhs right-hand sides and keep
    the fastest per shape; returns the number of table entries added."""
    n = _I()
    _chk(lib().smg_tune_mg(mg.h, nrhs, ctypes.byref(n)))
    return n.value


def tune_save(path):
    """The table is host memory: these calls need no GPU."""
    _hchk(load_library().smg_tune_save(str(path).encode()))


def tune_load(path):
    _hchk(load_library().smg_tune_load(str(path).encode()))


def tune_clear():
    _hchk(load_library().smg_tune_clear())


def tune_get(family, shape):
    sh = np.ascontiguousarray(shape, np.int32)
    c = _I()
    _hchk(load_library().smg_tune_get(family, _p(sh), ctypes.byref(c)))
    return c.value


def tune_set(family, shape, choice):
    sh = np.ascontiguousarray(shape, np.int32)
    _hchk(load_library().smg_tune_set(family, _p(sh), choice))

