"""Even-odd multigrid on the Schur operator M^dag M (EvenOddMG, src/eo_multigrid.h)
restated in numpy over the oracle's Schur apply.

TEST INFRASTRUCTURE ONLY: imported by tests/ and nowhere else.  Half-lattice vectors
hold V = L*L complex numbers: the odd sites in ascending site order (the reference's
Lattice::odd_sites, src/lattice.h:28-41), two spins each.
"""
import numpy as np

SCHUR_DAG_SCHUR = 7   # orc_eo_apply kind of M^dag M


def _mr(A, b, niter):
    """mr_smooth_op from x = 0 (src/smoother.cpp:26-40)."""
    x = np.zeros_like(b)
    for _ in range(niter):
        r = b - A(x)
        Ar = A(r)
        ArAr = np.vdot(Ar, Ar)
        if abs(ArAr) < 1e-30:
            break
        x = x + (np.vdot(Ar, r) / ArAr) * r
    return x


def near_null(A, starts, outer):
    """compute_near_null_space_generic (src/multigrid.cpp:545-587): inverse iteration
    with 5 MR sweeps per step, each vector orthogonalised against the earlier ones,
    then a final Gram-Schmidt pass."""
    vecs = []
    for i, s in enumerate(starts):
        v = s / np.linalg.norm(s)
        for _ in range(outer):
            xn = _mr(A, v, 5)
            for j in range(i):
                xn = xn - np.vdot(vecs[j], xn) * vecs[j]
            nv = np.linalg.norm(xn)
            if nv > 1e-14:
                v = xn / nv
            else:
                break
        vecs.append(v)
    for i in range(len(vecs)):
        for j in range(i):
            vecs[i] = vecs[i] - np.vdot(vecs[j], vecs[i]) * vecs[j]
        nv = np.linalg.norm(vecs[i])
        if nv > 1e-14:
            vecs[i] = vecs[i] / nv
    return vecs


def prolongator(L, bx, by, vecs):
    """EvenOddProlongator (src/eo_multigrid.h:26-155) as a dense V x (nblocks k) matrix:
    block b's vector kk on its odd sites (half-index order), modified Gram-Schmidt
    per block."""
    k = len(vecs)
    nbx, nby = L // bx, L // by
    odd = [s for s in range(L * L) if ((s % L) + (s // L)) % 2 == 1]
    blocks = [[] for _ in range(nbx * nby)]
    for io, s in enumerate(odd):
        blocks[(s % L) // bx + nbx * ((s // L) // by)].append(io)
    P = np.zeros((L * L, nbx * nby * k), np.complex128)
    for b, ios in enumerate(blocks):
        rows = np.array([[2 * io, 2 * io + 1] for io in ios]).ravel()
        bv = []
        for kk in range(k):
            v = vecs[kk][rows].copy()
            for w in bv:
                v = v - np.vdot(w, v) * w
            nrm = np.linalg.norm(v)
            if nrm > 1e-14:
                v = v / nrm
            bv.append(v)
            P[rows, b * k + kk] = v
    return P


def lambda_max(A, v0, niter=30):
    """estimate_lambda_max (src/smoother.cpp:43-59) from v0."""
    v = v0 / np.linalg.norm(v0)
    lam = 0.0
    for _ in range(niter):
        Av = A(v)
        lam = np.linalg.norm(Av)
        if lam < 1e-30:
            break
        v = Av / lam
    return lam


class EvenOddMG:
    """EvenOddMG::build and ::precondition (src/eo_multigrid.h:229-287) on the oracle's
    Dirac operator Do (an oracle.Dirac)."""

    def __init__(self, Do, L, bx, by, k, null_iters, pre, post, starts, lam_start):
        self.A = lambda v: Do.eo(SCHUR_DAG_SCHUR, v)
        self.P = prolongator(L, bx, by, near_null(self.A, list(starts), null_iters))
        AP = np.stack([self.A(self.P[:, j]) for j in range(self.P.shape[1])], axis=1)
        self.mat = self.P.conj().T @ AP     # EvenOddCoarseOp::build, column by column
        self.omega = 0.8 / lambda_max(self.A, lam_start)
        self.pre, self.post = pre, post

    def _richardson(self, x, b, n):
        for _ in range(n):     # richardson_smooth_op, src/smoother.cpp:63-73
            x = x + self.omega * (b - self.A(x))
        return x

    def precondition(self, b):
        x = self._richardson(np.zeros_like(b), b, self.pre)
        rc = self.P.conj().T @ (b - self.A(x))
        x = x + self.P @ np.linalg.solve(self.mat, rc)
        return self._richardson(x, b, self.post)
