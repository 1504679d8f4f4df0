// eigen.h -- eigensolvers and deflation: the dense Hermitian eigensolver of the
// small Lanczos problems (host), thick-restart Lanczos on a device operator, and
// the deflated solves built on its eigenpairs.
#pragma once
#include <vector>

#include "runtime.h"
#include "runtime_b.h"

namespace smg {

// lanczos_eigen (src/eigensolver_core.cpp:103-265) for a small dense Hermitian
// matrix on the host: A column-major (A[j*n + i]); ascending evals; evecs[row*n + col]
void dense_eigh(int n, const c64* A, double* evals, c64* evecs);

// trlm_eigensolver without the Chebyshev filter (src/eigensolver_core.cpp:691-1057):
// the n_ev smallest eigenpairs of the Hermitian operator A.  start: the device start
// vector (normalised inside).  The Krylov vectors stay on the device; the full
// reorthogonalisation is two passes of classical Gram-Schmidt (batched dots and one
// linear combination per pass) in place of the reference's sequential modified
// Gram-Schmidt, equal in exact arithmetic.
struct Eigenpairs {
    std::vector<double> vals;
    Field<double> vecs;   // n_ev vectors of A.n(), one after the other
    int restarts = 0;
    int iterations = 0;   // operator applications
    bool converged = false;
};
Eigenpairs trlm(LinOp<double>& A, int n_ev, int n_kr, int max_restarts, double tol, const c64* start,
                hipStream_t st);

// Extreme eigenvalue estimates of the Hermitian operator A: the smallest and largest
// eigenvalues of the tridiagonal of m Lanczos steps (full reorthogonalisation) from
// the device start vector.  lo bounds the smallest eigenvalue from above, hi the
// largest from below.  No reference counterpart: they give the spectral interval of
// the coarsest level's Chebyshev iteration.
void lanczos_extremes(LinOp<double>& A, int m, const c64* start, double& lo, double& hi, hipStream_t st);

// SparseCoarseOp::solve with deflation vectors (src/coarse_op.h:492-547): x0 =
// sum_i (v_i^dag b / lambda_i) v_i, then CG; returns the iterations
int deflated_coarse_cg(LinOp<double>& A, const Eigenpairs& E, const c64* b, c64* x, int maxit, double tol,
                       hipStream_t st);

// cg_solve_deflated (src/solvers.cpp:111-213): projected (preconditioned) CG with the
// deflation space U (ndefl device vectors), E = U^dag A U inverted on the host
SolveResult cg_deflated(LinOp<double>& A, MG* mg, const c64* U, int ndefl, const c64* b, c64* x, int maxit,
                        double tol, hipStream_t st);

// rr_evolve (src/eigen_evolve.cpp:48-108): the space E.vecs Rayleigh-Ritz rotated on A
// (M_ij = <v_i, A v_j>, i <= j, diagonalised by dense_eigh), E.vals the Ritz values
void rr_evolve(LinOp<double>& A, Eigenpairs& E, hipStream_t st);
// CoarseDeflState::lowmode_solution (src/hmc.h:311-323): out = P (sum_i v_i v_i^dag / l_i)
// P^dag v over the l_i > 1e-14, P the aggregation whose coarse space holds E
void lowmode_solution(const Aggregation& P, const Eigenpairs& E, const c64* v, c64* out, hipStream_t st);

// lobpcg_update (src/eigensolver_core.cpp:279-533) on the k vectors X0 (device,
// [k][n]) as one block of N = next power of two >= k columns: every operator and
// preconditioner application is one multi-right-hand-side launch, the Gram matrices
// of [X, W, P] are block contractions, and the orthonormalisations are Cholesky QR
// (equal in exact arithmetic to the reference's modified Gram-Schmidt, with AX
// updated as the reference updates it).  mg: the FP64 cycle as preconditioner, or
// null for the identity.  k <= 32.
// The Ritz pairs harvested inside fgmres_solve_generic (src/solvers.cpp:601-611,
// 693-751): after a restart cycle of kk >= n_ritz steps, the n_ritz lowest
// eigenpairs (dense_eigh) of G = H~^dag H~, H~ the unreduced (kk+1) x kk Hessenberg,
// give the normalised vectors sum_j evecs[j][p] Z_j; beyond 2 n_ritz stored pairs the
// n_ritz of smallest value are kept.  Later preconditioned vectors are deflated by
// every stored pair of value > 1e-14: z += (<v_q, v> / l_q) v_q.
struct RitzHarvest {
    int n_ritz;
    size_t n;
    std::vector<double> vals;
    Field<double> vecs;   // room for 3 n_ritz vectors of n
    DevMem coef;
    RitzHarvest(int n_ritz, int restart, size_t n);
    int count() const { return (int)vals.size(); }
    void deflate(const c64* v, c64* z, hipStream_t st);
    // Hraw row-major with ld columns (rows 0..kk); Z: kk vectors, stride n
    void harvest(const std::vector<c64>& Hraw, int ld, int kk, const c64* Z, hipStream_t st);
};

struct LobpcgResult {
    std::vector<double> vals;
    Field<double> vecs;   // k vectors of A.n(), one after the other
    int iterations = 0;
};
// exact_ax: the re-orthonormalisation keeps AX = A X (AX <- AX R^-1) in place of the
// reference's update
LobpcgResult lobpcg(LinOpB<double>& A, MG* mg, int k, const c64* X0, int max_iter, double tol, hipStream_t st,
                    bool exact_ax = false);

}  // namespace smg
