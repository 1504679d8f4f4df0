// smg_kernels.h -- host-side launchers of the gfx950 kernels (implemented in *.hip).
#pragma once
#include "smg_common.h"

namespace smg {

// Epilogue of an operator kernel.
enum Epi : int {
    EPI_NONE = 0,   // dst = A src
    EPI_RESID = 1,  // dst = b - A src
    EPI_DOTS = 2,   // dst = A src; reduction {Re<dst,aux>, Im<dst,aux>, |dst|^2} -> Fin
    // dst = A src; reduction {Re<dst,aux>, Im<dst,aux>, Re<dst,b>, Im<dst,b>, |dst|^2} -> Fin
    // (the BF16-vector coarse apply of the coarsest CG: aux = p, b = r)
    EPI_CG = 3,
    // one fused Chebyshev step of the BF16-vector coarse apply (src = d, the previous
    // direction, read at the neighbours): q = A d, then at each output r -= q,
    // dst = c1 d + c2 r (the new direction), aux += dst (x); no reduction
    EPI_CHEB = 4,
};

// ---------------------------------------------------------------- Dirac (dirac.hip)
template <typename T>
struct DiracView {
    const cplx<T>* U;  // [2][V]
    const T* C;        // clover diagonal 0.5*csw*F01 [V], nullptr if csw == 0
    T diag;            // 2r + m
    T r;
    int L;
};

// Full-lattice Wilson-clover(-twisted) apply: dst = g5^g5out D(mu) g5^g5in src.
// D^dag = (g5in = g5out = 1, mu -> -mu_t); g5 D = (g5out = 1).  nrhs vectors at
// stride vstride complex.  Returns the number of workgroups per rhs (partials).
template <typename T>
void launch_dslash(const DiracView<T>& D, cplx<T>* dst, const cplx<T>* src, T mu, int g5in, int g5out,
                   Epi epi, const cplx<T>* b, const cplx<T>* aux, const Fin& fin, int nrhs, size_t vstride,
                   hipStream_t st);

// Half-lattice hopping with fused diagonal.  dst parity p (0 even, 1 odd).
//   mode 0: dst = H src
//   mode 1: dst = Dinv_p(mu) H src
//   mode 2: dst = g5^g5out [ D_p(mu) g5^g5in aux - H src ]
// g5in applies to src reads in modes 0/1 and to aux in mode 2.
template <typename T>
void launch_eo_hop(const DiracView<T>& D, cplx<T>* dst, const cplx<T>* src, int p, int mode, T mu, int g5in,
                   int g5out, const cplx<T>* aux, Epi epi, const cplx<T>* b, const cplx<T>* aux2, const Fin& fin,
                   hipStream_t st);
// dst = D_p(mu)^{-1} src (inv = 1) or D_p(mu) src (inv = 0) on parity p
template <typename T>
void launch_eo_diag(const DiracView<T>& D, cplx<T>* dst, const cplx<T>* src, int p, int inv, T mu, hipStream_t st);
// even-odd pack/unpack: full[s] <-> half_p[eo(s)]
template <typename T>
void launch_eo_gather(int L, cplx<T>* half, const cplx<T>* full, int p, hipStream_t st);
template <typename T>
void launch_eo_scatter(int L, cplx<T>* full, const cplx<T>* half, int p, hipStream_t st);

void launch_clover(int L, const c64* U, double* F01, hipStream_t st);
void launch_delta_dirac(const DiracView<double>& D, c64* dst, const c64* src, const double* pi, double dt,
                        hipStream_t st);
// sum over even sites of 0.5 log(d0^2+mu^2) + 0.5 log(d1^2+mu^2) (FIN_STORE into fin.out)
void launch_logdet(const DiracView<double>& D, double mu, const Fin& fin, hipStream_t st);
// sum over sites of Re P (FIN_STORE into fin.out)
void launch_plaq_sum(int L, const c64* U, const Fin& fin, hipStream_t st);

// forces, FP64, F[mu*V + s]
// F (=|+=) scale * hopping force Re(i c_bwd - i c_fwd)
void launch_hopping_force(const DiracView<double>& D, const c64* chi, const c64* X, double scale, int accumulate,
                          double* F, hipStream_t st);
// F += factor * clover-derivative insertion of the per-site weight w
void launch_clover_insert(const DiracView<double>& D, const double* w, double factor, double* F, hipStream_t st);
// w = Re(chi0^* X0 - chi1^* X1) (full lattice)
void launch_sigma3_weight(int V, const c64* chi, const c64* X, double* w, hipStream_t st);
// fused backward: F = hopping(chi, X) + (csw/4) clover insertion of w(chi, X)
void launch_force_backward(const DiracView<double>& D, double csw, const c64* chi, const c64* X, double* F,
                           hipStream_t st);
// the whole force from X: chi = D X recomputed per tile in LDS (never stored), then the
// hopping force and clover insertion as launch_force_backward
void launch_force_fused(const DiracView<double>& D, double csw, double mu, const c64* X, double* F,
                        hipStream_t st);
// e/o weights for the Schur force: w[s] on both parities
void launch_eo_force_weights(const DiracView<double>& D, const c64* chi_o, const c64* psi_o, const c64* tmp3,
                             const c64* tmp2, double* w, hipStream_t st);
// logdet weight: even sites 0.5 csw (d0/(d0^2+mu^2) - d1/(d1^2+mu^2)), odd 0
void launch_logdet_weight(const DiracView<double>& D, double csw, double mu, double* w, hipStream_t st);
void launch_gauge_force(int L, const c64* U, double beta, double* F, hipStream_t st);
// HMC molecular dynamics (hmc.hip): U <- U exp(i scale pi) with the FP32 copy
// rewritten; pi <- pi + scale (gf + ff); C = 0.5 csw F01 in FP64 and FP32
void launch_link_update(c64* U, c32* U32, const double* pi, double scale, size_t n, hipStream_t st);
void launch_mom_update(double* pi, const double* gf, const double* ff, double scale, size_t n, hipStream_t st);
void launch_clover_scale(const double* F01, double csw, double* C64, float* C32, int V, hipStream_t st);
// n_sweeps checkerboard U(1) heatbath sweeps of the Wilson action on the FP64 links
// U[mu*V + x] (L even), the first numbered first_sweep
void launch_heatbath(c64* U, int L, double beta, int n_sweeps, unsigned long long seed, unsigned first_sweep,
                     hipStream_t st);

// ---------------------------------------------------------------- BLAS-1 (blas.hip)
// reductions finished in-kernel by the scalar program of fin
template <typename T>
void launch_dot(const cplx<T>* a, const cplx<T>* b, size_t n, const Fin& fin, hipStream_t st);  // {<a,b>}
template <typename T>
void launch_norm2(const cplx<T>* a, size_t n, const Fin& fin, hipStream_t st);  // {|a|^2}
template <typename T>
void launch_dots(const cplx<T>* a, const cplx<T>* b, size_t n, const Fin& fin, hipStream_t st);  // {<a,b>, |a|^2}

// write up to 4 constants to dst (and zero the int flag at iflag if non-null)
void launch_set_scalars(double* dst, int count, double v0, double v1, double v2, double v3, int* iflag,
                        hipStream_t st);
// batched dots against up to 32 vectors: fin.out[2i..2i+1] = <V_i, w> (swap: <w, V_i>)
template <typename T>
struct VecPtrs {
    const cplx<T>* v[32];
};
template <typename T>
void launch_multi_dot(const VecPtrs<T>& V, int nv, const cplx<T>* w, size_t n, int swap, const Fin& fin,
                      hipStream_t st);
// y = sum c_i V_i (mode 0), y += sum c_i V_i (1), y -= sum c_i V_i (2); nv <= 32 vectors,
// complex coefficients c_i at coef[2i], coef[2i+1] in device memory (eigen.hip)
template <typename T>
void launch_lincomb(cplx<T>* y, const VecPtrs<T>& V, int nv, const double* coef, int mode, size_t n, hipStream_t st);
// block vectors (N vectors of n entries, v[i*N + j], N a power of two <= 32)
constexpr int kMaxBlocks = 6;
struct BlockPtrs {
    const c64* p[kMaxBlocks];
};
// G[a][b] = <L_a, R_b> over the nl*N left and nr*N right vectors, complex, into out
// (device, 2*nl*N*nr*N doubles); part holds blk_gram_grid(n) * that many doubles
int blk_gram_grid(size_t n);
void launch_blk_gram(const BlockPtrs& L, int nl, const BlockPtrs& R, int nr, int N, size_t n, double* part,
                     double* out, hipStream_t st);
// out_m = sum_s S_s C_m[s] (C_m: (ns*N) x N complex, row-major; out1/C1 may be null);
// only the first kact columns of each source are read
void launch_blk_combine(c64* out0, c64* out1, const BlockPtrs& S, int ns, int N, int kact, const c64* C0,
                        const c64* C1, size_t n, hipStream_t st);
// y_j = a_j x_j + b_j y_j per column j (x may be null)
void launch_blk_cols(c64* y, const c64* x, const c64* a, const c64* b, int N, size_t n, hipStream_t st);
// GCR orthogonalisation: beta_i = dots[i]/apap[i]; p -= sum beta_i P_i; Ap -= sum beta_i AP_i
template <typename T>
void launch_gcr_orth(cplx<T>* p, cplx<T>* Ap, const VecPtrs<T>& P, const VecPtrs<T>& AP, int nv,
                     const double* dots, const double* const* apap_slots, size_t n, hipStream_t st);

// y = a x + b y with host scalars (a, b complex); x, y may alias
template <typename T>
void launch_axpby(cplx<T>* y, const cplx<T>* x, c64 a, c64 b, size_t n, hipStream_t st);
// y += s * x, s read from device (negate: y -= s x)
template <typename T>
void launch_axpy_dev(cplx<T>* y, const cplx<T>* x, const double* s, int negate, size_t n, hipStream_t st);
// MR update: x += w r; r -= w Ar  (w at device s[0..1])
template <typename T>
void launch_mr_update(cplx<T>* x, cplx<T>* r, const cplx<T>* Ar, const double* s, size_t n, hipStream_t st);
// CG update: x += a p; r -= a Ap; reduction {|r|^2} (a at s[0..1]; skipped when flag != 0)
template <typename T>
void launch_cg_update(cplx<T>* x, cplx<T>* r, const cplx<T>* p, const cplx<T>* Ap, const double* s,
                      const int* flag, size_t n, const Fin& fin, hipStream_t st);
// p = z + beta p (beta at s[0..1], skipped when flag != 0)
template <typename T>
void launch_xpay_dev(cplx<T>* p, const cplx<T>* z, const double* s, const int* flag, size_t n, hipStream_t st);
// dst = src (precision conversion)
template <typename To, typename From>
void launch_convert(cplx<To>* dst, const cplx<From>* src, size_t n, hipStream_t st);
void launch_convert_half(ch16* dst, const c64* src, size_t n, hipStream_t st);

template <typename T>
void launch_zero(cplx<T>* p, size_t n, hipStream_t st);
// dst = src (same precision), as a kernel (graph-capturable)
template <typename T>
void launch_copy(cplx<T>* dst, const cplx<T>* src, size_t n, hipStream_t st);
// x64 += x32 (iterative refinement accumulate)
void launch_accum_mixed(c64* x, const c32* e, size_t n, hipStream_t st);
// g5 on a spinor field (negate odd components), in place
template <typename T>
void launch_g5(cplx<T>* v, size_t nsites, hipStream_t st);

// ---------------------------------------------------------------- MG (mg.hip)
struct AggView {
    int nx, ny, nd;  // fine lattice and dof per site
    int bx, by, k;   // aggregate shape and coarse dof per aggregate
    int nbx, nby;    // coarse lattice
    int chiral;      // 1: each null vector is split into its two chiralities (k = 2 * vectors)
};
// Chirality of dof c of a site: the first nd/2 dof are gamma5 = +1 (the upper spin
// component on the fine lattice, the first half of the coarse dof on a coarse one),
// the rest gamma5 = -1.  With chiral aggregation, coarse dof j of an aggregate has
// chirality j / (k/2) and its block vector vanishes on the other chirality.
__host__ __device__ inline int agg_chi_dof(const AggView& A, int c) { return c >= A.nd / 2 ? 1 : 0; }
__host__ __device__ inline int agg_chi_vec(const AggView& A, int j) { return j >= A.k / 2 ? 1 : 0; }
// copy null vector vec into block slot `slot` of bv[b][slot][loc*nd + c] (FP64);
// chi = 0 or 1 keeps only the dof of that chirality (zero elsewhere), chi = -1 all
void launch_block_gather(const AggView& A, c64* bv, const c64* vec, int slot, int chi, hipStream_t st);
// per-aggregate modified Gram-Schmidt (FP64), src/prolongator.h:111-130
void launch_block_orthonormalize(const AggView& A, c64* bv, hipStream_t st);
template <typename T>
void launch_prolong(const AggView& A, const cplx<T>* bv, const cplx<T>* coarse, cplx<T>* fine, int accumulate,
                    hipStream_t st);
template <typename T>
void launch_restrict(const AggView& A, const cplx<T>* bv, const cplx<T>* fine, cplx<T>* coarse, hipStream_t st);

template <typename TL>
struct CoarseView {
    const TL* link;  // [nb][9][k*k]
    int nx, ny, k;
    // 1: the backward blocks (-x, -y, -x+y, -x-y) are, bit for bit, the conjugate
    // transposes of the neighbours' forward blocks (+x, +y, +x-y, +x+y), so a kernel may
    // read 5 of the 9 blocks (a Hermitian operator on a lattice of extent >= 3)
    int herm = 0;
};
// the FP16 links of a Hermitian operator made exactly Hermitian: every backward block
// overwritten with the conjugate transpose of the neighbour's forward block (extents >= 3)
void launch_half_hermitian(ch16* link, int nx, int ny, int k, hipStream_t st);
template <typename T, typename TL>
void launch_coarse_apply(const CoarseView<TL>& C, cplx<T>* dst, const cplx<T>* src, Epi epi, const cplx<T>* b,
                         const cplx<T>* aux, const Fin& fin, hipStream_t st);
// dense coarsest solve: M += stencil (row-major n x n); in-place Gauss-Jordan
// inverse (HPD, no pivoting); y = M x
void launch_coarse_to_dense(const CoarseView<c64>& C, c64* M, size_t n, hipStream_t st);
void launch_dense_invert(c64* M, size_t n, c64* prow, c64* pcol, hipStream_t st);
template <typename T>
void launch_dense_matvec(const cplx<T>* M, const cplx<T>* x, cplx<T>* y, size_t n, hipStream_t st);
// Galerkin probing helpers
void launch_probe_vector(int nbx, int nby, int k, int cx, int cy, int cls, int kj, c64* e, hipStream_t st);
void launch_probe_extract(int nbx, int nby, int k, int cx, int cy, int cls, int kj, const c64* col, c64* link,
                          hipStream_t st);

// allocation-free dot product (noalloc.hip): partials, counter and result in a
// file-scope device array; noalloc_scratch() is the current device's address of it
double* noalloc_scratch();
size_t noalloc_dot_rows(size_t n);
void noalloc_dot(const void* a, const void* b, size_t n, int f64, double* out2, hipStream_t st);

}  // namespace smg
