// smg_batched.h -- launchers of the multi-right-hand-side ("batched") kernels.
//
// A batched field holds N right-hand sides with the RHS index innermost:
//   v[(i * N) + j], i = site*nd + dof, j = 0..N-1, N a power of two <= 64.
// Work-items map to (i, j) with j fastest, so a wave64 reads N consecutive
// complex numbers per dof (coalesced) and shares the site's links, clover term
// and aggregate vectors across the N lanes of one site (broadcast loads).
// Every scalar of the Krylov and smoother programs is per RHS: out[j*ostride..],
// aux[j*astride..].
#pragma once
#include "smg_kernels.h"

namespace smg {

struct FinB {
    int op = FIN_NONE;
    double* out = nullptr;
    int ostride = 2;
    double* aux = nullptr;
    int astride = 16;
    double* part = nullptr;
    unsigned* counter = nullptr;
    // CG early exit: the finish of the update writes *done = 1 once every RHS's
    // convergence flag (int at aux_j + 8) is set; kernels given skip return at
    // entry while *skip != 0 (grid-uniform, written by an earlier kernel)
    int* done = nullptr;
    const int* skip = nullptr;
    // set to 1 by a kernel that returned at entry on skip (the kernel after it in the
    // same iteration then returns too: its coefficients were not refreshed)
    int* late = nullptr;
    // EPI_CHEB: the step's coefficients d_new = c1 d + c2 r (fixed by the iteration
    // number and the spectral interval, the same for every right-hand side); what the
    // step writes (FP32 vectors): 0 r, d_new and x; 1 r only (the last step of a smoothing
    // that returns its residual; x too from a start); 2 x only (the last step of one
    // that does not)
    double cheb_c1 = 0.0, cheb_c2 = 0.0;
    int cheb_mode = 0;
    // FP32 vectors: the first step of a smoothing takes the place of its initial pass.
    // cheb_start 1 (from x = 0): the input d and r are both b, the direction b / theta
    // (cheb_s = 1 / theta scales A d and d), x_1 = b / theta is not read; 2 (from the
    // residual r0 = b - A x0): d and r are both r0, scaled the same way, and
    // x_1 = x0 + r0 / theta.  The new r then goes to cheb_rout (r is an input the step
    // must not overwrite); 0 continues (cheb_s = 1, r in place unless cheb_rout is given)
    int cheb_start = 0;
    double cheb_s = 1.0;
    void* cheb_rout = nullptr;
    // 0 x += d_new; 1 x left alone (neither read nor written); 2 x += d + d_new, d the
    // direction the step reads (the previous step's d_new, which the step before did not
    // add), plus cheb_xadd_s times the vector cheb_xadd where given (FP32 and FP64
    // vectors: the scaled residual a first step from x0 did not add)
    int cheb_xacc = 0;
    const void* cheb_xadd = nullptr;
    double cheb_xadd_s = 0.0;
};

// Per-RHS coefficient array views (device): complex a_j at s[2j], s[2j+1].
template <typename T>
void launch_b_dslash(const DiracView<T>& D, cplx<T>* dst, const cplx<T>* src, T mu, int g5in, int g5out, Epi epi,
                     const cplx<T>* b, const cplx<T>* aux, const FinB& fin, int N, hipStream_t st);
// D^dag D in one launch (the tile's intermediate D x kept in LDS); false when the shape
// does not fit it (L a multiple of 16, N a multiple of 16 in FP32 or 8 in FP64) or the
// tuner chose the two launches, and the caller runs launch_b_dslash twice
template <typename T>
bool launch_b_ddagd(const DiracView<T>& D, cplx<T>* dst, const cplx<T>* src, T mu, Epi epi, const cplx<T>* b,
                    const cplx<T>* aux, const FinB& fin, int N, hipStream_t st);
bool b_ddagd_fits(int L, int N, size_t real_bytes);   // the fused launch's shape condition
// whether launch_b_coarse_apply takes EPI_CHEB (FP16 links with FP32 vectors on the
// matrix cores: k in {8, 16, 32}, N in {16, 32, 64})
bool b_coarse_cheb_fits(int k, int N, bool f16_links_f32);
template <typename T, typename TL>
void launch_b_coarse_apply(const CoarseView<TL>& C, cplx<T>* dst, const cplx<T>* src, Epi epi, const cplx<T>* b,
                           const cplx<T>* aux, const FinB& fin, int N, hipStream_t st);
// LDS bytes of the staged (workgroup-per-aggregate) transfer for this aggregation, N and
// bytes per complex, or 0 where it does not apply; the tuner's TUNE_TRANSFER choice 1
// selects it
size_t b_transfer_lds_bytes(const AggView& A, int N, size_t cbytes, bool prolong);
// whether the matrix-core transfers (FP32; the tuner's TUNE_TRANSFER choice 2) apply:
// N a multiple of 16, 16 or 32 block vectors and 16, 32 or 64 fine entries per
// chirality half of an aggregate
bool b_transfer_mfma_fits(const AggView& A, int N);
template <typename T>
void launch_b_prolong(const AggView& A, const cplx<T>* bv, const cplx<T>* coarse, cplx<T>* fine, int accumulate,
                      int N, hipStream_t st, const cplx<T>* base = nullptr, double bs = 0.0);
template <typename T>
void launch_b_restrict(const AggView& A, const cplx<T>* bv, const cplx<T>* fine, cplx<T>* coarse, int N,
                       hipStream_t st);
// y = M x with the dense n x n inverse M; work holds b_dense_work(n, N) entries
// (the matrix-core path's column-slice partials; nullptr selects the plain kernel)
size_t b_dense_work(size_t n, int N);
template <typename T>
void launch_b_dense_apply(const cplx<T>* M, const cplx<T>* x, cplx<T>* y, size_t n, int N, cplx<T>* work,
                          hipStream_t st);

// reductions over the vector index, one result per RHS
template <typename T>
void launch_b_norm2(const cplx<T>* a, size_t n, const FinB& fin, int N, hipStream_t st);         // {|a_j|^2}
template <typename T>
void launch_b_dots(const cplx<T>* a, const cplx<T>* b, size_t n, const FinB& fin, int N, hipStream_t st);  // {<a,b>,|a|^2}
// fin.out[j*2*nv + 2i..] = <V_i[:,j], w[:,j]> for i < nv (nv <= 16)
template <typename T>
void launch_b_multi_dot(const VecPtrs<T>& V, int nv, const cplx<T>* w, size_t n, const FinB& fin, int N,
                        hipStream_t st);
// p -= sum_i beta_ij P_i ; Ap -= sum_i beta_ij AP_i, beta_ij = dots[j*32 + 2i..] / apap_i[j];
// with fin, also {<Ap, r>, |Ap|^2} of the new Ap per RHS reduced into fin
template <typename T>
void launch_b_gcr_orth(cplx<T>* p, cplx<T>* Ap, const VecPtrs<T>& P, const VecPtrs<T>& AP, int nv,
                       const double* dots, const double* const* apap, size_t n, int N, hipStream_t st,
                       const cplx<T>* r = nullptr, const FinB* fin = nullptr);
// mode 0: x += w_j r ; r -= w_j Ar   (w_j at s[2j..]); 1: x = w_j b ; r = b - w_j Ar (the
// first sweep from x = 0); 2: x += w_j r only (r is not read again)
template <typename T>
void launch_b_mr_update(cplx<T>* x, cplx<T>* r, const cplx<T>* Ar, const double* s, size_t n, int N, hipStream_t st,
                        int mode = 0, const cplx<T>* b = nullptr);
// x (+, with accumulate)= sum_{i < m} a_ij P_i, a_ij complex at coef[i][2j..] (m <= 16)
template <typename T>
void launch_b_combine(cplx<T>* x, const VecPtrs<T>& P, const double* const* coef, int m, int accumulate, size_t n,
                      int N, hipStream_t st);
// x += a_j p ; r -= a_j Ap with {|r_j|^2} -> fin (x = nullptr: r only)
template <typename T>
void launch_b_update(cplx<T>* x, cplx<T>* r, const cplx<T>* p, const cplx<T>* Ap, const double* s, size_t n,
                     const FinB& fin, int N, hipStream_t st);
// p = z + b_j p (b_j at s[2j..]); nothing while skip && *skip
template <typename T>
void launch_b_xpay(cplx<T>* p, const cplx<T>* z, const double* s, size_t n, int N, hipStream_t st,
                   const int* skip = nullptr);
// per-RHS CG program state: aux_j[5] = tol^2, flag_j = 0; *done = *late = 0
void launch_b_cg_init(double* aux, int astride, int N, double tol2, int* done, int* late, hipStream_t st);
// CG vector passes after the apply: r -= a_j Ap with |r|^2 finished by fin (returns
// at entry and sets *late when *fin.skip), then x += a_j p, p = r + b_j p (nothing
// once *late is set)
template <typename T>
void launch_b_cg_r(cplx<T>* r, const cplx<T>* Ap, const double* a, size_t n, const FinB& fin, int N, int* late,
                   hipStream_t st);
template <typename T>
void launch_b_cg_xp(cplx<T>* x, cplx<T>* p, const cplx<T>* r, const double* a, const double* b, size_t n, int N,
                    const int* late, hipStream_t st);
// y = a x + c y with one real scalar pair for all RHS
template <typename T>
void launch_b_axpby_real(cplx<T>* y, const cplx<T>* x, double a, double c, size_t n, hipStream_t st);
// layout transforms: separate vectors [j][i] (n entries each) <-> batched [i][j], with precision change
template <typename To, typename From>
void launch_b_pack(cplx<To>* batched, const cplx<From>* sep, size_t n, int N, hipStream_t st);
template <typename To, typename From>
void launch_b_unpack(cplx<To>* sep, const cplx<From>* batched, size_t n, int N, hipStream_t st);
// x64 += e32 (batched layouts); r = b - Ax with {|r_j|^2}
void launch_b_accum_mixed(c64* x, const c32* e, size_t n, int N, hipStream_t st);

// BF16 vectors (one dword per complex entry: bf16 Re low, bf16 Im high), N in
// {16, 32, 64}: the coarse apply with FP16 links (EPI_NONE, EPI_DOTS, or EPI_CG with
// the second partner b) and the CG's vector update; init: r = p = bf16(b), x = 0 with
// |r|^2 -> fin; to_f32 converts a batched BF16 vector to FP32
using cb16 = unsigned;
void launch_b_coarse_apply_h(const CoarseView<ch16>& C, cb16* dst, const cb16* src, Epi epi, const cb16* aux,
                             const FinB& fin, int N, hipStream_t st, const cb16* b = nullptr);
void launch_h_cg_init(const c32* b, cb16* r, cb16* p, cb16* x, size_t n, const FinB& fin, int N, hipStream_t st);
// one CG step after the apply q = A p with EPI_CG and FIN_CG_STEP: r -= a_j q,
// x += a_j p, p = r + b_j p (a_j at a[2j], b_j at aux[16j + 6]); returns at entry
// when *late is set
void launch_h_cg_step(cb16* x, cb16* p, cb16* r, const cb16* q, const double* a, const double* aux, size_t n, int N,
                      const int* late, hipStream_t st);
void launch_h_to_f32(c32* out, const cb16* in, size_t n, int N, hipStream_t st);
// the first step of the Chebyshev iteration from x = 0: r = bf16(b), d = bf16(b / theta),
// x = d; each later step is the coarse apply with EPI_CHEB
void launch_h_cheb_init(const c32* b, cb16* r, cb16* d, cb16* x, size_t n, int N, double theta, hipStream_t st);
void launch_h_from_f32(cb16* out, const c32* in, size_t n, int N, hipStream_t st);

}  // namespace smg
