// runtime_b.h -- multi-right-hand-side solve path (N right-hand sides per launch).
#pragma once
#include <map>
#include "runtime.h"
#include "smg_batched.h"

namespace smg {

template <typename T>
struct LinOpB {
    int N;
    explicit LinOpB(int n_) : N(n_) {}
    virtual ~LinOpB() = default;
    virtual size_t n() const = 0;   // entries per right-hand side
    virtual void apply(cplx<T>* dst, const cplx<T>* src, Epi epi, const cplx<T>* b, const cplx<T>* aux,
                       const FinB& fin, hipStream_t st) = 0;
    void apply(cplx<T>* dst, const cplx<T>* src, hipStream_t st) {
        apply(dst, src, EPI_NONE, nullptr, nullptr, FinB(), st);
    }
    // whether cheb_step runs as one launch for this operator and N
    virtual bool fused_cheb() const { return false; }
    // one Chebyshev step in the apply's epilogue, as f's EPI_CHEB fields give it:
    // r_new = r - s A d, d_new = c1 s d + c2 r_new, x updated (FinB)
    virtual void cheb_step(cplx<T>* d_new, const cplx<T>* d, const cplx<T>* r, cplx<T>* x, const FinB& f,
                           hipStream_t st) {
        SMG_REQUIRE(false, "Chebyshev step: no fused form for this operator");
    }
};

template <typename T>
struct DdagDOpB : LinOpB<T> {
    const Dirac& D;
    Field<T> tmp;
    DdagDOpB(const Dirac& d, int N) : LinOpB<T>(N), D(d), tmp(2 * (size_t)d.V * N) {}
    size_t n() const override { return 2 * (size_t)D.V; }
    void apply(cplx<T>* dst, const cplx<T>* src, Epi epi, const cplx<T>* b, const cplx<T>* aux, const FinB& fin,
               hipStream_t st) override;
    bool fused_cheb() const override { return b_ddagd_fits(D.L, this->N, sizeof(T)); }
    void cheb_step(cplx<T>* d_new, const cplx<T>* d, const cplx<T>* r, cplx<T>* x, const FinB& f,
                   hipStream_t st) override;
    using LinOpB<T>::apply;
};

template <typename T, typename TL>
struct CoarseLinOpB : LinOpB<T> {
    const CoarseOperator& C;
    CoarseLinOpB(const CoarseOperator& c, int N) : LinOpB<T>(N), C(c) {}
    size_t n() const override { return C.n(); }
    void apply(cplx<T>* dst, const cplx<T>* src, Epi epi, const cplx<T>* b, const cplx<T>* aux, const FinB& fin,
               hipStream_t st) override;
    bool fused_cheb() const override {
        return b_coarse_cheb_fits(C.k, this->N, std::is_same<T, float>::value && std::is_same<TL, ch16>::value);
    }
    void cheb_step(cplx<T>* d_new, const cplx<T>* d, const cplx<T>* r, cplx<T>* x, const FinB& f,
                   hipStream_t st) override;
    using LinOpB<T>::apply;
};

// The W/V cycle of MG on N right-hand sides at once (same hierarchy, per-RHS
// smoother and coarse-solve scalars).
template <typename T>
struct CycleB;

template <typename T>
CycleB<T>& batched_cycle(MG& mg, int N);
template <typename T>
void mg_precondition_b(MG& mg, const cplx<T>* b, cplx<T>* x, int N, hipStream_t st);
// the coarsest-level solve of the batched FP32 cycle on N right-hand sides (FP32
// batched layout, coarsest dimension): the BF16 CG, the BF16 Chebyshev iteration, FP32
// CG or the dense inverse, as cfg selects
void coarsest_solve_b32(MG& mg, const c32* b, c32* x, int N, hipStream_t st);
// drop the batched cycles built for mg (before mg is destroyed)
void release_batched_cycles(MG* mg);
// the autotuner's timing search over the launch shapes of mg's batched cycle for N
// right-hand sides (tune.h); returns the number of new table entries
int tune_mg(MG& mg, int N, hipStream_t st);

// GCR(trunc) on N right-hand sides in lockstep (independent Krylov spaces, per-RHS
// scalars), MG preconditioned; batched layouts; returns per-RHS relative residuals
// in res (host, N entries) and the iteration count.
template <typename T>
int gcr_solve_b(LinOpB<T>& A, const cplx<T>* b, cplx<T>* x, MG* mg, int maxit, double tol, int trunc,
                double* res, hipStream_t st);

struct BatchedSolveResult {
    int iterations = 0;   // inner iterations (summed over refinements)
    int restarts = 0;
    std::vector<double> true_residual;
};
// FP64 (prec = 0) or mixed (prec = 1: FP32 GCR + FP32/FP16 cycle in FP64 refinement)
// solve of N right-hand sides given in the separate-vector layout [j][i] (FP64).
BatchedSolveResult gcr_solve_batched(const Dirac& D, MG* mg, int prec, const c64* B, c64* X, int N, int maxit,
                                     double tol, int trunc, double inner_tol, hipStream_t st);

}  // namespace smg
