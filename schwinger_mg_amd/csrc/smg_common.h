// smg_common.h -- shared device/host helpers for the gfx950 Schwinger-MG library.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <cstdint>
#include <cstddef>
#include <stdexcept>
#include <string>

namespace smg {

// ------------------------------------------------------------------ complex
template <typename T>
struct alignas(2 * sizeof(T)) cplx {
    T re, im;
};
using c64 = cplx<double>;
using c32 = cplx<float>;

template <typename T>
__host__ __device__ __forceinline__ cplx<T> mk(T a, T b) { return cplx<T>{a, b}; }
template <typename T>
__host__ __device__ __forceinline__ cplx<T> operator+(cplx<T> a, cplx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <typename T>
__host__ __device__ __forceinline__ cplx<T> operator-(cplx<T> a, cplx<T> b) { return {a.re - b.re, a.im - b.im}; }
template <typename T>
__host__ __device__ __forceinline__ cplx<T> operator-(cplx<T> a) { return {-a.re, -a.im}; }
template <typename T>
__host__ __device__ __forceinline__ cplx<T> operator*(cplx<T> a, cplx<T> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T>
__host__ __device__ __forceinline__ cplx<T> operator*(T s, cplx<T> a) { return {s * a.re, s * a.im}; }
template <typename T>
__host__ __device__ __forceinline__ cplx<T> conj(cplx<T> a) { return {a.re, -a.im}; }
// conj(a) * b
template <typename T>
__host__ __device__ __forceinline__ cplx<T> cdot(cplx<T> a, cplx<T> b) {
    return {a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re};
}
// i * a
template <typename T>
__host__ __device__ __forceinline__ cplx<T> mul_i(cplx<T> a) { return {-a.im, a.re}; }
template <typename T>
__host__ __device__ __forceinline__ T norm2(cplx<T> a) { return a.re * a.re + a.im * a.im; }
// a / b, Smith's algorithm
template <typename T>
__host__ __device__ __forceinline__ cplx<T> cdiv(cplx<T> a, cplx<T> b) {
    if (fabs(b.re) >= fabs(b.im)) {
        T r = b.im / b.re, d = b.re + r * b.im;
        return {(a.re + r * a.im) / d, (a.im - r * a.re) / d};
    }
    T r = b.re / b.im, d = b.im + r * b.re;
    return {(a.re * r + a.im) / d, (a.im * r - a.re) / d};
}
template <typename To, typename From>
__host__ __device__ __forceinline__ cplx<To> cvt(cplx<From> a) { return {(To)a.re, (To)a.im}; }

// half-precision complex storage for coarse links
struct alignas(4) ch16 {
    __half re, im;
};
template <typename T>
struct Store { using type = cplx<T>; };
__device__ __forceinline__ c32 load_c(const ch16* p) {
    ch16 v = *p;
    return {__half2float(v.re), __half2float(v.im)};
}
__device__ __forceinline__ c32 load_c(const c32* p) { return *p; }
__device__ __forceinline__ c64 load_c(const c64* p) { return *p; }

// ------------------------------------------------------------------ errors
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
#define SMG_HIP(call)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            throw ::smg::Error(-2, std::string(#call) + ": " + hipGetErrorString(e_));      \
    } while (0)
#define SMG_REQUIRE(cond, msg)                                                              \
    do {                                                                                    \
        if (!(cond)) throw ::smg::Error(-1, msg);                                           \
    } while (0)
#define SMG_LAUNCH_CHECK() SMG_HIP(hipGetLastError())

// ------------------------------------------------------------------ reductions
constexpr int kBlock = 256;         // threads per workgroup (4 wave64)
constexpr int kMaxRedBlocks = 1024; // partial sums per reduction

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Block-wide sum of NC doubles; result valid in thread 0.  blockDim.x == kBlock.
template <int NC>
__device__ __forceinline__ void block_sum(double (&v)[NC]) {
    __shared__ double sm[kBlock / 64][NC];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < NC; c++) v[c] = wave_sum(v[c]);
    if (lane == 0)
#pragma unroll
        for (int c = 0; c < NC; c++) sm[wid][c] = v[c];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int c = 0; c < NC; c++) {
            double a = 0;
            for (int w = 0; w < kBlock / 64; w++) a += sm[w][c];
            v[c] = a;
        }
    }
    __syncthreads();
}

// ------------------------------------------------------------------ fused finalize
// Scalar programs run on the summed reduction (see FinOp in smg_kernels.h).
enum FinOp : int {
    FIN_NONE = -1,      // no reduction requested
    FIN_STORE = 0,      // out[c] = sum[c]
    FIN_MR = 1,         // sum = {Re<Ar,r>, Im<Ar,r>, |Ar|^2}; out = <Ar,r>/|Ar|^2 (0 if |Ar|^2 < 1e-30)
    FIN_CG_INIT = 2,    // sum = {|r|^2}: aux[0..1] = rr, aux[2] = aux[5]*rr, aux[3] = 0, flag = rr < 1e-60
    FIN_CG_ALPHA = 3,   // sum = {Re<Ap,p>, Im<Ap,p>, .}: out = rz/pAp, pAp = conj(sum), rz = aux[0..1]
    FIN_CG_BETA = 4,    // sum = {|r|^2}: iteration count aux[3], stop test vs aux[2], out = rr/aux[0]
    FIN_PCG_BETA = 5,   // sum = {Re<r,z>, Im<r,z>}: out = sum/aux[0..1]; aux[0..1] = sum
    FIN_GCR_ALPHA = 6,  // sum = {Re<Ap,r>, Im<Ap,r>, |Ap|^2}: aux[0] = |Ap|^2; out = <Ap,r>/|Ap|^2 or 0
    // sum = {Re<q,p>, Im<q,p>, Re<q,r>, Im<q,r>, |q|^2} with q = A p: one whole CG step,
    // out = alpha, aux[6..7] = beta, |r - alpha q|^2 from the sums (see fin_program)
    FIN_CG_STEP = 7,
};
// aux layout of the CG programs: aux[0..1] rz, aux[2] stop threshold, aux[3] iterations,
// aux[4] last |r|^2, aux[5] tol^2, aux[6..7] beta (FIN_CG_STEP), int flag at (int*)(aux + 8)

// A reduction request handed to a kernel: the kernel writes one partial row per
// workgroup to part, and the workgroup that finishes last sums the rows in index
// order (deterministic) and runs the scalar program `op` -- no separate finalize
// launch.  counter is a device word that is zero between kernels (the last
// workgroup resets it).
struct Fin {
    int op = FIN_NONE;
    double* out = nullptr;
    double* aux = nullptr;
    double* part = nullptr;
    unsigned* counter = nullptr;
};

__device__ __forceinline__ void fin_program(int op, const double* v, int nc, double* out, double* aux) {
    int* flag = aux ? reinterpret_cast<int*>(aux + 8) : nullptr;
    switch (op) {
    case FIN_STORE:
        for (int c = 0; c < nc; c++) out[c] = v[c];
        break;
    case FIN_MR: {
        // src/smoother.cpp:34-37: omega = <Ar,r>/<Ar,Ar>; |<Ar,Ar>| < 1e-30 ends the
        // sweeps, which a zero step reproduces (x and r stay unchanged from then on)
        const double d = v[2];
        if (fabs(d) < 1e-30) {
            out[0] = 0.0;
            out[1] = 0.0;
        } else {
            out[0] = v[0] / d;
            out[1] = v[1] / d;
        }
        break;
    }
    case FIN_CG_INIT: {
        const double rr = v[0];
        aux[0] = rr;
        aux[1] = 0.0;
        aux[2] = aux[5] * rr;
        aux[3] = 0.0;
        aux[4] = rr;
        *flag = (rr < 1e-60) ? 1 : 0;   // |b| < 1e-30: x = 0
        break;
    }
    case FIN_CG_ALPHA: {
        // src/coarse_op.h:520-523 (pAp = <p, Ap> = conj(<Ap, p>)); a finished RHS takes a zero step
        if (*flag) {
            out[0] = 0.0;
            out[1] = 0.0;
            break;
        }
        const cplx<double> pAp = {v[0], -v[1]};
        if (sqrt(pAp.re * pAp.re + pAp.im * pAp.im) < 1e-30) {
            *flag = 2;
            out[0] = 0.0;
            out[1] = 0.0;
            break;
        }
        const cplx<double> a = cdiv(cplx<double>{aux[0], aux[1]}, pAp);
        out[0] = a.re;
        out[1] = a.im;
        break;
    }
    case FIN_CG_BETA: {
        // src/coarse_op.h:529-540: stop on |r|/|b| < tol (or NaN/Inf), beta = rr_new/rr
        if (*flag) break;
        const double rr = v[0];
        aux[3] += 1.0;
        aux[4] = rr;
        if (rr < aux[2] || isnan(rr) || isinf(rr)) {
            *flag = 1;
            break;
        }
        out[0] = rr / aux[0];
        out[1] = 0.0;
        aux[0] = rr;
        aux[1] = 0.0;
        break;
    }
    case FIN_CG_STEP: {
        // FIN_CG_ALPHA and FIN_CG_BETA in one program, for an update kernel that applies
        // r -= alpha q, x += alpha p, p = r + beta p in one pass.  The new |r|^2 comes from
        // the apply's sums, |r - alpha q|^2 = |r|^2 - 2 Re(alpha conj(<q,r>)) + |alpha|^2 |q|^2
        // (<q,r> = sum conj(q) r).  A finished RHS takes zero steps from then on (alpha =
        // beta = 0: x and r stay, p = r).
        aux[6] = 0.0;
        aux[7] = 0.0;
        out[0] = 0.0;
        out[1] = 0.0;
        if (*flag) break;
        const cplx<double> pAp = {v[0], -v[1]};
        if (sqrt(pAp.re * pAp.re + pAp.im * pAp.im) < 1e-30) {
            *flag = 2;
            break;
        }
        const cplx<double> a = cdiv(cplx<double>{aux[0], aux[1]}, pAp);
        out[0] = a.re;
        out[1] = a.im;
        const double rr = aux[0] - 2.0 * (a.re * v[2] + a.im * v[3]) + (a.re * a.re + a.im * a.im) * v[4];
        aux[3] += 1.0;
        aux[4] = rr;
        if (rr < aux[2] || isnan(rr) || isinf(rr)) {
            *flag = 1;
            break;
        }
        aux[6] = rr / aux[0];
        aux[0] = rr;
        aux[1] = 0.0;
        break;
    }
    case FIN_PCG_BETA: {
        const cplx<double> b = cdiv(cplx<double>{v[0], v[1]}, cplx<double>{aux[0], aux[1]});
        out[0] = b.re;
        out[1] = b.im;
        aux[0] = v[0];
        aux[1] = v[1];
        break;
    }
    case FIN_GCR_ALPHA: {
        const double apap = v[2];
        aux[0] = apap;
        if (apap < 1e-30) {
            out[0] = 0.0;
            out[1] = 0.0;
        } else {
            out[0] = v[0] / apap;
            out[1] = v[1] / apap;
        }
        break;
    }
    default:
        break;
    }
}

// Cross-workgroup hand-off of reduction partials.  The partials, and the
// completion counters, are written and read with device-scope (agent) atomic
// accesses, which are coherent across the XCDs' L2 caches; nothing else in L2 is
// written back or invalidated (a device-scope __threadfence() would do both, on
// every workgroup, and evict the data the other workgroups are streaming).  A
// wave's stores are complete once s_waitcnt returns, so every wave waits for its
// own partial stores before the workgroup barrier that precedes the counter
// increment; the workgroup that sees the last increment reads the partials with
// device-scope loads.
__device__ __forceinline__ void part_store(double* p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double part_load(const double* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// s_waitcnt orders the hardware, but the compiler models it as touching no memory, so
// compiler-only fences on both sides keep the partial stores before it and the
// counter increment after it (no cache writeback is emitted for a signal fence).
__device__ __forceinline__ void stores_complete() {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}
__device__ __forceinline__ unsigned counter_arrive(unsigned* c) {
    return __hip_atomic_fetch_add(c, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void counter_reset(unsigned* c) {
    __hip_atomic_store(c, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// End of a reducing kernel: every thread holds its partial v[NC]; all threads of
// every workgroup must call this.  nblk = total workgroups of the launch.
template <int NC>
__device__ __forceinline__ void reduce_finish(double (&v)[NC], const Fin& f, int nblk, int blk) {
    block_sum<NC>(v);
    __shared__ int last;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int c = 0; c < NC; c++) part_store(f.part + (size_t)blk * NC + c, v[c]);
        stores_complete();
        const unsigned prev = counter_arrive(f.counter);
        last = (prev == (unsigned)nblk - 1u);
    }
    __syncthreads();
    if (!last) return;
    double w[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) w[c] = 0.0;
    // fixed-order sum: thread t adds rows t, t + kBlock, ... ; then the block tree
    for (int i = threadIdx.x; i < nblk; i += kBlock)
#pragma unroll
        for (int c = 0; c < NC; c++) w[c] += part_load(f.part + (size_t)i * NC + c);
    block_sum<NC>(w);
    if (threadIdx.x == 0) {
        fin_program(f.op, w, NC, f.out, f.aux);
        counter_reset(f.counter);
    }
}

// Number of partial-sum blocks used for a reduction over n items: a function of
// n only, so the summation order is deterministic.
inline int red_blocks(size_t n) {
    size_t b = (n + kBlock - 1) / kBlock;
    return (int)(b < 1 ? 1 : (b > (size_t)kMaxRedBlocks ? kMaxRedBlocks : b));
}
inline int grid_for(size_t n) { return (int)((n + kBlock - 1) / kBlock); }

}  // namespace smg
