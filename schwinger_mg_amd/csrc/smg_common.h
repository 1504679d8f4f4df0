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

// Number of partial-sum blocks used for a reduction over n items: a function of
// n only, so the summation order is deterministic.
inline int red_blocks(size_t n) {
    size_t b = (n + kBlock - 1) / kBlock;
    return (int)(b < 1 ? 1 : (b > (size_t)kMaxRedBlocks ? kMaxRedBlocks : b));
}
inline int grid_for(size_t n) { return (int)((n + kBlock - 1) / kBlock); }

}  // namespace smg
