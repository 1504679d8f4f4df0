// blas.hip -- BLAS-1 and reductions (src/linalg.cpp:8-40) for gfx950.
//
// Reductions accumulate in FP64 whatever the storage precision: per-thread
// partial, wave64 shuffle tree, LDS across the 4 waves, one partial per
// workgroup; the workgroup that finishes last sums the partials in index order
// (deterministic) and runs the scalar program (reduce_finish), so no separate
// finalize launch.  Scalars that feed later kernels (MR omega, CG alpha/beta,
// GCR alpha) stay in device memory: a smoother or coarse solve issues no host
// synchronisation.
#include "smg_kernels.h"

namespace smg {

template <typename T>
__global__ __launch_bounds__(kBlock) void k_dot(const cplx<T>* __restrict__ a, const cplx<T>* __restrict__ b,
                                                size_t n, Fin fin) {
    double v[2] = {0.0, 0.0};
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        const cplx<T> p = cdot(a[i], b[i]);
        v[0] += (double)p.re;
        v[1] += (double)p.im;
    }
    reduce_finish<2>(v, fin, gridDim.x, blockIdx.x);
}
template <typename T>
void launch_dot(const cplx<T>* a, const cplx<T>* b, size_t n, const Fin& fin, hipStream_t st) {
    k_dot<T><<<red_blocks(n), kBlock, 0, st>>>(a, b, n, fin);
    SMG_LAUNCH_CHECK();
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_norm2(const cplx<T>* __restrict__ a, size_t n, Fin fin) {
    double v[1] = {0.0};
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        const cplx<T> p = a[i];
        v[0] += (double)p.re * (double)p.re + (double)p.im * (double)p.im;
    }
    reduce_finish<1>(v, fin, gridDim.x, blockIdx.x);
}
template <typename T>
void launch_norm2(const cplx<T>* a, size_t n, const Fin& fin, hipStream_t st) {
    k_norm2<T><<<red_blocks(n), kBlock, 0, st>>>(a, n, fin);
    SMG_LAUNCH_CHECK();
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_dots(const cplx<T>* __restrict__ a, const cplx<T>* __restrict__ b,
                                                 size_t n, Fin fin) {
    double v[3] = {0.0, 0.0, 0.0};
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        const cplx<T> x = a[i];
        const cplx<T> p = cdot(x, b[i]);
        v[0] += (double)p.re;
        v[1] += (double)p.im;
        v[2] += (double)x.re * (double)x.re + (double)x.im * (double)x.im;
    }
    reduce_finish<3>(v, fin, gridDim.x, blockIdx.x);
}
template <typename T>
void launch_dots(const cplx<T>* a, const cplx<T>* b, size_t n, const Fin& fin, hipStream_t st) {
    k_dots<T><<<red_blocks(n), kBlock, 0, st>>>(a, b, n, fin);
    SMG_LAUNCH_CHECK();
}

__global__ void k_set_scalars(double* dst, int count, double v0, double v1, double v2, double v3, int* iflag) {
    if (threadIdx.x != 0) return;
    const double v[4] = {v0, v1, v2, v3};
    for (int i = 0; i < count; i++) dst[i] = v[i];
    if (iflag) *iflag = 0;
}
void launch_set_scalars(double* dst, int count, double v0, double v1, double v2, double v3, int* iflag,
                        hipStream_t st) {
    k_set_scalars<<<1, 64, 0, st>>>(dst, count, v0, v1, v2, v3, iflag);
    SMG_LAUNCH_CHECK();
}

// <V_i, w> for i < nv (swap: <w, V_i>), FP64 accumulation
template <typename T>
__global__ __launch_bounds__(kBlock) void k_multi_dot(VecPtrs<T> V, int nv, const cplx<T>* __restrict__ w,
                                                      size_t n, int swap, Fin fin) {
    double acc[64];
#pragma unroll
    for (int c = 0; c < 64; c++) acc[c] = 0.0;
    for (size_t e = (size_t)blockIdx.x * kBlock + threadIdx.x; e < n; e += (size_t)gridDim.x * kBlock) {
        const cplx<T> we = w[e];
#pragma unroll
        for (int i = 0; i < 32; i++) {
            if (i < nv) {
                const cplx<T> p = cdot(V.v[i][e], we);
                acc[2 * i] += (double)p.re;
                acc[2 * i + 1] += swap ? -(double)p.im : (double)p.im;
            }
        }
    }
    reduce_finish<64>(acc, fin, gridDim.x, blockIdx.x);
}
template <typename T>
void launch_multi_dot(const VecPtrs<T>& V, int nv, const cplx<T>* w, size_t n, int swap, const Fin& fin,
                      hipStream_t st) {
    SMG_REQUIRE(nv >= 1 && nv <= 32, "multi_dot: 1..32 vectors");
    k_multi_dot<T><<<red_blocks(n), kBlock, 0, st>>>(V, nv, w, n, swap, fin);
    SMG_LAUNCH_CHECK();
}

template <typename T>
struct ApapPtrs {
    const double* a[32];
};
template <typename T>
__global__ __launch_bounds__(kBlock) void k_gcr_orth(cplx<T>* __restrict__ p, cplx<T>* __restrict__ Ap,
                                                     VecPtrs<T> P, VecPtrs<T> AP, int nv,
                                                     const double* __restrict__ dots, ApapPtrs<T> apap, size_t n) {
    __shared__ cplx<T> beta[32];
    if (threadIdx.x < nv) {
        const double d = *apap.a[threadIdx.x];
        beta[threadIdx.x] = cplx<T>{(T)(dots[2 * threadIdx.x] / d), (T)(dots[2 * threadIdx.x + 1] / d)};
    }
    __syncthreads();
    const size_t e = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (e >= n) return;
    cplx<T> pe = p[e], ae = Ap[e];
    for (int i = 0; i < nv; i++) {
        pe = pe - beta[i] * P.v[i][e];
        ae = ae - beta[i] * AP.v[i][e];
    }
    p[e] = pe;
    Ap[e] = ae;
}
template <typename T>
void launch_gcr_orth(cplx<T>* p, cplx<T>* Ap, const VecPtrs<T>& P, const VecPtrs<T>& AP, int nv,
                     const double* dots, const double* const* apap_slots, size_t n, hipStream_t st) {
    ApapPtrs<T> a;
    for (int i = 0; i < nv; i++) a.a[i] = apap_slots[i];
    k_gcr_orth<T><<<grid_for(n), kBlock, 0, st>>>(p, Ap, P, AP, nv, dots, a, n);
    SMG_LAUNCH_CHECK();
}

template <typename T>
__global__ void k_axpby(cplx<T>* __restrict__ y, const cplx<T>* __restrict__ x, c64 a, c64 b, size_t n) {
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const cplx<T> ta = cvt<T>(a), tb = cvt<T>(b);
    y[i] = ta * x[i] + tb * y[i];
}
template <typename T>
void launch_axpby(cplx<T>* y, const cplx<T>* x, c64 a, c64 b, size_t n, hipStream_t st) {
    k_axpby<T><<<grid_for(n), kBlock, 0, st>>>(y, x, a, b, n);
    SMG_LAUNCH_CHECK();
}

template <typename T>
__global__ void k_axpy_dev(cplx<T>* __restrict__ y, const cplx<T>* __restrict__ x, const double* __restrict__ s,
                           int negate, size_t n) {
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    cplx<T> a = {(T)s[0], (T)s[1]};
    if (negate) a = -a;
    y[i] = y[i] + a * x[i];
}
template <typename T>
void launch_axpy_dev(cplx<T>* y, const cplx<T>* x, const double* s, int negate, size_t n, hipStream_t st) {
    k_axpy_dev<T><<<grid_for(n), kBlock, 0, st>>>(y, x, s, negate, n);
    SMG_LAUNCH_CHECK();
}

template <typename T>
__global__ void k_mr_update(cplx<T>* __restrict__ x, cplx<T>* __restrict__ r, const cplx<T>* __restrict__ Ar,
                            const double* __restrict__ s, size_t n) {
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const cplx<T> w = {(T)s[0], (T)s[1]};
    const cplx<T> ri = r[i];
    x[i] = x[i] + w * ri;
    r[i] = ri - w * Ar[i];
}
template <typename T>
void launch_mr_update(cplx<T>* x, cplx<T>* r, const cplx<T>* Ar, const double* s, size_t n, hipStream_t st) {
    k_mr_update<T><<<grid_for(n), kBlock, 0, st>>>(x, r, Ar, s, n);
    SMG_LAUNCH_CHECK();
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_update(cplx<T>* __restrict__ x, cplx<T>* __restrict__ r,
                                                      const cplx<T>* __restrict__ p,
                                                      const cplx<T>* __restrict__ Ap, const double* __restrict__ s,
                                                      const int* __restrict__ flag, size_t n, Fin fin) {
    double v[1] = {0.0};
    const bool live = (*flag == 0);
    if (live) {
        const cplx<T> a = {(T)s[0], (T)s[1]};
        for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
            x[i] = x[i] + a * p[i];
            const cplx<T> ri = r[i] - a * Ap[i];
            r[i] = ri;
            v[0] += (double)ri.re * (double)ri.re + (double)ri.im * (double)ri.im;
        }
    }
    reduce_finish<1>(v, fin, gridDim.x, blockIdx.x);
}
template <typename T>
void launch_cg_update(cplx<T>* x, cplx<T>* r, const cplx<T>* p, const cplx<T>* Ap, const double* s, const int* flag,
                      size_t n, const Fin& fin, hipStream_t st) {
    k_cg_update<T><<<red_blocks(n), kBlock, 0, st>>>(x, r, p, Ap, s, flag, n, fin);
    SMG_LAUNCH_CHECK();
}

template <typename T>
__global__ void k_xpay_dev(cplx<T>* __restrict__ p, const cplx<T>* __restrict__ z, const double* __restrict__ s,
                           const int* __restrict__ flag, size_t n) {
    if (flag && *flag) return;
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const cplx<T> b = {(T)s[0], (T)s[1]};
    p[i] = z[i] + b * p[i];
}
template <typename T>
void launch_xpay_dev(cplx<T>* p, const cplx<T>* z, const double* s, const int* flag, size_t n, hipStream_t st) {
    k_xpay_dev<T><<<grid_for(n), kBlock, 0, st>>>(p, z, s, flag, n);
    SMG_LAUNCH_CHECK();
}

template <typename To, typename From>
__global__ void k_convert(cplx<To>* __restrict__ d, const cplx<From>* __restrict__ s, size_t n) {
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) d[i] = cvt<To>(s[i]);
}
template <typename To, typename From>
void launch_convert(cplx<To>* dst, const cplx<From>* src, size_t n, hipStream_t st) {
    k_convert<To, From><<<grid_for(n), kBlock, 0, st>>>(dst, src, n);
    SMG_LAUNCH_CHECK();
}
__global__ void k_convert_half(ch16* __restrict__ d, const c64* __restrict__ s, size_t n) {
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) d[i] = ch16{__float2half((float)s[i].re), __float2half((float)s[i].im)};
}
void launch_convert_half(ch16* dst, const c64* src, size_t n, hipStream_t st) {
    k_convert_half<<<grid_for(n), kBlock, 0, st>>>(dst, src, n);
    SMG_LAUNCH_CHECK();
}

template <typename T>
__global__ void k_zero(cplx<T>* p, size_t n) {
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = cplx<T>{(T)0, (T)0};
}
template <typename T>
void launch_zero(cplx<T>* p, size_t n, hipStream_t st) {
    k_zero<T><<<grid_for(n), kBlock, 0, st>>>(p, n);
    SMG_LAUNCH_CHECK();
}

template <typename T>
__global__ void k_copy(cplx<T>* __restrict__ d, const cplx<T>* __restrict__ src, size_t n) {
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) d[i] = src[i];
}
template <typename T>
void launch_copy(cplx<T>* dst, const cplx<T>* src, size_t n, hipStream_t st) {
    k_copy<T><<<grid_for(n), kBlock, 0, st>>>(dst, src, n);
    SMG_LAUNCH_CHECK();
}

__global__ void k_accum_mixed(c64* __restrict__ x, const c32* __restrict__ e, size_t n) {
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) x[i] = x[i] + cvt<double>(e[i]);
}
void launch_accum_mixed(c64* x, const c32* e, size_t n, hipStream_t st) {
    k_accum_mixed<<<grid_for(n), kBlock, 0, st>>>(x, e, n);
    SMG_LAUNCH_CHECK();
}

template <typename T>
__global__ void k_g5(cplx<T>* v, size_t nsites) {
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < nsites) v[2 * i + 1] = -v[2 * i + 1];
}
template <typename T>
void launch_g5(cplx<T>* v, size_t nsites, hipStream_t st) {
    k_g5<T><<<grid_for(nsites), kBlock, 0, st>>>(v, nsites);
    SMG_LAUNCH_CHECK();
}

#define SMG_INST(T)                                                                                          \
    template void launch_dot<T>(const cplx<T>*, const cplx<T>*, size_t, const Fin&, hipStream_t);           \
    template void launch_norm2<T>(const cplx<T>*, size_t, const Fin&, hipStream_t);                         \
    template void launch_dots<T>(const cplx<T>*, const cplx<T>*, size_t, const Fin&, hipStream_t);          \
    template void launch_copy<T>(cplx<T>*, const cplx<T>*, size_t, hipStream_t);                            \
    template void launch_axpby<T>(cplx<T>*, const cplx<T>*, c64, c64, size_t, hipStream_t);                 \
    template void launch_axpy_dev<T>(cplx<T>*, const cplx<T>*, const double*, int, size_t, hipStream_t);    \
    template void launch_mr_update<T>(cplx<T>*, cplx<T>*, const cplx<T>*, const double*, size_t, hipStream_t); \
    template void launch_cg_update<T>(cplx<T>*, cplx<T>*, const cplx<T>*, const cplx<T>*, const double*,    \
                                      const int*, size_t, const Fin&, hipStream_t);                          \
    template void launch_xpay_dev<T>(cplx<T>*, const cplx<T>*, const double*, const int*, size_t, hipStream_t); \
    template void launch_zero<T>(cplx<T>*, size_t, hipStream_t);                                             \
    template void launch_g5<T>(cplx<T>*, size_t, hipStream_t);                                              \
    template void launch_multi_dot<T>(const VecPtrs<T>&, int, const cplx<T>*, size_t, int, const Fin&, hipStream_t); \
    template void launch_gcr_orth<T>(cplx<T>*, cplx<T>*, const VecPtrs<T>&, const VecPtrs<T>&, int,         \
                                     const double*, const double* const*, size_t, hipStream_t);
SMG_INST(double)
SMG_INST(float)
#undef SMG_INST
template void launch_convert<float, double>(cplx<float>*, const cplx<double>*, size_t, hipStream_t);
template void launch_convert<double, float>(cplx<double>*, const cplx<float>*, size_t, hipStream_t);
template void launch_convert<double, double>(cplx<double>*, const cplx<double>*, size_t, hipStream_t);
template void launch_convert<float, float>(cplx<float>*, const cplx<float>*, size_t, hipStream_t);

}  // namespace smg
