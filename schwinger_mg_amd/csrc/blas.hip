// blas.hip -- BLAS-1 and reductions (src/linalg.cpp:8-40) for gfx950.
//
// Reductions accumulate in FP64 whatever the storage precision: per-thread
// partial, wave64 shuffle tree, LDS across the 4 waves, one partial per
// workgroup, then a single-workgroup finalize in fixed order (deterministic).
// Scalars that feed later kernels (MR omega, CG alpha/beta, GCR alpha) stay in
// device memory, so a smoother or coarse solve issues no host synchronisation.
#include "smg_kernels.h"

namespace smg {

template <typename T>
__global__ __launch_bounds__(kBlock) void k_dot(const cplx<T>* __restrict__ a, const cplx<T>* __restrict__ b,
                                                size_t n, double* part) {
    double v[2] = {0.0, 0.0};
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        const cplx<T> p = cdot(a[i], b[i]);
        v[0] += (double)p.re;
        v[1] += (double)p.im;
    }
    block_sum<2>(v);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = v[0];
        part[2 * blockIdx.x + 1] = v[1];
    }
}
template <typename T>
int launch_dot_part(const cplx<T>* a, const cplx<T>* b, size_t n, double* part, hipStream_t st) {
    int g = red_blocks(n);
    k_dot<T><<<g, kBlock, 0, st>>>(a, b, n, part);
    SMG_LAUNCH_CHECK();
    return g;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_norm2(const cplx<T>* __restrict__ a, size_t n, double* part) {
    double v[1] = {0.0};
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        const cplx<T> p = a[i];
        v[0] += (double)p.re * (double)p.re + (double)p.im * (double)p.im;
    }
    block_sum<1>(v);
    if (threadIdx.x == 0) part[blockIdx.x] = v[0];
}
template <typename T>
int launch_norm2_part(const cplx<T>* a, size_t n, double* part, hipStream_t st) {
    int g = red_blocks(n);
    k_norm2<T><<<g, kBlock, 0, st>>>(a, n, part);
    SMG_LAUNCH_CHECK();
    return g;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_dots(const cplx<T>* __restrict__ a, const cplx<T>* __restrict__ b,
                                                 size_t n, double* part) {
    double v[3] = {0.0, 0.0, 0.0};
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        const cplx<T> x = a[i];
        const cplx<T> p = cdot(x, b[i]);
        v[0] += (double)p.re;
        v[1] += (double)p.im;
        v[2] += (double)x.re * (double)x.re + (double)x.im * (double)x.im;
    }
    block_sum<3>(v);
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x] = v[0];
        part[3 * blockIdx.x + 1] = v[1];
        part[3 * blockIdx.x + 2] = v[2];
    }
}
template <typename T>
int launch_dots_part(const cplx<T>* a, const cplx<T>* b, size_t n, double* part, hipStream_t st) {
    int g = red_blocks(n);
    k_dots<T><<<g, kBlock, 0, st>>>(a, b, n, part);
    SMG_LAUNCH_CHECK();
    return g;
}

// Finalize: sum partial rows, then one scalar program in thread 0 (see FinOp).
__global__ __launch_bounds__(kBlock) void k_finalize(const double* __restrict__ part, int nparts, int nc,
                                                     double* __restrict__ out, int op, double* aux) {
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = threadIdx.x; i < nparts; i += kBlock)
        for (int c = 0; c < nc; c++) v[c] += part[(size_t)i * nc + c];
    block_sum<4>(v);
    if (threadIdx.x != 0) return;
    int* flag = aux ? reinterpret_cast<int*>(aux + 8) : nullptr;
    switch (op) {
    case FIN_STORE:
        for (int c = 0; c < nc; c++) out[c] = v[c];
        break;
    case FIN_MR: {
        // src/smoother.cpp:34-37: omega = <Ar,r>/<Ar,Ar>; |<Ar,Ar>| < 1e-30 ends the sweeps,
        // which a zero step reproduces (x and r stay unchanged from then on)
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
        // src/coarse_op.h:520-523 (pAp = <p, Ap> = conj(<Ap, p>))
        if (*flag) break;
        const c64 pAp = {v[0], -v[1]};
        if (sqrt(pAp.re * pAp.re + pAp.im * pAp.im) < 1e-30) {
            *flag = 2;
            out[0] = 0.0;
            out[1] = 0.0;
            break;
        }
        const c64 a = cdiv(c64{aux[0], aux[1]}, pAp);
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
    case FIN_PCG_BETA: {
        const c64 b = cdiv(c64{v[0], v[1]}, c64{aux[0], aux[1]});
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
    }
}
void launch_finalize(const double* part, int nparts, int nc, double* out, int op, double* aux, hipStream_t st) {
    k_finalize<<<1, kBlock, 0, st>>>(part, nparts, nc, out, op, aux);
    SMG_LAUNCH_CHECK();
}

__global__ void k_finalize_many(const double* __restrict__ part, int nparts, int nc, double* __restrict__ out) {
    const int c = threadIdx.x;
    if (c >= nc) return;
    double a = 0.0;
    for (int i = 0; i < nparts; i++) a += part[(size_t)i * nc + c];
    out[c] = a;
}
void launch_finalize_many(const double* part, int nparts, int nc, double* out, hipStream_t st) {
    SMG_REQUIRE(nc <= 64, "finalize_many: at most 64 components");
    k_finalize_many<<<1, 64, 0, st>>>(part, nparts, nc, out);
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
                                                      size_t n, int swap, double* part) {
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
    block_sum<64>(acc);
    if (threadIdx.x == 0)
        for (int c = 0; c < 2 * nv; c++) part[(size_t)blockIdx.x * 2 * nv + c] = acc[c];
}
template <typename T>
int launch_multi_dot(const VecPtrs<T>& V, int nv, const cplx<T>* w, size_t n, int swap, double* part,
                     hipStream_t st) {
    SMG_REQUIRE(nv >= 1 && nv <= 32, "multi_dot: 1..32 vectors");
    int g = red_blocks(n);
    k_multi_dot<T><<<g, kBlock, 0, st>>>(V, nv, w, n, swap, part);
    SMG_LAUNCH_CHECK();
    return g;
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
                                                      const int* __restrict__ flag, size_t n, double* part) {
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
    block_sum<1>(v);
    if (threadIdx.x == 0) part[blockIdx.x] = v[0];
}
template <typename T>
int launch_cg_update(cplx<T>* x, cplx<T>* r, const cplx<T>* p, const cplx<T>* Ap, const double* s, const int* flag,
                     size_t n, double* part, hipStream_t st) {
    int g = red_blocks(n);
    k_cg_update<T><<<g, kBlock, 0, st>>>(x, r, p, Ap, s, flag, n, part);
    SMG_LAUNCH_CHECK();
    return g;
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

__global__ __launch_bounds__(kBlock) void k_resid_norm(c64* __restrict__ r, const c64* __restrict__ b,
                                                       const c64* __restrict__ Ax, size_t n, double* part) {
    double v[1] = {0.0};
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        const c64 d = b[i] - Ax[i];
        r[i] = d;
        v[0] += d.re * d.re + d.im * d.im;
    }
    block_sum<1>(v);
    if (threadIdx.x == 0) part[blockIdx.x] = v[0];
}
int launch_resid_norm(c64* r, const c64* b, const c64* Ax, size_t n, double* part, hipStream_t st) {
    int g = red_blocks(n);
    k_resid_norm<<<g, kBlock, 0, st>>>(r, b, Ax, n, part);
    SMG_LAUNCH_CHECK();
    return g;
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
    template int launch_dot_part<T>(const cplx<T>*, const cplx<T>*, size_t, double*, hipStream_t);          \
    template int launch_norm2_part<T>(const cplx<T>*, size_t, double*, hipStream_t);                        \
    template int launch_dots_part<T>(const cplx<T>*, const cplx<T>*, size_t, double*, hipStream_t);         \
    template void launch_axpby<T>(cplx<T>*, const cplx<T>*, c64, c64, size_t, hipStream_t);                 \
    template void launch_axpy_dev<T>(cplx<T>*, const cplx<T>*, const double*, int, size_t, hipStream_t);    \
    template void launch_mr_update<T>(cplx<T>*, cplx<T>*, const cplx<T>*, const double*, size_t, hipStream_t); \
    template int launch_cg_update<T>(cplx<T>*, cplx<T>*, const cplx<T>*, const cplx<T>*, const double*,     \
                                     const int*, size_t, double*, hipStream_t);                              \
    template void launch_xpay_dev<T>(cplx<T>*, const cplx<T>*, const double*, const int*, size_t, hipStream_t); \
    template void launch_zero<T>(cplx<T>*, size_t, hipStream_t);                                             \
    template void launch_g5<T>(cplx<T>*, size_t, hipStream_t);                                              \
    template int launch_multi_dot<T>(const VecPtrs<T>&, int, const cplx<T>*, size_t, int, double*, hipStream_t); \
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
