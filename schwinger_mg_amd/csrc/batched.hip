// batched.hip -- multi-right-hand-side kernels (layout: see smg_batched.h).
//
// Work-item t -> (i, j) = (t / N, t % N): the N lanes of one site or dof are
// consecutive, so field loads are N-wide coalesced runs and the per-site data
// (links, clover term, aggregate vectors, coarse links) is one broadcast load
// shared by those lanes.  Reductions are per RHS: lanes with equal j are summed
// by xor-shuffles over the lane bits above log2(N), then across the waves in LDS;
// the last workgroup sums the partial rows in index order and runs the scalar
// program of every RHS.
#include "smg_batched.h"

namespace smg {

constexpr int kRedGrid = 1024;   // workgroups of a reducing batched kernel

__device__ __forceinline__ int lg2(int N) { return 31 - __clz(N); }

// Per-RHS reduce-and-finish.  v[NC] holds this lane's partials for RHS j = lane % N.
template <int NC, int NMAX>
__device__ __forceinline__ void reduce_finish_b(double (&v)[NC], const FinB& f, int N, int nblk, int blk) {
    __shared__ double sm[kBlock / 64][NC * NMAX];
    __shared__ int last;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int off = N; off < 64; off <<= 1)
#pragma unroll
        for (int c = 0; c < NC; c++) v[c] += __shfl_xor(v[c], off, 64);
    if (lane < N)
#pragma unroll
        for (int c = 0; c < NC; c++) sm[wid][c * N + lane] = v[c];
    __syncthreads();
    const int tot = NC * N;
    for (int t = threadIdx.x; t < tot; t += kBlock)
        f.part[(size_t)blk * tot + t] = (sm[0][t] + sm[1][t]) + (sm[2][t] + sm[3][t]);
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) last = (atomicAdd(f.counter, 1u) == (unsigned)nblk - 1u);
    __syncthreads();
    if (!last) return;
    __threadfence();
    // component c = t % tot, rows g, g + G, ... for group g = t / tot; groups combined in order
    const int G = tot >= kBlock ? 1 : kBlock / tot;
    double* acc = &sm[0][0];   // G * tot <= kBlock*... fits: tot*G <= max(tot, kBlock) <= 4*NC*NMAX
    for (int t = threadIdx.x; t < tot * G; t += kBlock) {
        const int c = t % tot, g = t / tot;
        double a = 0.0;
        for (int i = g; i < nblk; i += G) a += f.part[(size_t)i * tot + c];
        acc[t] = a;
    }
    __syncthreads();
    if (threadIdx.x < N) {
        const int j = threadIdx.x;
        double w[NC];
#pragma unroll
        for (int c = 0; c < NC; c++) {
            double a = 0.0;
            for (int g = 0; g < G; g++) a += acc[g * tot + c * N + j];
            w[c] = a;
        }
        fin_program(f.op, w, NC, f.out + (size_t)j * f.ostride, f.aux ? f.aux + (size_t)j * f.astride : nullptr);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        *f.counter = 0u;
        __threadfence();
    }
}

static inline int b_grid(size_t items, int cap) {
    size_t g = (items + kBlock - 1) / kBlock;
    if (g < 1) g = 1;
    return (int)(g > (size_t)cap ? cap : g);
}

// ----------------------------------------------------------------- Wilson-clover
template <typename T>
__device__ __forceinline__ void ld_b(const cplx<T>* __restrict__ p, size_t site, int N, int j, int g5, cplx<T>& a,
                                     cplx<T>& b) {
    a = p[(2 * site) * N + j];
    b = p[(2 * site + 1) * N + j];
    if (g5) b = -b;
}

template <typename T>
__device__ __forceinline__ void hop_b(T r, cplx<T> u0, cplx<T> u0b, cplx<T> u1, cplx<T> u1b, cplx<T> xf0, cplx<T> xf1,
                                      cplx<T> xb0, cplx<T> xb1, cplx<T> yf0, cplx<T> yf1, cplx<T> yb0, cplx<T> yb1,
                                      cplx<T>& h0, cplx<T>& h1) {
    const cplx<T> up0 = u0 * xf0, up1 = u0 * xf1;
    const cplx<T> ud = conj(u0b);
    const cplx<T> q0 = ud * xb0, q1 = ud * xb1;
    const cplx<T> vp0 = u1 * yf0, vp1 = u1 * yf1;
    const cplx<T> wd = conj(u1b);
    const cplx<T> w0 = wd * yb0, w1 = wd * yb1;
    const T hr = (T)0.5 * r;
    const cplx<T> s0 = (up0 + q0) + (vp0 + w0), s1 = (up1 + q1) + (vp1 + w1);
    const cplx<T> t0 = (q1 - up1) + mul_i(vp1 - w1), t1 = (q0 - up0) - mul_i(vp0 - w0);
    h0 = -(hr * s0) - (T)0.5 * t0;
    h1 = -(hr * s1) - (T)0.5 * t1;
}

template <typename T, int EPI>
__global__ __launch_bounds__(kBlock) void k_b_dslash(DiracView<T> D, cplx<T>* __restrict__ dst,
                                                     const cplx<T>* __restrict__ src, T mu, int g5in, int g5out,
                                                     const cplx<T>* __restrict__ bvec,
                                                     const cplx<T>* __restrict__ aux, FinB fin, int N) {
    const int L = D.L, V = L * L, lN = lg2(N);
    const size_t items = (size_t)V << lN;
    double acc[3] = {0.0, 0.0, 0.0};
    for (size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x; t < items; t += (size_t)gridDim.x * kBlock) {
        const int s = (int)(t >> lN), j = (int)(t & (N - 1));
        const int x = s % L, y = s / L;
        const int sxp = (x + 1 == L) ? s - x : s + 1;
        const int sxm = (x == 0) ? s + L - 1 : s - 1;
        const int syp = (y + 1 == L) ? x : s + L;
        const int sym = (y == 0) ? s + V - L : s - L;
        cplx<T> s0, s1, xf0, xf1, xb0, xb1, yf0, yf1, yb0, yb1;
        ld_b(src, s, N, j, g5in, s0, s1);
        ld_b(src, sxp, N, j, g5in, xf0, xf1);
        ld_b(src, sxm, N, j, g5in, xb0, xb1);
        ld_b(src, syp, N, j, g5in, yf0, yf1);
        ld_b(src, sym, N, j, g5in, yb0, yb1);
        const cplx<T>* U0 = D.U;
        const cplx<T>* U1 = D.U + V;
        cplx<T> h0, h1;
        hop_b(D.r, U0[s], U0[sxm], U1[s], U1[sym], xf0, xf1, xb0, xb1, yf0, yf1, yb0, yb1, h0, h1);
        const T C = D.C ? D.C[s] : (T)0;
        cplx<T> o0 = mk<T>(D.diag + C, mu) * s0 + h0, o1 = mk<T>(D.diag - C, -mu) * s1 + h1;
        if (g5out) o1 = -o1;
        const size_t i0 = (2 * (size_t)s) * N + j, i1 = i0 + N;
        if (EPI == EPI_RESID) {
            o0 = bvec[i0] - o0;
            o1 = bvec[i1] - o1;
        }
        dst[i0] = o0;
        dst[i1] = o1;
        if (EPI == EPI_DOTS) {
            const cplx<T> d0 = cdot(o0, aux[i0]), d1 = cdot(o1, aux[i1]);
            acc[0] += (double)d0.re + (double)d1.re;
            acc[1] += (double)d0.im + (double)d1.im;
            acc[2] += (double)norm2(o0) + (double)norm2(o1);
        }
    }
    if (EPI == EPI_DOTS) reduce_finish_b<3, 64>(acc, fin, N, gridDim.x, blockIdx.x);
}

template <typename T>
void launch_b_dslash(const DiracView<T>& D, cplx<T>* dst, const cplx<T>* src, T mu, int g5in, int g5out, Epi epi,
                     const cplx<T>* b, const cplx<T>* aux, const FinB& fin, int N, hipStream_t st) {
    const size_t items = (size_t)D.L * D.L * N;
    if (epi == EPI_NONE)
        k_b_dslash<T, EPI_NONE><<<b_grid(items, 1 << 16), kBlock, 0, st>>>(D, dst, src, mu, g5in, g5out, b, aux, fin, N);
    else if (epi == EPI_RESID)
        k_b_dslash<T, EPI_RESID><<<b_grid(items, 1 << 16), kBlock, 0, st>>>(D, dst, src, mu, g5in, g5out, b, aux, fin, N);
    else
        k_b_dslash<T, EPI_DOTS><<<b_grid(items, kRedGrid), kBlock, 0, st>>>(D, dst, src, mu, g5in, g5out, b, aux, fin, N);
    SMG_LAUNCH_CHECK();
}

// ----------------------------------------------------------------- coarse operator
template <typename T, typename TL, int EPI>
__global__ __launch_bounds__(kBlock) void k_b_coarse(CoarseView<TL> C, cplx<T>* __restrict__ dst,
                                                     const cplx<T>* __restrict__ src,
                                                     const cplx<T>* __restrict__ bvec,
                                                     const cplx<T>* __restrict__ aux, FinB fin, int N) {
    const int k = C.k, lN = lg2(N);
    const size_t items = ((size_t)C.nx * C.ny * k) << lN;
    double acc[3] = {0.0, 0.0, 0.0};
    for (size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x; t < items; t += (size_t)gridDim.x * kBlock) {
        const size_t row = t >> lN;   // b*k + ki
        const int j = (int)(t & (N - 1));
        const int b = (int)(row / k), ki = (int)(row % k);
        const int bx = b % C.nx, by = b / C.nx;
        const int xm = (bx == 0 ? C.nx - 1 : bx - 1), xp = (bx + 1 == C.nx ? 0 : bx + 1);
        const int ym = (by == 0 ? C.ny - 1 : by - 1), yp = (by + 1 == C.ny ? 0 : by + 1);
        const int nbr[9] = {b, xp + by * C.nx, xm + by * C.nx, bx + yp * C.nx, bx + ym * C.nx,
                            xp + yp * C.nx, xp + ym * C.nx, xm + yp * C.nx, xm + ym * C.nx};
        cplx<T> out = {(T)0, (T)0};
        for (int d = 0; d < 9; d++) {
            const TL* lrow = C.link + (((size_t)b * 9 + d) * k + ki) * k;
            const cplx<T>* sv = src + ((size_t)nbr[d] * k) * N + j;
            cplx<T> sum = {(T)0, (T)0};
            for (int kj = 0; kj < k; kj++) sum = sum + cvt<T>(load_c(lrow + kj)) * sv[(size_t)kj * N];
            out = out + sum;
        }
        if (EPI == EPI_RESID) out = bvec[t] - out;
        dst[t] = out;
        if (EPI == EPI_DOTS) {
            const cplx<T> p = cdot(out, aux[t]);
            acc[0] += p.re;
            acc[1] += p.im;
            acc[2] += norm2(out);
        }
    }
    if (EPI == EPI_DOTS) reduce_finish_b<3, 64>(acc, fin, N, gridDim.x, blockIdx.x);
}
template <typename T, typename TL>
void launch_b_coarse_apply(const CoarseView<TL>& C, cplx<T>* dst, const cplx<T>* src, Epi epi, const cplx<T>* b,
                           const cplx<T>* aux, const FinB& fin, int N, hipStream_t st) {
    const size_t items = (size_t)C.nx * C.ny * C.k * N;
    if (epi == EPI_NONE)
        k_b_coarse<T, TL, EPI_NONE><<<b_grid(items, 1 << 16), kBlock, 0, st>>>(C, dst, src, b, aux, fin, N);
    else if (epi == EPI_RESID)
        k_b_coarse<T, TL, EPI_RESID><<<b_grid(items, 1 << 16), kBlock, 0, st>>>(C, dst, src, b, aux, fin, N);
    else
        k_b_coarse<T, TL, EPI_DOTS><<<b_grid(items, kRedGrid), kBlock, 0, st>>>(C, dst, src, b, aux, fin, N);
    SMG_LAUNCH_CHECK();
}

// ----------------------------------------------------------------- transfers
template <typename T>
__global__ __launch_bounds__(kBlock) void k_b_prolong(AggView A, const cplx<T>* __restrict__ bv,
                                                      const cplx<T>* __restrict__ coarse, cplx<T>* __restrict__ fine,
                                                      int accumulate, int N) {
    const int lN = lg2(N);
    const size_t items = ((size_t)A.nx * A.ny * A.nd) << lN;
    const int bsize = A.bx * A.by * A.nd;
    for (size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x; t < items; t += (size_t)gridDim.x * kBlock) {
        const size_t f = t >> lN;
        const int j = (int)(t & (N - 1));
        const int s = (int)(f / A.nd), c = (int)(f % A.nd);
        const int x = s % A.nx, y = s / A.nx;
        const int b = x / A.bx + A.nbx * (y / A.by);
        const int loc = (x % A.bx) + A.bx * (y % A.by);
        const cplx<T>* v = bv + (size_t)b * A.k * bsize + loc * A.nd + c;
        const cplx<T>* cc = coarse + ((size_t)b * A.k) * N + j;
        cplx<T> acc = {(T)0, (T)0};
        for (int kk = 0; kk < A.k; kk++) acc = acc + cc[(size_t)kk * N] * v[(size_t)kk * bsize];
        fine[t] = accumulate ? fine[t] + acc : acc;
    }
}
template <typename T>
void launch_b_prolong(const AggView& A, const cplx<T>* bv, const cplx<T>* coarse, cplx<T>* fine, int accumulate,
                      int N, hipStream_t st) {
    const size_t items = (size_t)A.nx * A.ny * A.nd * N;
    k_b_prolong<T><<<b_grid(items, 1 << 16), kBlock, 0, st>>>(A, bv, coarse, fine, accumulate, N);
    SMG_LAUNCH_CHECK();
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_b_restrict(AggView A, const cplx<T>* __restrict__ bv,
                                                       const cplx<T>* __restrict__ fine,
                                                       cplx<T>* __restrict__ coarse, int N) {
    const int lN = lg2(N);
    const size_t items = ((size_t)A.nbx * A.nby * A.k) << lN;
    const int bsites = A.bx * A.by, bsize = bsites * A.nd;
    for (size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x; t < items; t += (size_t)gridDim.x * kBlock) {
        const size_t row = t >> lN;   // b*k + kk
        const int j = (int)(t & (N - 1));
        const int b = (int)(row / A.k), kk = (int)(row % A.k);
        const int bxi = b % A.nbx, byi = b / A.nbx;
        const cplx<T>* v = bv + ((size_t)b * A.k + kk) * bsize;
        cplx<T> acc = {(T)0, (T)0};
        for (int loc = 0; loc < bsites; loc++) {
            const int s = (bxi * A.bx + loc % A.bx) + A.nx * (byi * A.by + loc / A.bx);
            const cplx<T>* fp = fine + ((size_t)s * A.nd) * N + j;
            cplx<T> a = cdot(v[loc * A.nd], fp[0]);
            for (int c = 1; c < A.nd; c++) a = a + cdot(v[loc * A.nd + c], fp[(size_t)c * N]);
            acc = acc + a;
        }
        coarse[t] = acc;
    }
}
template <typename T>
void launch_b_restrict(const AggView& A, const cplx<T>* bv, const cplx<T>* fine, cplx<T>* coarse, int N,
                       hipStream_t st) {
    const size_t items = (size_t)A.nbx * A.nby * A.k * N;
    k_b_restrict<T><<<b_grid(items, 1 << 16), kBlock, 0, st>>>(A, bv, fine, coarse, N);
    SMG_LAUNCH_CHECK();
}

// Y = M X, M row-major n x n (the dense coarsest inverse); one work-item per (row, j)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_b_dense(const cplx<T>* __restrict__ M, const cplx<T>* __restrict__ X,
                                                    cplx<T>* __restrict__ Y, size_t n, int N) {
    const int lN = lg2(N);
    const size_t items = n << lN;
    for (size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x; t < items; t += (size_t)gridDim.x * kBlock) {
        const size_t row = t >> lN;
        const int j = (int)(t & (N - 1));
        const cplx<T>* mr = M + row * n;
        cplx<T> acc = {(T)0, (T)0};
        for (size_t c = 0; c < n; c++) acc = acc + mr[c] * X[(c << lN) + j];
        Y[t] = acc;
    }
}
template <typename T>
void launch_b_dense_apply(const cplx<T>* M, const cplx<T>* x, cplx<T>* y, size_t n, int N, hipStream_t st) {
    k_b_dense<T><<<b_grid(n * N, 1 << 16), kBlock, 0, st>>>(M, x, y, n, N);
    SMG_LAUNCH_CHECK();
}

// ----------------------------------------------------------------- BLAS-1
template <typename T>
__global__ __launch_bounds__(kBlock) void k_b_norm2(const cplx<T>* __restrict__ a, size_t items, FinB fin, int N) {
    double v[1] = {0.0};
    for (size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x; t < items; t += (size_t)gridDim.x * kBlock)
        v[0] += (double)norm2(a[t]);
    reduce_finish_b<1, 64>(v, fin, N, gridDim.x, blockIdx.x);
}
template <typename T>
void launch_b_norm2(const cplx<T>* a, size_t n, const FinB& fin, int N, hipStream_t st) {
    k_b_norm2<T><<<b_grid(n * N, kRedGrid), kBlock, 0, st>>>(a, n * N, fin, N);
    SMG_LAUNCH_CHECK();
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_b_dots(const cplx<T>* __restrict__ a, const cplx<T>* __restrict__ b,
                                                   size_t items, FinB fin, int N) {
    double v[3] = {0.0, 0.0, 0.0};
    for (size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x; t < items; t += (size_t)gridDim.x * kBlock) {
        const cplx<T> x = a[t];
        const cplx<T> p = cdot(x, b[t]);
        v[0] += (double)p.re;
        v[1] += (double)p.im;
        v[2] += (double)norm2(x);
    }
    reduce_finish_b<3, 64>(v, fin, N, gridDim.x, blockIdx.x);
}
template <typename T>
void launch_b_dots(const cplx<T>* a, const cplx<T>* b, size_t n, const FinB& fin, int N, hipStream_t st) {
    k_b_dots<T><<<b_grid(n * N, kRedGrid), kBlock, 0, st>>>(a, b, n * N, fin, N);
    SMG_LAUNCH_CHECK();
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_b_multi_dot(VecPtrs<T> V, int nv, const cplx<T>* __restrict__ w,
                                                        size_t items, FinB fin, int N) {
    double v[32];
#pragma unroll
    for (int c = 0; c < 32; c++) v[c] = 0.0;
    for (size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x; t < items; t += (size_t)gridDim.x * kBlock) {
        const cplx<T> we = w[t];
#pragma unroll
        for (int i = 0; i < 16; i++)
            if (i < nv) {
                const cplx<T> p = cdot(V.v[i][t], we);
                v[2 * i] += (double)p.re;
                v[2 * i + 1] += (double)p.im;
            }
    }
    reduce_finish_b<32, 32>(v, fin, N, gridDim.x, blockIdx.x);
}
template <typename T>
void launch_b_multi_dot(const VecPtrs<T>& V, int nv, const cplx<T>* w, size_t n, const FinB& fin, int N,
                        hipStream_t st) {
    SMG_REQUIRE(nv >= 1 && nv <= 16 && N <= 32, "batched multi_dot: 1..16 vectors, N <= 32");
    k_b_multi_dot<T><<<b_grid(n * N, 256), kBlock, 0, st>>>(V, nv, w, n * N, fin, N);
    SMG_LAUNCH_CHECK();
}

template <typename T>
struct ApapB {
    const double* a[16];
};
template <typename T>
__global__ __launch_bounds__(kBlock) void k_b_gcr_orth(cplx<T>* __restrict__ p, cplx<T>* __restrict__ Ap,
                                                       VecPtrs<T> P, VecPtrs<T> AP, int nv,
                                                       const double* __restrict__ dots, ApapB<T> apap, size_t items,
                                                       int N) {
    __shared__ cplx<T> beta[16 * 64];
    for (int q = threadIdx.x; q < nv * N; q += kBlock) {
        const int i = q / N, j = q % N;
        const double d = apap.a[i][j];
        const double* dd = dots + (size_t)j * 32 + 2 * i;
        beta[i * N + j] = d > 0.0 ? cplx<T>{(T)(dd[0] / d), (T)(dd[1] / d)} : cplx<T>{(T)0, (T)0};
    }
    __syncthreads();
    for (size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x; t < items; t += (size_t)gridDim.x * kBlock) {
        const int j = (int)(t & (N - 1));
        cplx<T> pe = p[t], ae = Ap[t];
        for (int i = 0; i < nv; i++) {
            const cplx<T> bb = beta[i * N + j];
            pe = pe - bb * P.v[i][t];
            ae = ae - bb * AP.v[i][t];
        }
        p[t] = pe;
        Ap[t] = ae;
    }
}
template <typename T>
void launch_b_gcr_orth(cplx<T>* p, cplx<T>* Ap, const VecPtrs<T>& P, const VecPtrs<T>& AP, int nv,
                       const double* dots, const double* const* apap, size_t n, int N, hipStream_t st) {
    ApapB<T> a;
    for (int i = 0; i < nv; i++) a.a[i] = apap[i];
    k_b_gcr_orth<T><<<b_grid(n * N, 1 << 14), kBlock, 0, st>>>(p, Ap, P, AP, nv, dots, a, n * N, N);
    SMG_LAUNCH_CHECK();
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_b_mr_update(cplx<T>* __restrict__ x, cplx<T>* __restrict__ r,
                                                        const cplx<T>* __restrict__ Ar, const double* __restrict__ s,
                                                        size_t items, int N) {
    for (size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x; t < items; t += (size_t)gridDim.x * kBlock) {
        const int j = (int)(t & (N - 1));
        const cplx<T> w = {(T)s[2 * j], (T)s[2 * j + 1]};
        const cplx<T> ri = r[t];
        x[t] = x[t] + w * ri;
        r[t] = ri - w * Ar[t];
    }
}
template <typename T>
void launch_b_mr_update(cplx<T>* x, cplx<T>* r, const cplx<T>* Ar, const double* s, size_t n, int N, hipStream_t st) {
    k_b_mr_update<T><<<b_grid(n * N, 1 << 16), kBlock, 0, st>>>(x, r, Ar, s, n * N, N);
    SMG_LAUNCH_CHECK();
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_b_update(cplx<T>* __restrict__ x, cplx<T>* __restrict__ r,
                                                     const cplx<T>* __restrict__ p, const cplx<T>* __restrict__ Ap,
                                                     const double* __restrict__ s, size_t items, FinB fin, int N) {
    double v[1] = {0.0};
    for (size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x; t < items; t += (size_t)gridDim.x * kBlock) {
        const int j = (int)(t & (N - 1));
        const cplx<T> a = {(T)s[2 * j], (T)s[2 * j + 1]};
        x[t] = x[t] + a * p[t];
        const cplx<T> ri = r[t] - a * Ap[t];
        r[t] = ri;
        v[0] += (double)norm2(ri);
    }
    reduce_finish_b<1, 64>(v, fin, N, gridDim.x, blockIdx.x);
}
template <typename T>
void launch_b_update(cplx<T>* x, cplx<T>* r, const cplx<T>* p, const cplx<T>* Ap, const double* s, size_t n,
                     const FinB& fin, int N, hipStream_t st) {
    k_b_update<T><<<b_grid(n * N, kRedGrid), kBlock, 0, st>>>(x, r, p, Ap, s, n * N, fin, N);
    SMG_LAUNCH_CHECK();
}

template <typename T>
__global__ void k_b_xpay(cplx<T>* __restrict__ p, const cplx<T>* __restrict__ z, const double* __restrict__ s,
                         size_t items, int N) {
    for (size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x; t < items; t += (size_t)gridDim.x * kBlock) {
        const int j = (int)(t & (N - 1));
        const cplx<T> b = {(T)s[2 * j], (T)s[2 * j + 1]};
        p[t] = z[t] + b * p[t];
    }
}
template <typename T>
void launch_b_xpay(cplx<T>* p, const cplx<T>* z, const double* s, size_t n, int N, hipStream_t st) {
    k_b_xpay<T><<<b_grid(n * N, 1 << 16), kBlock, 0, st>>>(p, z, s, n * N, N);
    SMG_LAUNCH_CHECK();
}

// per-RHS CG program state: aux_j[5] = tol^2, flag_j = 0
__global__ void k_b_cg_init(double* aux, int astride, int N, double tol2) {
    const int j = threadIdx.x;
    if (j < N) {
        aux[(size_t)j * astride + 5] = tol2;
        *reinterpret_cast<int*>(aux + (size_t)j * astride + 8) = 0;
    }
}
void launch_b_cg_init(double* aux, int astride, int N, double tol2, hipStream_t st) {
    k_b_cg_init<<<1, 64, 0, st>>>(aux, astride, N, tol2);
    SMG_LAUNCH_CHECK();
}

template <typename T>
__global__ void k_b_axpby_real(cplx<T>* __restrict__ y, const cplx<T>* __restrict__ x, T a, T c, size_t items) {
    for (size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x; t < items; t += (size_t)gridDim.x * kBlock)
        y[t] = a * x[t] + c * y[t];
}
template <typename T>
void launch_b_axpby_real(cplx<T>* y, const cplx<T>* x, double a, double c, size_t n, hipStream_t st) {
    k_b_axpby_real<T><<<b_grid(n, 1 << 16), kBlock, 0, st>>>(y, x, (T)a, (T)c, n);
    SMG_LAUNCH_CHECK();
}

// separate [j][i] <-> batched [i][j] through an LDS tile of 64 i x N j
template <typename To, typename From, int TO_BATCHED>
__global__ __launch_bounds__(kBlock) void k_b_transpose(cplx<To>* __restrict__ out, const cplx<From>* __restrict__ in,
                                                        size_t n, int N) {
    __shared__ cplx<To> tile[64 * 65];
    const size_t i0 = (size_t)blockIdx.x * 64;
    const int rows = (int)((n - i0) < 64 ? (n - i0) : 64);
    if (TO_BATCHED) {
        for (int q = threadIdx.x; q < 64 * N; q += kBlock) {
            const int j = q / 64, i = q % 64;
            if (i < rows) tile[i * 65 + j] = cvt<To>(in[(size_t)j * n + i0 + i]);
        }
        __syncthreads();
        for (int q = threadIdx.x; q < 64 * N; q += kBlock) {
            const int i = q / N, j = q % N;
            if (i < rows) out[(i0 + i) * N + j] = tile[i * 65 + j];
        }
    } else {
        for (int q = threadIdx.x; q < 64 * N; q += kBlock) {
            const int i = q / N, j = q % N;
            if (i < rows) tile[i * 65 + j] = cvt<To>(in[(i0 + i) * N + j]);
        }
        __syncthreads();
        for (int q = threadIdx.x; q < 64 * N; q += kBlock) {
            const int j = q / 64, i = q % 64;
            if (i < rows) out[(size_t)j * n + i0 + i] = tile[i * 65 + j];
        }
    }
}
template <typename To, typename From>
void launch_b_pack(cplx<To>* batched, const cplx<From>* sep, size_t n, int N, hipStream_t st) {
    SMG_REQUIRE(N >= 1 && N <= 64 && (N & (N - 1)) == 0, "N must be a power of two <= 64");
    k_b_transpose<To, From, 1><<<(unsigned)((n + 63) / 64), kBlock, 0, st>>>(batched, sep, n, N);
    SMG_LAUNCH_CHECK();
}
template <typename To, typename From>
void launch_b_unpack(cplx<To>* sep, const cplx<From>* batched, size_t n, int N, hipStream_t st) {
    SMG_REQUIRE(N >= 1 && N <= 64 && (N & (N - 1)) == 0, "N must be a power of two <= 64");
    k_b_transpose<To, From, 0><<<(unsigned)((n + 63) / 64), kBlock, 0, st>>>(sep, batched, n, N);
    SMG_LAUNCH_CHECK();
}

__global__ void k_b_accum_mixed(c64* __restrict__ x, const c32* __restrict__ e, size_t items) {
    for (size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x; t < items; t += (size_t)gridDim.x * kBlock)
        x[t] = x[t] + cvt<double>(e[t]);
}
void launch_b_accum_mixed(c64* x, const c32* e, size_t n, int N, hipStream_t st) {
    k_b_accum_mixed<<<b_grid(n * N, 1 << 16), kBlock, 0, st>>>(x, e, n * N);
    SMG_LAUNCH_CHECK();
}

#define SMG_BINST(T)                                                                                             \
    template void launch_b_dslash<T>(const DiracView<T>&, cplx<T>*, const cplx<T>*, T, int, int, Epi,           \
                                     const cplx<T>*, const cplx<T>*, const FinB&, int, hipStream_t);            \
    template void launch_b_prolong<T>(const AggView&, const cplx<T>*, const cplx<T>*, cplx<T>*, int, int,        \
                                      hipStream_t);                                                              \
    template void launch_b_restrict<T>(const AggView&, const cplx<T>*, const cplx<T>*, cplx<T>*, int, hipStream_t); \
    template void launch_b_dense_apply<T>(const cplx<T>*, const cplx<T>*, cplx<T>*, size_t, int, hipStream_t);   \
    template void launch_b_norm2<T>(const cplx<T>*, size_t, const FinB&, int, hipStream_t);                     \
    template void launch_b_dots<T>(const cplx<T>*, const cplx<T>*, size_t, const FinB&, int, hipStream_t);      \
    template void launch_b_multi_dot<T>(const VecPtrs<T>&, int, const cplx<T>*, size_t, const FinB&, int,        \
                                        hipStream_t);                                                            \
    template void launch_b_gcr_orth<T>(cplx<T>*, cplx<T>*, const VecPtrs<T>&, const VecPtrs<T>&, int,            \
                                       const double*, const double* const*, size_t, int, hipStream_t);           \
    template void launch_b_mr_update<T>(cplx<T>*, cplx<T>*, const cplx<T>*, const double*, size_t, int,          \
                                        hipStream_t);                                                            \
    template void launch_b_update<T>(cplx<T>*, cplx<T>*, const cplx<T>*, const cplx<T>*, const double*, size_t,  \
                                     const FinB&, int, hipStream_t);                                             \
    template void launch_b_axpby_real<T>(cplx<T>*, const cplx<T>*, double, double, size_t, hipStream_t);    \
    template void launch_b_xpay<T>(cplx<T>*, const cplx<T>*, const double*, size_t, int, hipStream_t);
SMG_BINST(double)
SMG_BINST(float)
#undef SMG_BINST
template void launch_b_coarse_apply<double, c64>(const CoarseView<c64>&, c64*, const c64*, Epi, const c64*,
                                                 const c64*, const FinB&, int, hipStream_t);
template void launch_b_coarse_apply<float, c32>(const CoarseView<c32>&, c32*, const c32*, Epi, const c32*,
                                                const c32*, const FinB&, int, hipStream_t);
template void launch_b_coarse_apply<float, ch16>(const CoarseView<ch16>&, c32*, const c32*, Epi, const c32*,
                                                 const c32*, const FinB&, int, hipStream_t);
template void launch_b_pack<double, double>(c64*, const c64*, size_t, int, hipStream_t);
template void launch_b_pack<float, double>(c32*, const c64*, size_t, int, hipStream_t);
template void launch_b_unpack<double, double>(c64*, const c64*, size_t, int, hipStream_t);
template void launch_b_unpack<double, float>(c64*, const c32*, size_t, int, hipStream_t);

}  // namespace smg
