// mg.hip -- aggregation transfer operators and the 9-point coarse operator.
//
// Aggregates are geometric bx x by blocks of a (coarse or fine) lattice with nd
// dof per site; block vectors are stored per aggregate, bv[b][j][loc*nd + c]
// (src/prolongator.h:21-23), so a prolong reads each aggregate's vectors as one
// contiguous run.  The coarse operator is the 9-point stencil of
// src/coarse_op.h:154-297: link[b][d][ki*k + kj].
#include "smg_kernels.h"

namespace smg {

__device__ __forceinline__ void agg_site(const AggView& A, int b, int loc, int& s) {
    const int bxi = b % A.nbx, byi = b / A.nbx;
    const int lx = loc % A.bx, ly = loc / A.bx;
    s = (bxi * A.bx + lx) + A.nx * (byi * A.by + ly);
}

// src/prolongator.h:63-74: copy vector kv into the block storage
__global__ void k_block_gather(AggView A, c64* __restrict__ bv, const c64* __restrict__ vec, int kv) {
    const int bsites = A.bx * A.by, bsize = bsites * A.nd;
    const int nb = A.nbx * A.nby;
    const size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= (size_t)nb * bsize) return;
    const int b = (int)(t / bsize), e = (int)(t % bsize);
    const int loc = e / A.nd, c = e % A.nd;
    int s;
    agg_site(A, b, loc, s);
    bv[((size_t)b * A.k + kv) * bsize + e] = vec[(size_t)s * A.nd + c];
}
void launch_block_gather(const AggView& A, c64* bv, const c64* vec, int kv, hipStream_t st) {
    const size_t n = (size_t)A.nbx * A.nby * A.bx * A.by * A.nd;
    k_block_gather<<<grid_for(n), kBlock, 0, st>>>(A, bv, vec, kv);
    SMG_LAUNCH_CHECK();
}

// src/prolongator.h:111-130: modified Gram-Schmidt per aggregate, one wave64 per
// aggregate, lanes striding over the block vector.
__global__ __launch_bounds__(64) void k_block_orth(AggView A, c64* __restrict__ bv) {
    const int bsize = A.bx * A.by * A.nd;
    const int b = blockIdx.x;
    const int lane = threadIdx.x;
    c64* base = bv + (size_t)b * A.k * bsize;
    for (int kv = 0; kv < A.k; kv++) {
        c64* v = base + (size_t)kv * bsize;
        for (int j = 0; j < kv; j++) {
            const c64* w = base + (size_t)j * bsize;
            double pr = 0.0, pi = 0.0;
            for (int i = lane; i < bsize; i += 64) {
                const c64 p = cdot(w[i], v[i]);
                pr += p.re;
                pi += p.im;
            }
            pr = wave_sum(pr);
            pi = wave_sum(pi);
            const c64 proj = {pr, pi};
            for (int i = lane; i < bsize; i += 64) v[i] = v[i] - proj * w[i];
        }
        double nn = 0.0;
        for (int i = lane; i < bsize; i += 64) nn += norm2(v[i]);
        nn = sqrt(wave_sum(nn));
        if (nn > 1e-14) {
            const double inv = 1.0 / nn;
            for (int i = lane; i < bsize; i += 64) v[i] = c64{v[i].re * inv, v[i].im * inv};
        }
    }
}
void launch_block_orthonormalize(const AggView& A, c64* bv, hipStream_t st) {
    k_block_orth<<<A.nbx * A.nby, 64, 0, st>>>(A, bv);
    SMG_LAUNCH_CHECK();
}

// src/prolongator.h:77-90
template <typename T>
__global__ __launch_bounds__(kBlock) void k_prolong(AggView A, const cplx<T>* __restrict__ bv,
                                                    const cplx<T>* __restrict__ coarse, cplx<T>* __restrict__ fine,
                                                    int accumulate) {
    const size_t f = (size_t)blockIdx.x * kBlock + threadIdx.x;
    const size_t nf = (size_t)A.nx * A.ny * A.nd;
    if (f >= nf) return;
    const int s = (int)(f / A.nd), c = (int)(f % A.nd);
    const int x = s % A.nx, y = s / A.nx;
    const int b = x / A.bx + A.nbx * (y / A.by);
    const int loc = (x % A.bx) + A.bx * (y % A.by);
    const int bsize = A.bx * A.by * A.nd;
    const cplx<T>* v = bv + (size_t)b * A.k * bsize + loc * A.nd + c;
    const cplx<T>* cc = coarse + (size_t)b * A.k;
    cplx<T> acc = {(T)0, (T)0};
    for (int j = 0; j < A.k; j++) acc = acc + cc[j] * v[(size_t)j * bsize];
    fine[f] = accumulate ? fine[f] + acc : acc;
}
template <typename T>
void launch_prolong(const AggView& A, const cplx<T>* bv, const cplx<T>* coarse, cplx<T>* fine, int accumulate,
                    hipStream_t st) {
    const size_t nf = (size_t)A.nx * A.ny * A.nd;
    k_prolong<T><<<grid_for(nf), kBlock, 0, st>>>(A, bv, coarse, fine, accumulate);
    SMG_LAUNCH_CHECK();
}

// src/prolongator.h:94-108: one work-item per coarse dof (b, j), summing the
// aggregate's sites in ascending site order.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_restrict(AggView A, const cplx<T>* __restrict__ bv,
                                                     const cplx<T>* __restrict__ fine,
                                                     cplx<T>* __restrict__ coarse) {
    const int nb = A.nbx * A.nby;
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= nb * A.k) return;
    const int b = t / A.k, j = t % A.k;
    const int bsites = A.bx * A.by, bsize = bsites * A.nd;
    const cplx<T>* v = bv + ((size_t)b * A.k + j) * bsize;
    cplx<T> acc = {(T)0, (T)0};
    for (int loc = 0; loc < bsites; loc++) {
        int s;
        agg_site(A, b, loc, s);
        const cplx<T>* fp = fine + (size_t)s * A.nd;
        cplx<T> a = cdot(v[loc * A.nd], fp[0]);
        for (int c = 1; c < A.nd; c++) a = a + cdot(v[loc * A.nd + c], fp[c]);
        acc = acc + a;
    }
    coarse[t] = acc;
}
template <typename T>
void launch_restrict(const AggView& A, const cplx<T>* bv, const cplx<T>* fine, cplx<T>* coarse, hipStream_t st) {
    const int n = A.nbx * A.nby * A.k;
    k_restrict<T><<<grid_for(n), kBlock, 0, st>>>(A, bv, fine, coarse);
    SMG_LAUNCH_CHECK();
}

// src/coarse_op.h:205-222
__device__ __forceinline__ int cnb(int nbx, int nby, int b, int d) {
    const int dx = (d == 1 || d == 5 || d == 6) ? 1 : ((d == 2 || d == 7 || d == 8) ? -1 : 0);
    const int dy = (d == 3 || d == 5 || d == 7) ? 1 : ((d == 4 || d == 6 || d == 8) ? -1 : 0);
    const int bx = b % nbx, by = b / nbx;
    return (bx + dx + nbx) % nbx + ((by + dy + nby) % nby) * nbx;
}

// src/coarse_op.h:282-297: one work-item per coarse dof (b, ki)
template <typename T, typename TL, int EPI>
__global__ __launch_bounds__(kBlock) void k_coarse_apply(CoarseView<TL> C, cplx<T>* __restrict__ dst,
                                                         const cplx<T>* __restrict__ src,
                                                         const cplx<T>* __restrict__ bvec,
                                                         const cplx<T>* __restrict__ aux, double* part) {
    const int k = C.k, nb = C.nx * C.ny;
    const int t = blockIdx.x * kBlock + threadIdx.x;
    double acc[3] = {0.0, 0.0, 0.0};
    if (t < nb * k) {
        const int b = t / k, ki = t % k;
        cplx<T> out = {(T)0, (T)0};
#pragma unroll 1
        for (int d = 0; d < 9; d++) {
            const int n = cnb(C.nx, C.ny, b, d);
            const TL* row = C.link + (((size_t)b * 9 + d) * k + ki) * k;
            const cplx<T>* sv = src + (size_t)n * k;
            cplx<T> sum = {(T)0, (T)0};
            for (int kj = 0; kj < k; kj++) sum = sum + cvt<T>(load_c(row + kj)) * sv[kj];
            out = out + sum;
        }
        if (EPI == EPI_RESID) out = bvec[t] - out;
        dst[t] = out;
        if (EPI == EPI_DOTS) {
            const cplx<T> p = cdot(out, aux[t]);
            acc[0] = p.re;
            acc[1] = p.im;
            acc[2] = norm2(out);
        }
    }
    if (EPI == EPI_DOTS) {
        block_sum<3>(acc);
        if (threadIdx.x == 0) {
            double* q = part + (size_t)blockIdx.x * 3;
            q[0] = acc[0];
            q[1] = acc[1];
            q[2] = acc[2];
        }
    }
}
template <typename T, typename TL>
int launch_coarse_apply(const CoarseView<TL>& C, cplx<T>* dst, const cplx<T>* src, Epi epi, const cplx<T>* b,
                        const cplx<T>* aux, double* part, hipStream_t st) {
    const int n = C.nx * C.ny * C.k;
    const int g = grid_for(n);
    if (epi == EPI_NONE) k_coarse_apply<T, TL, EPI_NONE><<<g, kBlock, 0, st>>>(C, dst, src, b, aux, part);
    else if (epi == EPI_RESID) k_coarse_apply<T, TL, EPI_RESID><<<g, kBlock, 0, st>>>(C, dst, src, b, aux, part);
    else k_coarse_apply<T, TL, EPI_DOTS><<<g, kBlock, 0, st>>>(C, dst, src, b, aux, part);
    SMG_LAUNCH_CHECK();
    return g;
}

// ---- Galerkin by coloured probing (the reference probes one column at a time,
// src/coarse_op.h:241-279; a colour class of aggregates >= 3 apart in both
// directions is probed at once).
__device__ __forceinline__ bool in_class(int nbx, int cx, int cy, int cls, int J) {
    return ((J % nbx) % cx == cls % cx) && ((J / nbx) % cy == cls / cx);
}
__global__ void k_probe_vector(int nbx, int nby, int k, int cx, int cy, int cls, int kj, c64* e) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= nbx * nby * k) return;
    const int J = t / k, kk = t % k;
    e[t] = c64{(kk == kj && in_class(nbx, cx, cy, cls, J)) ? 1.0 : 0.0, 0.0};
}
void launch_probe_vector(int nbx, int nby, int k, int cx, int cy, int cls, int kj, c64* e, hipStream_t st) {
    k_probe_vector<<<grid_for(nbx * nby * k), kBlock, 0, st>>>(nbx, nby, k, cx, cy, cls, kj, e);
    SMG_LAUNCH_CHECK();
}
__global__ void k_probe_extract(int nbx, int nby, int k, int cx, int cy, int cls, int kj, const c64* col,
                                c64* link) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= nbx * nby * 9 * k) return;
    const int J = t / (9 * k), d = (t / k) % 9, ki = t % k;
    if (!in_class(nbx, cx, cy, cls, J)) return;
    const int I = cnb(nbx, nby, J, d);
    // on lattices with < 3 aggregates along a direction several directions reach the
    // same aggregate: store the coupling once, in the first such direction
    for (int d2 = 0; d2 < d; d2++)
        if (cnb(nbx, nby, J, d2) == I) return;
    const int rev[9] = {0, 2, 1, 4, 3, 8, 7, 6, 5};
    link[(((size_t)I * 9 + rev[d]) * k + ki) * k + kj] = col[(size_t)I * k + ki];
}
void launch_probe_extract(int nbx, int nby, int k, int cx, int cy, int cls, int kj, const c64* col, c64* link,
                          hipStream_t st) {
    k_probe_extract<<<grid_for(nbx * nby * 9 * k), kBlock, 0, st>>>(nbx, nby, k, cx, cy, cls, kj, col, link);
    SMG_LAUNCH_CHECK();
}

template void launch_prolong<double>(const AggView&, const c64*, const c64*, c64*, int, hipStream_t);
template void launch_prolong<float>(const AggView&, const c32*, const c32*, c32*, int, hipStream_t);
template void launch_restrict<double>(const AggView&, const c64*, const c64*, c64*, hipStream_t);
template void launch_restrict<float>(const AggView&, const c32*, const c32*, c32*, hipStream_t);
template int launch_coarse_apply<double, c64>(const CoarseView<c64>&, c64*, const c64*, Epi, const c64*,
                                              const c64*, double*, hipStream_t);
template int launch_coarse_apply<float, c32>(const CoarseView<c32>&, c32*, const c32*, Epi, const c32*,
                                             const c32*, double*, hipStream_t);
template int launch_coarse_apply<float, ch16>(const CoarseView<ch16>&, c32*, const c32*, Epi, const c32*,
                                              const c32*, double*, hipStream_t);

}  // namespace smg
