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

// src/prolongator.h:63-74: copy one null vector into the block storage (with
// chiral aggregation, one chirality of it)
__global__ void k_block_gather(AggView A, c64* __restrict__ bv, const c64* __restrict__ vec, int slot, int chi) {
    const int bsites = A.bx * A.by, bsize = bsites * A.nd;
    const int nb = A.nbx * A.nby;
    const size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= (size_t)nb * bsize) return;
    const int b = (int)(t / bsize), e = (int)(t % bsize);
    const int loc = e / A.nd, c = e % A.nd;
    int s;
    agg_site(A, b, loc, s);
    const bool keep = chi < 0 || agg_chi_dof(A, c) == chi;
    bv[((size_t)b * A.k + slot) * bsize + e] = keep ? vec[(size_t)s * A.nd + c] : c64{0.0, 0.0};
}
void launch_block_gather(const AggView& A, c64* bv, const c64* vec, int slot, int chi, hipStream_t st) {
    const size_t n = (size_t)A.nbx * A.nby * A.bx * A.by * A.nd;
    k_block_gather<<<grid_for(n), kBlock, 0, st>>>(A, bv, vec, slot, chi);
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
    // chiral aggregation: only the coarse dof of this dof's chirality contribute
    const int j0 = A.chiral ? agg_chi_dof(A, c) * (A.k / 2) : 0, j1 = A.chiral ? j0 + A.k / 2 : A.k;
    cplx<T> acc = {(T)0, (T)0};
    for (int j = j0; j < j1; j++) acc = acc + cc[j] * v[(size_t)j * bsize];
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
    // chiral aggregation: the block vector vanishes off its chirality
    const int c0 = A.chiral ? agg_chi_vec(A, j) * (A.nd / 2) : 0, c1 = A.chiral ? c0 + A.nd / 2 : A.nd;
    cplx<T> acc = {(T)0, (T)0};
    for (int loc = 0; loc < bsites; loc++) {
        int s;
        agg_site(A, b, loc, s);
        const cplx<T>* fp = fine + (size_t)s * A.nd;
        cplx<T> a = cdot(v[loc * A.nd + c0], fp[c0]);
        for (int c = c0 + 1; c < c1; c++) a = a + cdot(v[loc * A.nd + c], fp[c]);
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

// src/coarse_op.h:282-297.  Generic form: one work-item per coarse dof (b, ki).
template <typename T, typename TL, int EPI>
__global__ __launch_bounds__(kBlock) void k_coarse_apply(CoarseView<TL> C, cplx<T>* __restrict__ dst,
                                                         const cplx<T>* __restrict__ src,
                                                         const cplx<T>* __restrict__ bvec,
                                                         const cplx<T>* __restrict__ aux, Fin fin) {
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
    if (EPI == EPI_DOTS) reduce_finish<3>(acc, fin, gridDim.x, blockIdx.x);
}

// Power-of-two k: one lane per link entry (b, ki, kj), so the k x k block of each
// direction is one coalesced read, and the row sums over kj are shuffles inside
// groups of K lanes.  The workgroups (at most kLanesGrid) stride over the nb*k*k
// entries: a reduction then finishes over few partial rows (each workgroup waits
// for its partials to reach memory before it retires), and the stride, a multiple
// of kBlock and so of K, keeps every row inside one wave and one iteration; the
// loop bound is uniform per workgroup, so every lane reaches every shuffle.
constexpr int kLanesGrid = 2048;
template <typename T, typename TL, int EPI, int K>
__global__ __launch_bounds__(kBlock) void k_coarse_apply_lanes(CoarseView<TL> C, cplx<T>* __restrict__ dst,
                                                               const cplx<T>* __restrict__ src,
                                                               const cplx<T>* __restrict__ bvec,
                                                               const cplx<T>* __restrict__ aux, Fin fin) {
    static_assert(kBlock % K == 0 && K <= 64, "a row of K lanes must sit inside one wave");
    constexpr int KK = K * K;
    const int nb = C.nx * C.ny;
    const size_t total = (size_t)nb * KK;
    double acc[3] = {0.0, 0.0, 0.0};
    for (size_t t0 = (size_t)blockIdx.x * kBlock; t0 < total; t0 += (size_t)gridDim.x * kBlock) {
        const size_t t = t0 + threadIdx.x;
        const int b = (int)(t / KK), e = (int)(t % KK), kj = e % K;
        cplx<T> sum = {(T)0, (T)0};
        if (b < nb) {
            const int bx = b % C.nx, by = b / C.nx;
            const int xm = (bx == 0 ? C.nx - 1 : bx - 1), xp = (bx + 1 == C.nx ? 0 : bx + 1);
            const int ym = (by == 0 ? C.ny - 1 : by - 1), yp = (by + 1 == C.ny ? 0 : by + 1);
            const int nbr[9] = {b, xp + by * C.nx, xm + by * C.nx, bx + yp * C.nx, bx + ym * C.nx,
                                xp + yp * C.nx, xp + ym * C.nx, xm + yp * C.nx, xm + ym * C.nx};
            const TL* lk = C.link + (size_t)b * 9 * KK + e;
#pragma unroll
            for (int d = 0; d < 9; d++) sum = sum + cvt<T>(load_c(lk + d * KK)) * src[(size_t)nbr[d] * K + kj];
        }
#pragma unroll
        for (int off = K / 2; off > 0; off >>= 1) {
            sum.re += __shfl_xor(sum.re, off, 64);
            sum.im += __shfl_xor(sum.im, off, 64);
        }
        if (b < nb && kj == 0) {
            const size_t o = (size_t)b * K + e / K;
            cplx<T> out = sum;
            if (EPI == EPI_RESID) out = bvec[o] - out;
            dst[o] = out;
            if (EPI == EPI_DOTS) {
                const cplx<T> p = cdot(out, aux[o]);
                acc[0] += p.re;
                acc[1] += p.im;
                acc[2] += norm2(out);
            }
        }
    }
    if (EPI == EPI_DOTS) reduce_finish<3>(acc, fin, gridDim.x, blockIdx.x);
}

template <typename T, typename TL, int EPI>
static void coarse_apply_dispatch(const CoarseView<TL>& C, cplx<T>* dst, const cplx<T>* src, const cplx<T>* b,
                                  const cplx<T>* aux, const Fin& fin, hipStream_t st) {
    const size_t nb = (size_t)C.nx * C.ny;
    const size_t lanes = nb * C.k * C.k;
    switch (C.k) {
#define SMG_CK(K)                                                                                         \
    case K:                                                                                                 \
        k_coarse_apply_lanes<T, TL, EPI, K><<<std::min(grid_for(lanes), kLanesGrid), kBlock, 0, st>>>(    \
            C, dst, src, b, aux, fin);                                                                      \
        break;
        SMG_CK(2)
        SMG_CK(4)
        SMG_CK(8)
        SMG_CK(16)
        SMG_CK(32)
#undef SMG_CK
    default:
        k_coarse_apply<T, TL, EPI><<<grid_for(nb * C.k), kBlock, 0, st>>>(C, dst, src, b, aux, fin);
    }
}

template <typename T, typename TL>
void launch_coarse_apply(const CoarseView<TL>& C, cplx<T>* dst, const cplx<T>* src, Epi epi, const cplx<T>* b,
                         const cplx<T>* aux, const Fin& fin, hipStream_t st) {
    if (epi == EPI_NONE) coarse_apply_dispatch<T, TL, EPI_NONE>(C, dst, src, b, aux, fin, st);
    else if (epi == EPI_RESID) coarse_apply_dispatch<T, TL, EPI_RESID>(C, dst, src, b, aux, fin, st);
    else coarse_apply_dispatch<T, TL, EPI_DOTS>(C, dst, src, b, aux, fin, st);
    SMG_LAUNCH_CHECK();
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

// ---- dense inverse of the coarsest operator (HPD: Gauss-Jordan without pivoting)
// M (row-major n x n, zeroed by the caller) += the 9-point stencil of C
__global__ void k_coarse_to_dense(CoarseView<c64> C, c64* __restrict__ M, size_t n) {
    const int k = C.k;
    const size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x;
    const size_t nb = (size_t)C.nx * C.ny;
    if (t >= nb * k * k) return;
    const int b = (int)(t / (k * k)), ki = (int)((t / k) % k), kj = (int)(t % k);
    const size_t row = (size_t)b * k + ki;
    for (int d = 0; d < 9; d++) {
        const int nbb = cnb(C.nx, C.ny, b, d);
        const c64 v = C.link[(((size_t)b * 9 + d) * k + ki) * k + kj];
        c64& m = M[row * n + (size_t)nbb * k + kj];
        m = m + v;
    }
}
// backward direction j (0..3) of a site: -x, -y, -x+y, -x-y (indices 2, 4, 7, 8 of the
// 9-direction order {0, +x, -x, +y, -y, +x+y, +x-y, -x+y, -x-y}); from the neighbour it
// reaches, the way back is +x, +y, +x-y, +x+y (indices 1, 3, 6, 5)
__global__ void k_half_hermitian(ch16* __restrict__ link, int nx, int ny, int k) {
    const size_t kk = (size_t)k * k;
    const size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= (size_t)nx * ny * 4 * kk) return;
    const int b = (int)(t / (4 * kk)), j = (int)(t / kk % 4);
    const int e = (int)(t % kk), ki = e / k, kj = e % k;
    const int d = (j == 0) ? 2 : (j == 1) ? 4 : (j == 2) ? 7 : 8;
    const int rd = (j == 0) ? 1 : (j == 1) ? 3 : (j == 2) ? 6 : 5;
    const int dx = (j == 1) ? 0 : -1, dy = (j == 0) ? 0 : (j == 1 || j == 3) ? -1 : 1;
    const int bx = b % nx, by = b / nx;
    const int n = (bx + dx + nx) % nx + ((by + dy + ny) % ny) * nx;
    const ch16 v = link[((size_t)n * 9 + rd) * kk + (size_t)kj * k + ki];
    link[((size_t)b * 9 + d) * kk + e] = ch16{v.re, __hneg(v.im)};
}
void launch_half_hermitian(ch16* link, int nx, int ny, int k, hipStream_t st) {
    SMG_REQUIRE(nx >= 3 && ny >= 3, "Hermitian FP16 links need coarse extents >= 3");
    k_half_hermitian<<<grid_for((size_t)nx * ny * 4 * k * k), kBlock, 0, st>>>(link, nx, ny, k);
    SMG_LAUNCH_CHECK();
}

void launch_coarse_to_dense(const CoarseView<c64>& C, c64* M, size_t n, hipStream_t st) {
    k_coarse_to_dense<<<grid_for((size_t)C.nx * C.ny * C.k * C.k), kBlock, 0, st>>>(C, M, n);
    SMG_LAUNCH_CHECK();
}
// step p of in-place Gauss-Jordan: the scaled pivot row and the pivot column
__global__ void k_gj_pivot(const c64* __restrict__ M, size_t n, size_t p, c64* __restrict__ prow,
                           c64* __restrict__ pcol) {
    const c64 piv = M[p * n + p];
    const c64 inv = cdiv(c64{1.0, 0.0}, piv);
    for (size_t j = threadIdx.x; j < n; j += blockDim.x) {
        prow[j] = (j == p) ? inv : M[p * n + j] * inv;
        pcol[j] = (j == p) ? c64{0.0, 0.0} : M[j * n + p];
    }
}
__global__ void k_gj_update(c64* __restrict__ M, size_t n, size_t p, const c64* __restrict__ prow,
                            const c64* __restrict__ pcol) {
    const size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= n * n) return;
    const size_t i = t / n, j = t % n;
    if (i == p) {
        M[t] = prow[j];
        return;
    }
    const c64 a = (j == p) ? c64{0.0, 0.0} : M[t];
    M[t] = a - pcol[i] * prow[j];
}
void launch_dense_invert(c64* M, size_t n, c64* prow, c64* pcol, hipStream_t st) {
    for (size_t p = 0; p < n; p++) {
        k_gj_pivot<<<1, kBlock, 0, st>>>(M, n, p, prow, pcol);
        k_gj_update<<<grid_for(n * n), kBlock, 0, st>>>(M, n, p, prow, pcol);
    }
    SMG_LAUNCH_CHECK();
}
// y = M x, one wave64 per row
template <typename T>
__global__ __launch_bounds__(kBlock) void k_dense_matvec(const cplx<T>* __restrict__ M, const cplx<T>* __restrict__ x,
                                                         cplx<T>* __restrict__ y, size_t n) {
    const size_t row = (size_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    double re = 0.0, im = 0.0;
    if (row < n) {
        const cplx<T>* mr = M + row * n;
        for (size_t j = lane; j < n; j += 64) {
            const cplx<T> p = mr[j] * x[j];
            re += (double)p.re;
            im += (double)p.im;
        }
    }
    re = wave_sum(re);
    im = wave_sum(im);
    if (row < n && lane == 0) y[row] = cplx<T>{(T)re, (T)im};
}
template <typename T>
void launch_dense_matvec(const cplx<T>* M, const cplx<T>* x, cplx<T>* y, size_t n, hipStream_t st) {
    k_dense_matvec<T><<<(unsigned)((n + 3) / 4), kBlock, 0, st>>>(M, x, y, n);
    SMG_LAUNCH_CHECK();
}
template void launch_dense_matvec<double>(const c64*, const c64*, c64*, size_t, hipStream_t);
template void launch_dense_matvec<float>(const c32*, const c32*, c32*, size_t, hipStream_t);

template void launch_prolong<double>(const AggView&, const c64*, const c64*, c64*, int, hipStream_t);
template void launch_prolong<float>(const AggView&, const c32*, const c32*, c32*, int, hipStream_t);
template void launch_restrict<double>(const AggView&, const c64*, const c64*, c64*, hipStream_t);
template void launch_restrict<float>(const AggView&, const c32*, const c32*, c32*, hipStream_t);
template void launch_coarse_apply<double, c64>(const CoarseView<c64>&, c64*, const c64*, Epi, const c64*,
                                               const c64*, const Fin&, hipStream_t);
template void launch_coarse_apply<float, c32>(const CoarseView<c32>&, c32*, const c32*, Epi, const c32*,
                                              const c32*, const Fin&, hipStream_t);
template void launch_coarse_apply<float, ch16>(const CoarseView<ch16>&, c32*, const c32*, Epi, const c32*,
                                               const c32*, const Fin&, hipStream_t);

}  // namespace smg
