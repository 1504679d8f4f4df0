// eigen.hip -- the block vector kernel of the Lanczos eigensolver: y = (y or 0) +/-
// sum_i c_i V_i over up to 32 vectors with complex coefficients c_i from device
// memory (so a multi-dot's result feeds it without a host round trip).  HBM-bound:
// one pass over the nv vectors and y.
#include "smg_kernels.h"

namespace smg {

template <typename T>
__global__ __launch_bounds__(kBlock) void k_lincomb(cplx<T>* __restrict__ y, VecPtrs<T> V, int nv,
                                                    const double* __restrict__ coef, int mode, size_t n) {
    for (size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x; t < n; t += (size_t)gridDim.x * kBlock) {
        c64 acc = {0.0, 0.0};
        for (int i = 0; i < nv; i++) {
            const c64 c = {coef[2 * i], coef[2 * i + 1]};
            const cplx<T> v = V.v[i][t];
            acc = acc + c * c64{(double)v.re, (double)v.im};
        }
        cplx<T> out;
        if (mode == 0) {
            out = cplx<T>{(T)acc.re, (T)acc.im};
        } else {
            const cplx<T> o = y[t];
            const double s = mode == 1 ? 1.0 : -1.0;
            out = cplx<T>{(T)(o.re + s * acc.re), (T)(o.im + s * acc.im)};
        }
        y[t] = out;
    }
}
template <typename T>
void launch_lincomb(cplx<T>* y, const VecPtrs<T>& V, int nv, const double* coef, int mode, size_t n, hipStream_t st) {
    SMG_REQUIRE(nv >= 1 && nv <= 32, "1..32 vectors per linear combination");
    k_lincomb<T><<<grid_for(n) < (1 << 16) ? grid_for(n) : (1 << 16), kBlock, 0, st>>>(y, V, nv, coef, mode, n);
    SMG_LAUNCH_CHECK();
}
template void launch_lincomb<double>(c64*, const VecPtrs<double>&, int, const double*, int, size_t, hipStream_t);


// ----------------------------------------------------------------- block vectors
// A block holds N vectors of n entries in the batched layout v[i*N + j] (vector j
// innermost), N a power of two <= 32.  The Gram kernel contracts a concatenation
// of nl left blocks with nr right blocks: G[a][b] = sum_i conj(L_a(i)) R_b(i) for
// the nl*N left and nr*N right vectors.  Rows are staged through LDS a tile at a
// time; each thread owns up to kGramPPT (a, b) pairs; per-workgroup partials are
// summed in a fixed order by k_blk_sum (deterministic).
constexpr int kGramPPT = 16;
constexpr int kGramLdsC64 = 3072;   // 48 KB of LDS per workgroup

template <int N>
__global__ __launch_bounds__(kBlock) void k_blk_gram(BlockPtrs Lb, int nl, BlockPtrs Rb, int nr, size_t n,
                                                     double* __restrict__ part) {
    __shared__ c64 tile[kGramLdsC64];
    const int CL = nl * N, CR = nr * N, CT = CL + CR;
    const int rows = kGramLdsC64 / CT;   // >= 1 (host checks)
    const int pairs = CL * CR;
    double acc[2 * kGramPPT];
#pragma unroll
    for (int p = 0; p < 2 * kGramPPT; p++) acc[p] = 0.0;
    const size_t ntiles = (n + rows - 1) / rows;
    for (size_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const size_t r0 = t * rows;
        // stage rows [r0, r0+rows) of every block: tile[r*CT + col]
        for (int e = threadIdx.x; e < rows * CT; e += kBlock) {
            const int r = e / CT, col = e % CT;
            const size_t row = r0 + r;
            c64 v = {0.0, 0.0};
            if (row < n) {
                if (col < CL) v = Lb.p[col / N][row * N + (col % N)];
                else v = Rb.p[(col - CL) / N][row * N + ((col - CL) % N)];
            }
            tile[e] = v;
        }
        __syncthreads();
#pragma unroll
        for (int p = 0; p < kGramPPT; p++) {
            const int q = threadIdx.x + p * kBlock;
            if (q < pairs) {
                const int a = q / CR, b = CL + q % CR;
                double re = acc[2 * p], im = acc[2 * p + 1];
                for (int r = 0; r < rows; r++) {
                    const c64 x = tile[r * CT + a], y = tile[r * CT + b];
                    re += x.re * y.re + x.im * y.im;
                    im += x.re * y.im - x.im * y.re;
                }
                acc[2 * p] = re;
                acc[2 * p + 1] = im;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int p = 0; p < kGramPPT; p++) {
        const int q = threadIdx.x + p * kBlock;
        if (q < pairs) {
            part[(size_t)blockIdx.x * 2 * pairs + 2 * q] = acc[2 * p];
            part[(size_t)blockIdx.x * 2 * pairs + 2 * q + 1] = acc[2 * p + 1];
        }
    }
}
// out[c] = sum over workgroups w (in order) of part[w*cnt + c]
__global__ __launch_bounds__(kBlock) void k_blk_sum(const double* __restrict__ part, int nblk, int cnt,
                                                    double* __restrict__ out) {
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= cnt) return;
    double s = 0.0;
    for (int w = 0; w < nblk; w++) s += part[(size_t)w * cnt + c];
    out[c] = s;
}
int blk_gram_grid(size_t n) {
    const size_t g = (n + 63) / 64;
    return (int)(g < 1 ? 1 : (g > 512 ? 512 : g));
}
void launch_blk_gram(const BlockPtrs& L, int nl, const BlockPtrs& R, int nr, int N, size_t n, double* part,
                     double* out, hipStream_t st) {
    SMG_REQUIRE(N >= 1 && N <= 32 && (N & (N - 1)) == 0, "block Gram: N must be a power of two <= 32");
    SMG_REQUIRE(nl >= 1 && nr >= 1 && nl <= kMaxBlocks && nr <= kMaxBlocks, "block Gram: 1..6 blocks a side");
    SMG_REQUIRE(nl * N * nr * N <= kGramPPT * kBlock, "block Gram: too many pairs for one launch");
    SMG_REQUIRE((nl + nr) * N <= kGramLdsC64, "block Gram: tile too wide");
    const int grid = blk_gram_grid(n);
    switch (N) {
#define SMG_GRAM_CASE(NN) \
    case NN: k_blk_gram<NN><<<grid, kBlock, 0, st>>>(L, nl, R, nr, n, part); break;
        SMG_GRAM_CASE(1) SMG_GRAM_CASE(2) SMG_GRAM_CASE(4) SMG_GRAM_CASE(8) SMG_GRAM_CASE(16) SMG_GRAM_CASE(32)
#undef SMG_GRAM_CASE
    }
    SMG_LAUNCH_CHECK();
    const int cnt = 2 * nl * N * nr * N;
    k_blk_sum<<<(cnt + kBlock - 1) / kBlock, kBlock, 0, st>>>(part, grid, cnt, out);
    SMG_LAUNCH_CHECK();
}

// out_m[i*N + c] = sum_{s < ns} sum_{j < kact} S_s[i*N + j] C_m[(s*N + j)*N + c] for the
// one or two outputs m (C_m complex, device memory); columns j >= kact of a source
// are padding and are not read; one thread per (i, c)
__global__ __launch_bounds__(kBlock) void k_blk_combine(c64* __restrict__ out0, c64* __restrict__ out1,
                                                        BlockPtrs S, int ns, int N, int kact,
                                                        const c64* __restrict__ C0, const c64* __restrict__ C1,
                                                        size_t total) {
    for (size_t e = (size_t)blockIdx.x * kBlock + threadIdx.x; e < total; e += (size_t)gridDim.x * kBlock) {
        const size_t i = e / N;
        const int c = (int)(e % N);
        c64 a0 = {0.0, 0.0}, a1 = {0.0, 0.0};
        for (int s = 0; s < ns; s++) {
            const c64* row = S.p[s] + i * N;
            for (int j = 0; j < kact; j++) {
                const c64 v = row[j];
                const size_t ci = ((size_t)s * N + j) * N + c;
                a0 = a0 + v * C0[ci];
                if (out1) a1 = a1 + v * C1[ci];
            }
        }
        out0[e] = a0;
        if (out1) out1[e] = a1;
    }
}
void launch_blk_combine(c64* out0, c64* out1, const BlockPtrs& S, int ns, int N, int kact, const c64* C0,
                        const c64* C1, size_t n, hipStream_t st) {
    SMG_REQUIRE(ns >= 1 && ns <= kMaxBlocks && N >= 1 && N <= 32 && kact >= 1 && kact <= N,
                "block combine: bad shape");
    const size_t total = n * N;
    const size_t g = (total + kBlock - 1) / kBlock;
    k_blk_combine<<<(int)(g > 65536 ? 65536 : g), kBlock, 0, st>>>(out0, out1, S, ns, N, kact, C0, C1, total);
    SMG_LAUNCH_CHECK();
}

// per-column scalings and updates of a block: y_j = a_j x_j + b_j y_j (a, b complex
// per column, device memory; x may be null for y_j = b_j y_j)
__global__ __launch_bounds__(kBlock) void k_blk_cols(c64* __restrict__ y, const c64* __restrict__ x,
                                                     const c64* __restrict__ a, const c64* __restrict__ b, int N,
                                                     size_t total) {
    const size_t e = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (e >= total) return;
    const int c = (int)(e % N);
    c64 v = b[c] * y[e];
    if (x) v = v + a[c] * x[e];
    y[e] = v;
}
void launch_blk_cols(c64* y, const c64* x, const c64* a, const c64* b, int N, size_t n, hipStream_t st) {
    const size_t total = n * N;
    k_blk_cols<<<(int)((total + kBlock - 1) / kBlock), kBlock, 0, st>>>(y, x, a, b, N, total);
    SMG_LAUNCH_CHECK();
}

}  // namespace smg
