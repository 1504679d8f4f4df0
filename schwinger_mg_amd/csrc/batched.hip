// batched.hip -- multi-right-hand-side kernels (layout: see smg_batched.h).
//
// Work-item t -> (i, j) = (t / N, t % N): the N lanes of one site or dof are
// consecutive, so field loads are N-wide coalesced runs and the per-site data
// (links, clover term, aggregate vectors, coarse links) is one broadcast load
// shared by those lanes.  Reductions are per RHS: lanes with equal j are summed
// by xor-shuffles over the lane bits above log2(N), then across the waves in LDS;
// the last workgroup sums the partial rows in index order and runs the scalar
// program of every RHS.
#include <type_traits>

#include "smg_batched.h"
#include "tune.h"

namespace smg {

// Workgroups of a reducing batched kernel: kRedGrid for the stencil with fused
// dots, kRedGridCoarse for the coarse apply with fused dots, kRedGridVec for the
// vector reductions and updates, kRedGridMulti for the multi-dot (up to 32
// partials per RHS).  Every workgroup of a reducing kernel waits for its partials
// to reach memory before it retires, so fewer workgroups that loop over more items
// are faster.  Measured on gfx950 at 32 RHS: the paired FP32 stencil with dots
// (L=512) 94 us at 512 workgroups, 106 us at 2048, 139 us at 8192; the k=32 coarse
// apply with dots 56 us at 256 workgroups and 63 us at 1024 (64^2 sites), 222 us
// and 250 us at 256 and 2048 (128^2); the coarsest-level CG update 35 us at 512
// and 51 us at 8192.  The partial rows, (nblk + nblk/32) * NC * N doubles, fit the
// 2M-double partial buffer, and the 1 + nblk/32 completion counters fit the
// 1024-entry counter array.
constexpr int kRedGrid = 512;
constexpr int kRedGridCoarse = 256;
constexpr int kRedGridVec = 512;
constexpr int kRedGridMulti = 1024;
static_assert((kRedGrid + kRedGrid / 32) * 3 * 64 <= (1 << 21), "partials exceed the reduction buffer");
static_assert((kRedGridMulti + kRedGridMulti / 32) * 32 * 32 <= (1 << 21), "partials exceed the reduction buffer");
static_assert(1 + kRedGridMulti / 32 <= 1024, "completion counters exceed the counter array");

__device__ __forceinline__ int lg2(int N) { return 31 - __clz(N); }

// Per-RHS reduce-and-finish.  v[NC] holds this lane's partials for RHS j = lane % N.
//
// Two-level, fixed-order tree: every workgroup writes its row of NC*N partials;
// the last workgroup to finish within a group of kGroup consecutive workgroups
// sums that group's rows in index order into a group row; the last group to
// finish sums the group rows in index order and runs the scalar program of
// every RHS.  Each level is at most kGroup (or nblk/kGroup) loads deep per
// column, all issued before the adds, so the finish costs a few memory
// latencies instead of one latency per workgroup.  The result depends only on
// nblk, not on the order in which workgroups finish.
// Counters: counter[0] counts finished groups, counter[1 + g] the workgroups of
// group g; each is reset to zero by the workgroup that consumed it.
constexpr int kGroup = 32;

// Column sums of rows [r0, r1) of a row-major table of tot columns, every column
// summed in a fixed order with all threads of the workgroup: thread t adds rows
// r0 + q, r0 + q + Q, ... of column t % tot (q = t / tot, Q = kBlock / tot parts),
// then the Q parts are combined in order through LDS.  out[c] (LDS) receives the
// sums.  At most 8 loads per thread are in flight, which keeps the register
// footprint of the calling (streaming) kernel small.
__device__ __forceinline__ void column_sums(const double* tab, int r0, int r1, int tot, double* red, double* out) {
    if (tot >= kBlock) {
        for (int c = threadIdx.x; c < tot; c += kBlock) {
            double a = 0.0;
#pragma unroll 8
            for (int i = r0; i < r1; i++) a += part_load(tab + (size_t)i * tot + c);
            out[c] = a;
        }
        __syncthreads();
        return;
    }
    const int Q = kBlock / tot, c = threadIdx.x % tot, q = threadIdx.x / tot;
    double a = 0.0;
    if (q < Q) {
#pragma unroll 8
        for (int i = r0 + q; i < r1; i += Q) a += part_load(tab + (size_t)i * tot + c);
    }
    red[threadIdx.x] = a;
    __syncthreads();
    if (threadIdx.x < tot) {
        double sum = 0.0;
        for (int qq = 0; qq < Q; qq++) sum += red[qq * tot + threadIdx.x];
        out[threadIdx.x] = sum;
    }
    __syncthreads();
}

template <int NC, int NMAX>
__device__ __forceinline__ void reduce_finish_b(double (&v)[NC], const FinB& f, int N, int nblk, int blk) {
    __shared__ double sm[kBlock / 64][NC * NMAX];
    __shared__ double red[kBlock];
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
        part_store(f.part + (size_t)blk * tot + t, (sm[0][t] + sm[1][t]) + (sm[2][t] + sm[3][t]));
    stores_complete();
    __syncthreads();
    const int ngrp = (nblk + kGroup - 1) / kGroup, g = blk / kGroup;
    const int g0 = g * kGroup, gn = min(nblk - g0, kGroup);
    if (threadIdx.x == 0) last = (counter_arrive(f.counter + 1 + g) == (unsigned)gn - 1u);
    __syncthreads();
    if (!last) return;
    double* out = &sm[0][0];   // tot <= NC * NMAX entries
    double* grow = f.part + (size_t)nblk * tot;   // group rows follow the workgroup rows
    column_sums(f.part, g0, g0 + gn, tot, red, out);
    for (int t = threadIdx.x; t < tot; t += kBlock) part_store(grow + (size_t)g * tot + t, out[t]);
    stores_complete();
    __syncthreads();
    if (threadIdx.x == 0) {
        counter_reset(f.counter + 1 + g);
        last = (counter_arrive(f.counter) == (unsigned)ngrp - 1u);
    }
    __syncthreads();
    if (!last) return;
    column_sums(grow, 0, ngrp, tot, red, out);
    if (threadIdx.x < N) {
        const int j = threadIdx.x;
        double w[NC];
#pragma unroll
        for (int c = 0; c < NC; c++) w[c] = out[c * N + j];
        fin_program(f.op, w, NC, f.out + (size_t)j * f.ostride, f.aux ? f.aux + (size_t)j * f.astride : nullptr);
    }
    if (f.done) {   // all RHS converged?  (flags written by this workgroup's threads above)
        __syncthreads();
        if (threadIdx.x == 0) {
            int all = 1;
            for (int j = 0; j < N; j++) all &= *reinterpret_cast<const int*>(f.aux + (size_t)j * f.astride + 8) != 0;
            *f.done = all;
        }
    }
    if (threadIdx.x == 0) counter_reset(f.counter);
}

static inline int b_grid(size_t items, int cap) {
    size_t g = (items + kBlock - 1) / kBlock;
    if (g < 1) g = 1;
    return (int)(g > (size_t)cap ? cap : g);
}

// ----------------------------------------------------------------- Wilson-clover
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

// P consecutive right-hand sides of one spin component: one 16-byte access for
// P = 2 in FP32 (and for P = 1 in FP64)
template <typename T, int P>
struct RhsRun {
    cplx<T> v[P];
    __device__ __forceinline__ void load(const cplx<T>* __restrict__ p) {
        if constexpr (P == 2 && sizeof(T) == 4) {
            const float4 w = *reinterpret_cast<const float4*>(p);
            v[0] = cplx<T>{w.x, w.y};
            v[1] = cplx<T>{w.z, w.w};
        } else {
#pragma unroll
            for (int q = 0; q < P; q++) v[q] = p[q];
        }
    }
    __device__ __forceinline__ void store(cplx<T>* __restrict__ p) const {
        if constexpr (P == 2 && sizeof(T) == 4) {
            *reinterpret_cast<float4*>(p) = float4{v[0].re, v[0].im, v[1].re, v[1].im};
        } else {
#pragma unroll
            for (int q = 0; q < P; q++) p[q] = v[q];
        }
    }
    // a streaming store: the result is not read again by this kernel, so it is written
    // with the non-temporal hint and does not displace the sources held in the caches
    __device__ __forceinline__ void store_nt(cplx<T>* __restrict__ p) const {
        using V4 = typename std::conditional<sizeof(T) == 4, float, double>::type;
        typedef V4 v2 __attribute__((ext_vector_type(2)));
        typedef float f4 __attribute__((ext_vector_type(4)));
        if constexpr (P == 2 && sizeof(T) == 4) {
            __builtin_nontemporal_store(f4{v[0].re, v[0].im, v[1].re, v[1].im}, reinterpret_cast<f4*>(p));
        } else {
#pragma unroll
            for (int q = 0; q < P; q++) __builtin_nontemporal_store(v2{v[q].re, v[q].im}, reinterpret_cast<v2*>(p + q));
        }
    }
};

// Per-RHS sums of a kernel whose lanes hold runs of P right-hand sides (lane l:
// RHS P*(l % (N/P)) + q in acc[c][q]) into the layout reduce_finish_b expects
// (lane j < N: RHS j; lanes >= N: zero).
template <int NC, int P>
__device__ __forceinline__ void runs_to_lanes(double (&acc)[NC][P], double (&v)[NC], int N) {
    if constexpr (P == 1) {
#pragma unroll
        for (int c = 0; c < NC; c++) v[c] = acc[c][0];
    } else {
        const int lane = threadIdx.x & 63, NP = N / P;
#pragma unroll
        for (int c = 0; c < NC; c++)
#pragma unroll
            for (int q = 0; q < P; q++)
                for (int off = NP; off < 64; off <<= 1) acc[c][q] += __shfl_xor(acc[c][q], off, 64);
        const int jl = lane & (N - 1);
#pragma unroll
        for (int c = 0; c < NC; c++) {
            double x = 0.0;
#pragma unroll
            for (int q = 0; q < P; q++) {
                const double y = __shfl(acc[c][q], jl / P, 64);
                if (jl % P == q) x = y;
            }
            v[c] = lane < N ? x : 0.0;
        }
    }
}
// Run length of the batched vector kernels: pairs of right-hand sides in FP32
// (N >= 2), single ones otherwise.  run_dispatch calls f with the run length as a
// compile-time constant.
template <typename T, typename F>
__host__ inline void run_dispatch(int N, F&& f) {
    if constexpr (sizeof(T) == 4) {
        if (N >= 2) {
            f(std::integral_constant<int, 2>{});
            return;
        }
    }
    f(std::integral_constant<int, 1>{});
}

// One work-item per (site, run of P right-hand sides).  P = 2 in FP32 halves the
// load instructions and the index arithmetic per byte: measured at L=512, 32 RHS,
// 51.5 us per launch against 76.5 us with one right-hand side per work-item.
template <typename T, int EPI, int P, bool NT = false>
__global__ __launch_bounds__(kBlock) void k_b_dslash(DiracView<T> D, cplx<T>* __restrict__ dst,
                                                     const cplx<T>* __restrict__ src, T mu, int g5in, int g5out,
                                                     const cplx<T>* __restrict__ bvec,
                                                     const cplx<T>* __restrict__ aux, FinB fin, int N) {
    const int L = D.L, V = L * L, lNP = lg2(N / P);
    const size_t items = (size_t)V << lNP;
    double acc[3][P];
#pragma unroll
    for (int q = 0; q < P; q++) acc[0][q] = acc[1][q] = acc[2][q] = 0.0;
    for (size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x; t < items; t += (size_t)gridDim.x * kBlock) {
        const int s = (int)(t >> lNP), j = (int)(t & (N / P - 1)) * P;
        const int x = s % L, y = s / L;
        const int sxp = (x + 1 == L) ? s - x : s + 1;
        const int sxm = (x == 0) ? s + L - 1 : s - 1;
        const int syp = (y + 1 == L) ? x : s + L;
        const int sym = (y == 0) ? s + V - L : s - L;
        const int ss[5] = {s, sxp, sxm, syp, sym};
        RhsRun<T, P> in[5][2];
#pragma unroll
        for (int n = 0; n < 5; n++)
#pragma unroll
            for (int sp = 0; sp < 2; sp++) in[n][sp].load(src + (2 * (size_t)ss[n] + sp) * N + j);
        const cplx<T> u0 = D.U[s], u0b = D.U[sxm], u1 = D.U[V + s], u1b = D.U[V + sym];
        const T C = D.C ? D.C[s] : (T)0;
        const size_t i0 = (2 * (size_t)s) * N + j, i1 = i0 + N;
        RhsRun<T, P> out[2], bb[2], ax[2];
        if (EPI == EPI_RESID) {
            bb[0].load(bvec + i0);
            bb[1].load(bvec + i1);
        }
        if (EPI == EPI_DOTS) {
            ax[0].load(aux + i0);
            ax[1].load(aux + i1);
        }
#pragma unroll
        for (int q = 0; q < P; q++) {
            const T g1 = g5in ? (T)-1 : (T)1;   // gamma5 on the input: lower spin negated
            cplx<T> h0, h1;
            hop_b(D.r, u0, u0b, u1, u1b, in[1][0].v[q], g1 * in[1][1].v[q], in[2][0].v[q], g1 * in[2][1].v[q],
                  in[3][0].v[q], g1 * in[3][1].v[q], in[4][0].v[q], g1 * in[4][1].v[q], h0, h1);
            cplx<T> o0 = mk<T>(D.diag + C, mu) * in[0][0].v[q] + h0;
            cplx<T> o1 = mk<T>(D.diag - C, -mu) * (g1 * in[0][1].v[q]) + h1;
            if (g5out) o1 = -o1;
            if (EPI == EPI_RESID) {
                o0 = bb[0].v[q] - o0;
                o1 = bb[1].v[q] - o1;
            }
            out[0].v[q] = o0;
            out[1].v[q] = o1;
            if (EPI == EPI_DOTS) {
                const cplx<T> d0 = cdot(o0, ax[0].v[q]), d1 = cdot(o1, ax[1].v[q]);
                acc[0][q] += (double)d0.re + (double)d1.re;
                acc[1][q] += (double)d0.im + (double)d1.im;
                acc[2][q] += (double)norm2(o0) + (double)norm2(o1);
            }
        }
        if constexpr (NT) {
            out[0].store_nt(dst + i0);
            out[1].store_nt(dst + i1);
        } else {
            out[0].store(dst + i0);
            out[1].store(dst + i1);
        }
    }
    if (EPI == EPI_DOTS) {
        double v[3];
        runs_to_lanes<3, P>(acc, v, N);
        reduce_finish_b<3, 64>(v, fin, N, gridDim.x, blockIdx.x);
    }
}

template <typename T, int P, bool NT>
static void dslash_pn(const DiracView<T>& D, cplx<T>* dst, const cplx<T>* src, T mu, int g5in, int g5out, Epi epi,
                      const cplx<T>* b, const cplx<T>* aux, const FinB& fin, int N, hipStream_t st) {
    const size_t items = (size_t)D.L * D.L * N / P;
    if (epi == EPI_NONE)
        k_b_dslash<T, EPI_NONE, P, NT><<<b_grid(items, 1 << 16), kBlock, 0, st>>>(D, dst, src, mu, g5in, g5out, b,
                                                                                aux, fin, N);
    else if (epi == EPI_RESID)
        k_b_dslash<T, EPI_RESID, P, NT><<<b_grid(items, 1 << 16), kBlock, 0, st>>>(D, dst, src, mu, g5in, g5out, b,
                                                                                 aux, fin, N);
    else {
        const int choice = tune_lookup(TuneKey{TUNE_DSLASH_DOTS, D.L, N, (int)sizeof(T), 0, 0});
        const int g = choice < 0 ? kRedGrid : kTuneGrids[choice];
        k_b_dslash<T, EPI_DOTS, P, NT><<<b_grid(items, g), kBlock, 0, st>>>(D, dst, src, mu, g5in, g5out, b, aux,
                                                                          fin, N);
    }
}
// the store hint the tuner picked for this shape and epilogue (default: ordinary stores)
template <typename T, int P>
static void dslash_p(const DiracView<T>& D, cplx<T>* dst, const cplx<T>* src, T mu, int g5in, int g5out, Epi epi,
                     const cplx<T>* b, const cplx<T>* aux, const FinB& fin, int N, hipStream_t st) {
    const int nt = tune_lookup(TuneKey{TUNE_DSLASH_STORE, D.L, N, (int)sizeof(T), (int)epi, 0});
    if (nt == 1) dslash_pn<T, P, true>(D, dst, src, mu, g5in, g5out, epi, b, aux, fin, N, st);
    else dslash_pn<T, P, false>(D, dst, src, mu, g5in, g5out, epi, b, aux, fin, N, st);
}

template <typename T>
void launch_b_dslash(const DiracView<T>& D, cplx<T>* dst, const cplx<T>* src, T mu, int g5in, int g5out, Epi epi,
                     const cplx<T>* b, const cplx<T>* aux, const FinB& fin, int N, hipStream_t st) {
    run_dispatch<T>(N, [&](auto pc) {
        dslash_p<T, decltype(pc)::value>(D, dst, src, mu, g5in, g5out, epi, b, aux, fin, N, st);
    });
    SMG_LAUNCH_CHECK();
}

// ----------------------------------------------------------------- fused D^dag D
// D^dag D x in one launch.  A workgroup takes a kDdTX x kDdTY tile of sites and R
// right-hand sides.  It forms the intermediate t = D x on the tile and its one-site ring
// in LDS, with the arithmetic of the two-launch form's first pass (k_b_dslash, g5in =
// g5out = 0, mu).  Then it forms g5 D(-mu) g5 t on the tile from LDS (the second pass).
// The intermediate never goes to memory: the launch reads x once (the ring's halo again
// from L2) and writes the result once, where the two-launch form writes and re-reads t.
// The ring is recomputed by the neighbouring tiles: 180 of 128 sites for 16 x 8 tiles.
// A workgroup keeps one group of R right-hand sides for the whole launch, so that its
// threads' dot-product partials belong to fixed right-hand sides.
constexpr int kDdTX = 16, kDdTY = 8, kDdRX = kDdTX + 2, kDdRY = kDdTY + 2, kDdRing = kDdRX * kDdRY;
template <typename T>
constexpr int ddagd_rhs() { return sizeof(T) == 4 ? 16 : 8; }

// The operands of a Chebyshev step (EPI_CHEB): d_new (dst) apart from everything it
// reads; x (aux) apart from d, r and the new r; the new r (fin.cheb_rout, required from a
// start) apart from d, r, x and d_new; r in place (b) only apart from d.  From a start d
// and r may be the same vector (b, or the residual).
static bool cheb_ptrs_ok(const void* dst, const void* src, const void* b, const void* aux, const FinB& f) {
    const void* ro = f.cheb_rout;
    if (!b || !aux || f.cheb_mode < 0 || f.cheb_mode > 2 || f.cheb_start < 0 || f.cheb_start > 2) return false;
    if (dst == src || dst == b || dst == aux || (ro && dst == ro)) return false;
    if (aux == src || aux == b || (ro && aux == ro)) return false;
    if (ro && (ro == src || ro == b)) return false;
    // x left alone (1) or caught up (2), the latter possibly with a term read from
    // cheb_xadd, which the step must not write
    if (f.cheb_xacc < 0 || f.cheb_xacc > 2) return false;
    const void* xa = f.cheb_xadd;
    if (xa && (f.cheb_xacc != 2 || xa == aux || xa == ro || (xa == dst && f.cheb_mode == 0) ||
               (xa == b && f.cheb_mode != 2)))
        return false;
    if (f.cheb_start == 0) return src != b;
    return ro != nullptr;
}

template <typename T, int EPI, int P, bool NT>
__global__ __launch_bounds__(kBlock) void k_b_ddagd(DiracView<T> D, cplx<T>* __restrict__ dst,
                                                    const cplx<T>* __restrict__ src, T mu,
                                                    const cplx<T>* __restrict__ bvec,
                                                    const cplx<T>* __restrict__ aux, FinB fin, int N) {
    constexpr int R = ddagd_rhs<T>(), RP = R / P;
    static_assert(kBlock % RP == 0, "a thread's runs must keep their right-hand sides");
    __shared__ cplx<T> tmp[kDdRing][2][R];
    const int L = D.L, V = L * L;
    const int G = N / R, ntx = L / kDdTX, ntile = ntx * (L / kDdTY);
    const int j0 = (int)(blockIdx.x % G) * R;
    double acc[3][P];
#pragma unroll
    for (int q = 0; q < P; q++) acc[0][q] = acc[1][q] = acc[2][q] = 0.0;
    for (int tile = blockIdx.x / G; tile < ntile; tile += gridDim.x / G) {
        const int x0 = (tile % ntx) * kDdTX, y0 = (tile / ntx) * kDdTY;
        // t = D(mu) x on the tile and its ring
        for (int i = threadIdx.x; i < kDdRing * RP; i += kBlock) {
            const int rs = i / RP, run = i % RP;
            const int x = (x0 - 1 + rs % kDdRX + L) % L, y = (y0 - 1 + rs / kDdRX + L) % L;
            const int s = x + L * y, j = j0 + run * P;
            const int sxp = (x + 1 == L) ? s - x : s + 1;
            const int sxm = (x == 0) ? s + L - 1 : s - 1;
            const int syp = (y + 1 == L) ? x : s + L;
            const int sym = (y == 0) ? s + V - L : s - L;
            const int ss[5] = {s, sxp, sxm, syp, sym};
            RhsRun<T, P> in[5][2];
#pragma unroll
            for (int n = 0; n < 5; n++)
#pragma unroll
                for (int sp = 0; sp < 2; sp++) in[n][sp].load(src + (2 * (size_t)ss[n] + sp) * N + j);
            const cplx<T> u0 = D.U[s], u0b = D.U[sxm], u1 = D.U[V + s], u1b = D.U[V + sym];
            const T C = D.C ? D.C[s] : (T)0;
#pragma unroll
            for (int q = 0; q < P; q++) {
                cplx<T> h0, h1;
                hop_b(D.r, u0, u0b, u1, u1b, in[1][0].v[q], in[1][1].v[q], in[2][0].v[q], in[2][1].v[q],
                      in[3][0].v[q], in[3][1].v[q], in[4][0].v[q], in[4][1].v[q], h0, h1);
                tmp[rs][0][run * P + q] = mk<T>(D.diag + C, mu) * in[0][0].v[q] + h0;
                tmp[rs][1][run * P + q] = mk<T>(D.diag - C, -mu) * in[0][1].v[q] + h1;
            }
        }
        __syncthreads();
        // out = g5 D(-mu) g5 t on the tile
        for (int i = threadIdx.x; i < kDdTX * kDdTY * RP; i += kBlock) {
            const int ts = i / RP, run = i % RP;
            const int lx = ts % kDdTX, ly = ts / kDdTX;
            const int x = x0 + lx, y = y0 + ly, s = x + L * y, j = j0 + run * P;
            const int sxm = (x == 0) ? s + L - 1 : s - 1;
            const int sym = (y == 0) ? s + V - L : s - L;
            const int c = (lx + 1) + (ly + 1) * kDdRX;
            const int cs[5] = {c, c + 1, c - 1, c + kDdRX, c - kDdRX};
            cplx<T> in[5][2][P];
#pragma unroll
            for (int n = 0; n < 5; n++)
#pragma unroll
                for (int sp = 0; sp < 2; sp++)
#pragma unroll
                    for (int q = 0; q < P; q++) in[n][sp][q] = tmp[cs[n]][sp][run * P + q];
            const cplx<T> u0 = D.U[s], u0b = D.U[sxm], u1 = D.U[V + s], u1b = D.U[V + sym];
            const T C = D.C ? D.C[s] : (T)0;
            const size_t i0 = (2 * (size_t)s) * N + j, i1 = i0 + N;
            RhsRun<T, P> out[2], bb[2], ax[2], dd[2];
            // EPI_CHEB (FinB): what the step reads, by its start and mode
            const int cst = EPI == EPI_CHEB ? fin.cheb_start : 0, cmo = EPI == EPI_CHEB ? fin.cheb_mode : 0;
            const int xac = EPI == EPI_CHEB ? fin.cheb_xacc : 0;
            const bool need_d = cmo != 1 || cst != 0, need_x = cst != 1 && xac != 1 && (cmo != 1 || cst == 2);
            const cplx<T>* const xad = static_cast<const cplx<T>*>(fin.cheb_xadd);
            RhsRun<T, P> xa[2];
            if (EPI == EPI_CHEB && xac == 2 && xad) {   // the catch-up term added to x
                xa[0].load(xad + i0);
                xa[1].load(xad + i1);
            }
            if (EPI == EPI_RESID || EPI == EPI_CHEB) {   // b, or the Chebyshev residual r
                bb[0].load(bvec + i0);
                bb[1].load(bvec + i1);
            }
            if (EPI == EPI_DOTS || (EPI == EPI_CHEB && need_x)) {   // aux, or the iterate x
                ax[0].load(aux + i0);
                ax[1].load(aux + i1);
            }
            if (EPI == EPI_CHEB && need_d) {   // the direction at this site
                dd[0].load(src + i0);
                dd[1].load(src + i1);
            }
#pragma unroll
            for (int q = 0; q < P; q++) {
                const T g1 = (T)-1;   // gamma5 on the input: lower spin negated
                cplx<T> h0, h1;
                hop_b(D.r, u0, u0b, u1, u1b, in[1][0][q], g1 * in[1][1][q], in[2][0][q], g1 * in[2][1][q],
                      in[3][0][q], g1 * in[3][1][q], in[4][0][q], g1 * in[4][1][q], h0, h1);
                cplx<T> o0 = mk<T>(D.diag + C, -mu) * in[0][0][q] + h0;
                cplx<T> o1 = mk<T>(D.diag - C, mu) * (g1 * in[0][1][q]) + h1;
                o1 = -o1;   // gamma5 on the output
                if (EPI == EPI_RESID) {
                    o0 = bb[0].v[q] - o0;
                    o1 = bb[1].v[q] - o1;
                }
                if (EPI == EPI_CHEB) {   // the new residual r - s A d
                    o0 = bb[0].v[q] - (T)fin.cheb_s * o0;
                    o1 = bb[1].v[q] - (T)fin.cheb_s * o1;
                }
                out[0].v[q] = o0;
                out[1].v[q] = o1;
                if (EPI == EPI_DOTS) {
                    const cplx<T> d0 = cdot(o0, ax[0].v[q]), d1 = cdot(o1, ax[1].v[q]);
                    acc[0][q] += (double)d0.re + (double)d1.re;
                    acc[1][q] += (double)d0.im + (double)d1.im;
                    acc[2][q] += (double)norm2(o0) + (double)norm2(o1);
                }
                if (EPI == EPI_CHEB && need_d) {   // d_new = c1 d + c2 r, x += d_new
#pragma unroll
                    for (int sp = 0; sp < 2; sp++) {
                        const cplx<T> dp = (T)fin.cheb_s * dd[sp].v[q];   // the previous direction
                        cplx<T> xp = cst == 1 ? dp : (cst == 2 ? ax[sp].v[q] + dp : ax[sp].v[q]);
                        if (xac == 2) {   // x was left alone by the step before: add its direction too
                            xp = xp + dp;
                            if (xad) xp = xp + (T)fin.cheb_xadd_s * xa[sp].v[q];
                        }
                        dd[sp].v[q] = (T)fin.cheb_c1 * dp + (T)fin.cheb_c2 * out[sp].v[q];
                        ax[sp].v[q] = cmo == 1 ? xp : xp + dd[sp].v[q];
                    }
                }
            }
            if constexpr (EPI == EPI_CHEB) {   // r, d and x; r only; x only (fin.cheb_mode)
                cplx<T>* const rw = fin.cheb_rout ? static_cast<cplx<T>*>(fin.cheb_rout) : const_cast<cplx<T>*>(bvec);
                cplx<T>* const xw = const_cast<cplx<T>*>(aux);
                if (cmo != 2) {
                    out[0].store(rw + i0);
                    out[1].store(rw + i1);
                }
                if (cmo == 0) {
                    dd[0].store(dst + i0);
                    dd[1].store(dst + i1);
                }
                if ((cmo != 1 || cst != 0) && xac != 1) {
                    ax[0].store(xw + i0);
                    ax[1].store(xw + i1);
                }
            } else if constexpr (NT) {
                out[0].store_nt(dst + i0);
                out[1].store_nt(dst + i1);
            } else {
                out[0].store(dst + i0);
                out[1].store(dst + i1);
            }
        }
        __syncthreads();   // the next tile overwrites t
    }
    if (EPI == EPI_DOTS) {
        // lanes with equal lane % RP hold the same right-hand sides: fold them, then lane
        // j < N takes right-hand side j from lane (j - j0) / P of its group (zero outside)
        const int lane = threadIdx.x & 63;
#pragma unroll
        for (int c = 0; c < 3; c++)
#pragma unroll
            for (int q = 0; q < P; q++)
                for (int off = RP; off < 64; off <<= 1) acc[c][q] += __shfl_xor(acc[c][q], off, 64);
        const int jr = lane - j0;
        const bool mine = lane < N && jr >= 0 && jr < R;
        double v[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            double x = 0.0;
#pragma unroll
            for (int q = 0; q < P; q++) {
                const double y = __shfl(acc[c][q], mine ? jr / P : 0, 64);
                if (mine && jr % P == q) x = y;
            }
            v[c] = x;
        }
        reduce_finish_b<3, 64>(v, fin, N, gridDim.x, blockIdx.x);
    }
}

template <typename T, int P, bool NT>
static void ddagd_pn(const DiracView<T>& D, cplx<T>* dst, const cplx<T>* src, T mu, Epi epi, const cplx<T>* b,
                     const cplx<T>* aux, const FinB& fin, int N, int cap, hipStream_t st) {
    const int G = N / ddagd_rhs<T>();
    const int units = (D.L / kDdTX) * (D.L / kDdTY) * G;
    const int g = std::max(G, std::min(units, cap) / G * G);   // a multiple of G
    if (epi == EPI_NONE)
        k_b_ddagd<T, EPI_NONE, P, NT><<<g, kBlock, 0, st>>>(D, dst, src, mu, b, aux, fin, N);
    else if (epi == EPI_RESID)
        k_b_ddagd<T, EPI_RESID, P, NT><<<g, kBlock, 0, st>>>(D, dst, src, mu, b, aux, fin, N);
    else if (epi == EPI_CHEB)
        k_b_ddagd<T, EPI_CHEB, P, NT><<<g, kBlock, 0, st>>>(D, dst, src, mu, b, aux, fin, N);
    else
        k_b_ddagd<T, EPI_DOTS, P, NT><<<g, kBlock, 0, st>>>(D, dst, src, mu, b, aux, fin, N);
}

bool b_ddagd_fits(int L, int N, size_t real_bytes) {
    const int R = real_bytes == 4 ? ddagd_rhs<float>() : ddagd_rhs<double>();
    return L % kDdTX == 0 && L % kDdTY == 0 && N >= R && N % R == 0;
}

// The tuner's TUNE_DDAGD choice for this shape: 0 the two launches, 1 + 2 g + nt the
// fused launch with kTuneGrids[g] workgroups (reducing launches; the others take one
// workgroup per tile and group) and non-temporal stores when nt.  Without an entry: the
// fused launch, ordinary stores, kRedGrid workgroups with dot products.  Returns false
// when the caller runs the two launches.
template <typename T>
bool launch_b_ddagd(const DiracView<T>& D, cplx<T>* dst, const cplx<T>* src, T mu, Epi epi, const cplx<T>* b,
                    const cplx<T>* aux, const FinB& fin, int N, hipStream_t st) {
    SMG_REQUIRE(epi == EPI_NONE || epi == EPI_RESID || epi == EPI_DOTS || epi == EPI_CHEB, "fused D^dag D: epilogue");
    SMG_REQUIRE(epi != EPI_CHEB || cheb_ptrs_ok(dst, src, b, aux, fin),
                "fused D^dag D: the Chebyshev epilogue's r, x, d, d_new (and new r) must not overlap as given");
    if (!b_ddagd_fits(D.L, N, sizeof(T))) return false;
    const int choice = tune_lookup(TuneKey{TUNE_DDAGD, D.L, N, (int)sizeof(T), (int)epi, 0});
    if (choice == 0 && epi != EPI_CHEB) return false;   // the Chebyshev step has no two-launch form
    const bool nt = choice > 0 && ((choice - 1) & 1);
    const int cap = epi == EPI_DOTS ? (choice > 0 ? kTuneGrids[(choice - 1) >> 1] : kRedGrid) : 1 << 20;
    constexpr int P = sizeof(T) == 4 ? 2 : 1;
    if (nt) ddagd_pn<T, P, true>(D, dst, src, mu, epi, b, aux, fin, N, cap, st);
    else ddagd_pn<T, P, false>(D, dst, src, mu, epi, b, aux, fin, N, cap, st);
    SMG_LAUNCH_CHECK();
    return true;
}
template bool launch_b_ddagd<double>(const DiracView<double>&, c64*, const c64*, double, Epi, const c64*, const c64*,
                                     const FinB&, int, hipStream_t);
template bool launch_b_ddagd<float>(const DiracView<float>&, c32*, const c32*, float, Epi, const c32*, const c32*,
                                    const FinB&, int, hipStream_t);

// ----------------------------------------------------------------- coarse operator
// link i of a coarse view in arithmetic type T
template <typename T, typename TL>
__device__ __forceinline__ cplx<T> link_at(const CoarseView<TL>& C, size_t i) {
    return cvt<T>(load_c(C.link + i));
}
template <typename T, typename TL, int EPI>
__global__ __launch_bounds__(kBlock) void k_b_coarse(CoarseView<TL> C, cplx<T>* __restrict__ dst,
                                                     const cplx<T>* __restrict__ src,
                                                     const cplx<T>* __restrict__ bvec,
                                                     const cplx<T>* __restrict__ aux, FinB fin, int N) {
    if (EPI == EPI_DOTS && fin.skip && *fin.skip) return;   // CG finished for every RHS
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
            const size_t lrow = (((size_t)b * 9 + d) * k + ki) * k;
            const cplx<T>* sv = src + ((size_t)nbr[d] * k) * N + j;
            cplx<T> sum = {(T)0, (T)0};
            for (int kj = 0; kj < k; kj++) sum = sum + link_at<T>(C, lrow + kj) * sv[(size_t)kj * N];
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
// ----------------------------------------------------------------- MFMA coarse operator
// The complex products of the coarse stencil and of the dense coarsest inverse
// run on the matrix cores as real GEMMs: a complex (m x k) x (k x N) product is
//   [[Re A, -Im A], [Im A, Re A]] (2m x 2k, rows (i, c) and columns (kk, c')
//   interleaved) x (Re X, Im X) (2k x N, rows (kk, c')),
// whose output rows (i, c) are (Re, Im) of the complex result.  Tiles are the
// 16x16x4 MFMA shapes in FP64 (FP64 path) or FP32 (FP32 path, coarse links
// stored in FP16/FP32 and widened on load).  Lane layouts (checked on gfx950):
// A[i][kk] in lane i + 16 kk and B[kk][j] in lane j + 16 kk for both; the
// result D[i][j] sits in lane j + 16 (i / 4), register i % 4 for FP32, and in
// lane j + 16 (i % 4), register i / 4 for FP64.  The tile row i <-> (complex row
// ki, part c) assignment is chosen per precision so that lane (j, g) always
// holds (Re, Im) of rows ki = 2g and 2g + 1 in registers (0, 1) and (2, 3).
template <typename T>
struct MfmaTile;
template <>
struct MfmaTile<double> {
    typedef double acc __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc mma(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // tile row i = g + 4v  <->  ki = 2g + v/2, c = v%2
    static __device__ __forceinline__ int row_ki(int i) { return 2 * (i & 3) + (i >> 3); }
    static __device__ __forceinline__ int row_c(int i) { return (i >> 2) & 1; }
};
template <>
struct MfmaTile<float> {
    typedef float acc __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc mma(float a, float b, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    // tile row i = 4g + v  <->  ki = 2g + v/2, c = v%2
    static __device__ __forceinline__ int row_ki(int i) { return i >> 1; }
    static __device__ __forceinline__ int row_c(int i) { return i & 1; }
};

// A-fragment entry of the real form of complex a for row part c and column part cp
template <typename T>
__device__ __forceinline__ T real_form(cplx<T> a, int c, int cp) {
    return c == cp ? a.re : (c ? a.im : -a.im);
}

// Q consecutive coarse links loaded as one (or a few) wide loads
template <typename TL, int Q>
struct alignas(sizeof(TL) * Q > 16 ? 16 : sizeof(TL) * Q) LinkRun {
    TL v[Q];
};

// One wave per coarse site b: out'[b] (2K x N) = sum_d A'_{b,d} (2K x 2K) B'_{nbr(b,d)} (2K x N).
// The real-form K index of MFMA (q, c') in lane group hi is (kj = hi*Q + q, part c'),
// Q = K/4, so a lane's Q links of one direction are contiguous (one wide load) and
// each source entry it loads feeds both parts c'.  In FP32 with an even number of
// 16-RHS tiles, tile t of lane column lo holds RHS 32(t/2) + 2 lo + t%2, so one
// 16-byte load gives a lane the sources of two tiles; otherwise RHS 16t + lo.
// Per site and direction: RT*Q*2*NT MFMAs (RT = K/8 row tiles).
template <typename T, typename TL, int EPI, int K, int NT>
__global__ __launch_bounds__(kBlock) void k_b_coarse_mfma(CoarseView<TL> C, cplx<T>* __restrict__ dst,
                                                          const cplx<T>* __restrict__ src,
                                                          const cplx<T>* __restrict__ bvec,
                                                          const cplx<T>* __restrict__ aux, FinB fin) {
    if (EPI == EPI_DOTS && fin.skip && *fin.skip) return;   // CG finished for every RHS
    constexpr int N = 16 * NT, RT = K / 8, Q = K / 4;
    constexpr bool PAIR = sizeof(T) == 4 && NT % 2 == 0;
    constexpr int NL = PAIR ? NT / 2 : NT;   // source loads per (kj) per lane
    constexpr int S = (sizeof(T) == 4 && K <= 8) ? 3 : 1;   // directions per pipeline stage (K <= 8)
    using MT = MfmaTile<T>;
    using acc_t = typename MT::acc;
    using Run = LinkRun<TL, Q>;
    const int lane = threadIdx.x & 63, lo = lane & 15, hi = lane >> 4;
    const int nx = C.nx, nsites = C.nx * C.ny;
    const int nw = kBlock / 64;
    const int a_ki = MT::row_ki(lo), a_c = MT::row_c(lo);
    double dacc[NT][3];
#pragma unroll
    for (int t = 0; t < NT; t++) dacc[t][0] = dacc[t][1] = dacc[t][2] = 0.0;
    for (int b = blockIdx.x * nw + (threadIdx.x >> 6); b < nsites; b += gridDim.x * nw) {
        const int bx = b % nx, by = b / nx;
        const int xm = (bx == 0 ? nx - 1 : bx - 1), xp = (bx + 1 == nx ? 0 : bx + 1);
        const int ym = (by == 0 ? C.ny - 1 : by - 1), yp = (by + 1 == C.ny ? 0 : by + 1);
        acc_t acc[RT][NT];
#pragma unroll
        for (int r = 0; r < RT; r++)
#pragma unroll
            for (int t = 0; t < NT; t++) acc[r][t] = acc_t{0, 0, 0, 0};
        // software pipeline over the 9 directions in stages of S: the loads of stage
        // i + 1 are issued before the MFMAs of stage i (sched_barrier keeps the order)
        int nbr[9];
#pragma unroll
        for (int d = 0; d < 9; d++) {
            const int cx = (d == 1 || d == 5 || d == 6) ? xp : (d == 2 || d == 7 || d == 8) ? xm : bx;
            const int cy = (d == 3 || d == 5 || d == 7) ? yp : (d == 4 || d == 6 || d == 8) ? ym : by;
            nbr[d] = cx + cy * nx;
        }
        constexpr int SB = K <= 8 ? S : 1;   // staged directions (K <= 8 only)
        Run lv[2][SB][K <= 8 ? RT : 1];
        cplx<T> sv[2][SB][K <= 8 ? Q : 1][NT];   // sv[.][.][q][t] = src[nb][hi*Q + q][rhs(t, lo)]
        auto load_stage = [&](int buf, int st0) {
#pragma unroll
            for (int e = 0; e < S; e++) {
                const int d = st0 + e, nb = nbr[d];
#pragma unroll
                for (int r = 0; r < RT; r++)
                    lv[buf][e][r] = *reinterpret_cast<const Run*>(C.link + ((size_t)b * 9 + d) * (K * K) +
                                                                  (8 * r + a_ki) * K + hi * Q);
#pragma unroll
                for (int q = 0; q < Q; q++) {
                    const cplx<T>* row = src + ((size_t)nb * K + hi * Q + q) * N;
#pragma unroll
                    for (int u = 0; u < NL; u++) {
                        if constexpr (PAIR) {
                            const float4 w = *reinterpret_cast<const float4*>(row + 32 * u + 2 * lo);
                            sv[buf][e][q][2 * u] = cplx<T>{w.x, w.y};
                            sv[buf][e][q][2 * u + 1] = cplx<T>{w.z, w.w};
                        } else {
                            sv[buf][e][q][u] = row[16 * u + lo];
                        }
                    }
                }
            }
        };
        auto mma_stage = [&](int buf) {
#pragma unroll
            for (int e = 0; e < S; e++)
#pragma unroll
                for (int r = 0; r < RT; r++)
#pragma unroll
                    for (int q = 0; q < Q; q++) {
                        const cplx<T> L = cvt<T>(load_c(&lv[buf][e][r].v[q]));
                        const T a0 = real_form(L, a_c, 0), a1 = real_form(L, a_c, 1);
#pragma unroll
                        for (int t = 0; t < NT; t++) {
                            acc[r][t] = MT::mma(a0, sv[buf][e][q][t].re, acc[r][t]);
                            acc[r][t] = MT::mma(a1, sv[buf][e][q][t].im, acc[r][t]);
                        }
                    }
        };
        if constexpr (K <= 8) {
            load_stage(0, 0);
#pragma unroll
            for (int i = 0; i < 9 / S; i++) {
                if (i + 1 < 9 / S) load_stage((i + 1) & 1, (i + 1) * S);
                __builtin_amdgcn_sched_barrier(0);
                mma_stage(i & 1);
            }
        } else {
            // large k: a direction carries RT*Q*2*NT MFMAs, enough to cover its own
            // loads; the sources of the direction are held, the links streamed per
            // row tile
#pragma unroll 1
            for (int d = 0; d < 9; d++) {
                const int nb = nbr[d];
                cplx<T> sw[Q][NT];
#pragma unroll
                for (int q = 0; q < Q; q++) {
                    const cplx<T>* row = src + ((size_t)nb * K + hi * Q + q) * N;
#pragma unroll
                    for (int u = 0; u < NL; u++) {
                        if constexpr (PAIR) {
                            const float4 w = *reinterpret_cast<const float4*>(row + 32 * u + 2 * lo);
                            sw[q][2 * u] = cplx<T>{w.x, w.y};
                            sw[q][2 * u + 1] = cplx<T>{w.z, w.w};
                        } else {
                            sw[q][u] = row[16 * u + lo];
                        }
                    }
                }
                const TL* lrow = C.link + ((size_t)b * 9 + d) * (K * K) + a_ki * K + hi * Q;
#pragma unroll
                for (int r = 0; r < RT; r++) {
                    const Run lr = *reinterpret_cast<const Run*>(lrow + 8 * r * K);
#pragma unroll
                    for (int q = 0; q < Q; q++) {
                        const cplx<T> L = cvt<T>(load_c(&lr.v[q]));
                        const T a0 = real_form(L, a_c, 0), a1 = real_form(L, a_c, 1);
#pragma unroll
                        for (int t = 0; t < NT; t++) {
                            acc[r][t] = MT::mma(a0, sw[q][t].re, acc[r][t]);
                            acc[r][t] = MT::mma(a1, sw[q][t].im, acc[r][t]);
                        }
                    }
                }
            }
        }
#pragma unroll
        for (int r = 0; r < RT; r++)
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int ki = 8 * r + 2 * hi + h;
                const size_t row = ((size_t)b * K + ki) * N;
#pragma unroll
                for (int u = 0; u < NL; u++) {
                    cplx<T> o[2];
                    const int nt = PAIR ? 2 : 1;
                    const size_t idx = row + (PAIR ? 32 * u + 2 * lo : 16 * u + lo);
#pragma unroll
                    for (int e = 0; e < nt; e++) {
                        const int t = PAIR ? 2 * u + e : u;
                        o[e] = cplx<T>{acc[r][t][2 * h], acc[r][t][2 * h + 1]};
                    }
                    if constexpr (PAIR) {
                        if (EPI == EPI_RESID) {
                            const float4 w = *reinterpret_cast<const float4*>(bvec + idx);
                            o[0] = cplx<T>{w.x, w.y} - o[0];
                            o[1] = cplx<T>{w.z, w.w} - o[1];
                        }
                        *reinterpret_cast<float4*>(dst + idx) = float4{o[0].re, o[0].im, o[1].re, o[1].im};
                        if (EPI == EPI_DOTS) {
                            const float4 w = *reinterpret_cast<const float4*>(aux + idx);
                            const cplx<T> x0{w.x, w.y}, x1{w.z, w.w};
                            const cplx<T> p0 = cdot(o[0], x0), p1 = cdot(o[1], x1);
                            dacc[2 * u][0] += p0.re;
                            dacc[2 * u][1] += p0.im;
                            dacc[2 * u][2] += norm2(o[0]);
                            dacc[2 * u + 1][0] += p1.re;
                            dacc[2 * u + 1][1] += p1.im;
                            dacc[2 * u + 1][2] += norm2(o[1]);
                        }
                    } else {
                        if (EPI == EPI_RESID) o[0] = bvec[idx] - o[0];
                        dst[idx] = o[0];
                        if (EPI == EPI_DOTS) {
                            const cplx<T> p = cdot(o[0], aux[idx]);
                            dacc[u][0] += p.re;
                            dacc[u][1] += p.im;
                            dacc[u][2] += norm2(o[0]);
                        }
                    }
                }
            }
    }
    if (EPI == EPI_DOTS) {
        // lanes with equal lo hold partials of RHS rhs(t, lo): fold the four lane
        // groups, then lane j < N fetches RHS j from lane lo(j), tile t(j)
#pragma unroll
        for (int t = 0; t < NT; t++)
#pragma unroll
            for (int c = 0; c < 3; c++) {
                dacc[t][c] += __shfl_xor(dacc[t][c], 16, 64);
                dacc[t][c] += __shfl_xor(dacc[t][c], 32, 64);
            }
        int tj, lj;
        if constexpr (PAIR) {
            tj = 2 * ((lane & 63) >> 5) + (lane & 1);
            lj = (lane & 31) >> 1;
        } else {
            tj = (lane >> 4) & (NT - 1);
            lj = lane & 15;
        }
        double v[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int t = 0; t < NT; t++)
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const double x = __shfl(dacc[t][c], lj, 64);
                if (t == tj && lane < N) v[c] = x;
            }
        reduce_finish_b<3, 64>(v, fin, N, gridDim.x, blockIdx.x);
    }
}

// ----------------------------------------------------------------- bf16x3 coarse operator
// FP32 coarse apply with FP16 links on the bf16 matrix cores (16x16x32, FP32
// accumulation).  Every FP32 operand x is split as x = hi + lo with hi = bf16(x)
// and lo = bf16(x - hi); the product keeps hi*hi + hi*lo + lo*hi, dropping
// lo*lo (relative size 2^-16).  FP16 links are exact as hi + lo (11 significant
// bits), so only the sources are truncated, to about 2^-17 relative.  bf16 has
// FP32's exponent range, so no scaling is needed.  Three bf16 MFMAs cost about
// 3/16 of the FP32 MFMAs they replace.
//
// The 9 directions x K link columns of a coarse site form one K axis of 9K complex
// columns (d, kj), taken 16 at a time per MFMA: lane group hi holds the four
// columns 16m + 4hi + (0..3) (one direction, four consecutive kj), each as the
// real pair (Re, Im) -> 8 bf16 operands.  Lane layout of v_mfma_f32_16x16x32_bf16
// (checked on gfx950): A[i][k] in lane i + 16 (k/8), element k%8; B[k][j] in lane
// j + 16 (k/8), element k%8; D as the FP32 16x16x4 tile.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Split-precision operands by bit manipulation: hi is x with the low 16 bits of
// its FP32 pattern cleared (a bf16, truncated), lo = x - hi (exact in FP32, its
// upper 16 bits are the bf16 carried).  pack(a, b) places the bf16 upper halves
// of a and b in the low / high half of a dword: v_perm_b32 with bytes (7,6,3,2)
// of [b:a]; with the selector 0x03020706 the halves are swapped.
__device__ __forceinline__ unsigned hi16(float x) { return __float_as_uint(x) & 0xffff0000u; }
__device__ __forceinline__ float lo_part(float x) { return x - __uint_as_float(hi16(x)); }
__device__ __forceinline__ unsigned pack_bf(float a, float b, unsigned sel) {
    return __builtin_amdgcn_perm(__float_as_uint(b), __float_as_uint(a), sel);
}

// BF16 storage of a complex vector entry: one dword, the bf16 of Re in the low
// half and of Im in the high half (the layout of a B-operand pair of the MFMA),
// rounded to nearest even.
__device__ __forceinline__ unsigned bf16_rne(float x) {
    const unsigned u = __float_as_uint(x);
    return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
}
__device__ __forceinline__ unsigned pack_h(c32 v) { return bf16_rne(v.re) | (bf16_rne(v.im) << 16); }
__device__ __forceinline__ c32 unpack_h(unsigned w) {
    return c32{__uint_as_float(w << 16), __uint_as_float(w & 0xffff0000u)};
}

// One wave per coarse site; the 9K/16 chunks are pipelined, chunks m + 1 and m + 2
// in flight while chunk m is split and multiplied.  The complex products use the
// links in their stored order: A[i][k] of row tile rt holds link row ki = 16 rt + i
// and columns k = (kj, c') as (Re L, Im L), so every link entry is split into bf16
// hi + lo once, by one lane, and a lane's four columns of a row are one 16-byte load.
// The complex sign goes into the source operand, in two forms per right-hand side:
// (Re x, -Im x) gives Re(L x) and (Im x, Re x) gives Im(L x), so every tile of 16
// right-hand sides takes two MFMA column tiles (part 0: real, part 1: imaginary).
// A lane of the result holds rows 16 rt + 4 hi + v (v = 0..3) of one right-hand side,
// real part in part 0, imaginary part in part 1.  K = 8 fills rows 0..7 of the tile.
// HV: source, destination and the dot partner are BF16 vectors (cb16 entries), so the
// source needs no split (two MFMAs per product instead of three); EPI_RESID is FP32
// only.  HERM (C.herm set, K >= 16 so that a chunk lies in one direction): the blocks
// of the backward directions are read as the conjugate transposes of the neighbours'
// forward blocks, L[b][d] = L[nbr(b,d)][rev d]^dag, so a site reads 5 of its 9 link
// blocks and the links' HBM traffic falls to 5/9; a lane then gathers its four link
// columns of a row as four dwords of one column of the stored block.
template <typename TL, int EPI, int K, int NT, bool HV = false, bool HERM = false>
__global__ __launch_bounds__(kBlock)
void k_b_coarse_bf16(CoarseView<TL> C, void* __restrict__ dstv, const void* __restrict__ srcv,
                     const void* __restrict__ bv, const void* __restrict__ auxv, FinB fin) {
    static_assert(std::is_same<TL, ch16>::value, "bf16x3 path: FP16 links");
    static_assert(!HV || EPI != EPI_RESID, "BF16 vectors: no residual epilogue");
    static_assert(HV || EPI != EPI_CG, "the CG epilogue: BF16 vectors");
    static_assert(!HERM || K % 16 == 0, "Hermitian reads: a chunk of 16 columns inside one direction");
    static_assert(K == 8 || K == 16 || K == 32, "k in {8, 16, 32}");
    c32* __restrict__ dst = static_cast<c32*>(dstv);
    const c32* __restrict__ src = static_cast<const c32*>(srcv);
    const c32* __restrict__ aux = static_cast<const c32*>(auxv);
    unsigned* __restrict__ dsth = static_cast<unsigned*>(dstv);
    const unsigned* __restrict__ srch = static_cast<const unsigned*>(srcv);
    const unsigned* __restrict__ auxh = static_cast<const unsigned*>(auxv);
    const c32* __restrict__ bvec = static_cast<const c32*>(bv);
    const unsigned* __restrict__ bh2 = static_cast<const unsigned*>(bv);   // EPI_CG, EPI_CHEB: r
    // EPI_CHEB updates r (bv) and x (auxv) in place at this wave's own site only
    unsigned* const rwh = const_cast<unsigned*>(bh2);
    unsigned* const xwh = const_cast<unsigned*>(auxh);
    // FP32 vectors: the new residual to fin.cheb_rout where given (FinB's cheb_start)
    c32* const rwf = fin.cheb_rout ? static_cast<c32*>(fin.cheb_rout) : const_cast<c32*>(bvec);
    c32* const xwf = const_cast<c32*>(aux);
    const float cc1 = (float)fin.cheb_c1, cc2 = (float)fin.cheb_c2, ccs = (float)fin.cheb_s;
    const int cmode = fin.cheb_mode;   // EPI_CHEB: 0 r, d and x; 1 r only; 2 x only
    const int cst = HV ? 0 : fin.cheb_start;
    const int hxacc = HV ? fin.cheb_xacc : 0;   // BF16 Chebyshev: FinB's cheb_xacc
    const int fxac = HV ? 0 : fin.cheb_xacc;   // FP32 vectors: FinB's cheb_xacc and cheb_xadd
    const bool need_d = cmode != 1 || cst != 0, need_x = cst != 1 && fxac != 1 && (cmode != 1 || cst == 2);
    const c32* const fxad = static_cast<const c32*>(fin.cheb_xadd);
    const float fxas = (float)fin.cheb_xadd_s;
    constexpr bool RED = EPI == EPI_DOTS || EPI == EPI_CG;
    constexpr int NC = EPI == EPI_CG ? 5 : 3;   // reduced values per RHS
    if (RED && fin.skip && *fin.skip) {   // CG finished for every RHS
        if (fin.late && blockIdx.x == 0 && threadIdx.x == 0) *fin.late = 1;
        return;
    }
    constexpr int N = 16 * NT, RT = (K + 15) / 16, NCH = (9 * K + 15) / 16;
    constexpr bool PAIR = NT % 2 == 0;
    constexpr int NL = PAIR ? NT / 2 : NT;
    constexpr unsigned kSel = 0x07060302u;   // pack_bf: bf16(a) low half, bf16(b) high half
    const int lane = threadIdx.x & 63, lo = lane & 15, hi = lane >> 4;
    const int nx = C.nx, ny = C.ny, nsites = nx * ny;
    const int nw = kBlock / 64;
    // K = 8: rows 8..15 of the tile do not exist (zero links, no result)
    const bool row_ok = K >= 16 || lo < K;        // this lane's link row
    const bool out_ok = K >= 16 || 4 * hi < K;    // this lane's result rows
    const unsigned* links = reinterpret_cast<const unsigned*>(C.link);
    double dacc[NT][NC];
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
        for (int c = 0; c < NC; c++) dacc[t][c] = 0.0;
    for (int b = blockIdx.x * nw + (threadIdx.x >> 6); b < nsites; b += gridDim.x * nw) {
        const int bx = b % nx, by = b / nx;
        f32x4 acc[RT][NT][2];
#pragma unroll
        for (int r = 0; r < RT; r++)
#pragma unroll
            for (int t = 0; t < NT; t++) acc[r][t][0] = acc[r][t][1] = f32x4{0, 0, 0, 0};
        // raw operands of one chunk: the sources of four link columns (both RHS
        // tiles of a pair in one 16-byte load) and the link rows of the row tiles
        using SrcF = typename std::conditional<PAIR, float4, float2>::type;
        using SrcH = typename std::conditional<PAIR, uint2, unsigned>::type;
        using SrcV = typename std::conditional<HV, SrcH, SrcF>::type;
        struct Raw {
            SrcV s[4][NL];
            uint4 l[RT];
        };
        auto load_raw = [&](Raw& R, int m) {
            const int cc0 = 16 * m + 4 * hi;
            const bool valid = (9 * K) % 16 == 0 || cc0 < 9 * K;
            const int cc = valid ? cc0 : 0;
            const int d = cc / K, kj0 = cc % K;
            const int dx = (d == 1 || d == 5 || d == 6) ? 1 : ((d == 2 || d == 7 || d == 8) ? -1 : 0);
            const int dy = (d == 3 || d == 5 || d == 7) ? 1 : ((d == 4 || d == 6 || d == 8) ? -1 : 0);
            const int nb = (bx + dx + nx) % nx + ((by + dy + ny) % ny) * nx;
#pragma unroll
            for (int e2 = 0; e2 < 4; e2++) {
                const size_t row = ((size_t)nb * K + kj0 + e2) * N;
#pragma unroll
                for (int u = 0; u < NL; u++) {
                    const size_t o = row + (PAIR ? 32 * u + 2 * lo : 16 * u + lo);
                    if constexpr (HV) R.s[e2][u] = *reinterpret_cast<const SrcV*>(srch + o);
                    else R.s[e2][u] = *reinterpret_cast<const SrcV*>(src + o);
                }
            }
            const bool bwd = HERM && (d == 2 || d == 4 || d == 7 || d == 8);
#pragma unroll
            for (int r = 0; r < RT; r++) {
                const int ki = 16 * r + lo;
                if (!row_ok) {
                    R.l[r] = uint4{0u, 0u, 0u, 0u};
                } else if (bwd) {   // L[b][d][ki][kj] = conj(L[nb][rev d][kj][ki]): Im's sign bit flipped
                    const int rd = d == 2 ? 1 : d == 4 ? 3 : d == 7 ? 6 : 5;
                    const unsigned* lcol = links + ((size_t)nb * 9 + rd) * (K * K) + (size_t)kj0 * K + ki;
                    R.l[r] = uint4{lcol[0] ^ 0x80000000u, lcol[K] ^ 0x80000000u, lcol[2 * K] ^ 0x80000000u,
                                   lcol[3 * K] ^ 0x80000000u};
                } else {
                    R.l[r] = *reinterpret_cast<const uint4*>(links + ((size_t)b * 9 + d) * (K * K) + ki * K + kj0);
                }
            }
            if (!valid) {   // the partial last chunk (k = 8): these columns do not exist
#pragma unroll
                for (int e2 = 0; e2 < 4; e2++)
#pragma unroll
                    for (int u = 0; u < NL; u++) R.s[e2][u] = SrcV{};
#pragma unroll
                for (int r = 0; r < RT; r++) R.l[r] = uint4{0u, 0u, 0u, 0u};
            }
        };
        auto mma_chunk = [&](const Raw& R) {
            // link operands of every row tile: (Re L, Im L) as bf16 hi and lo
            u32x4 ah[RT], al[RT];
#pragma unroll
            for (int r = 0; r < RT; r++) {
                const unsigned lws[4] = {R.l[r].x, R.l[r].y, R.l[r].z, R.l[r].w};
#pragma unroll
                for (int e2 = 0; e2 < 4; e2++) {
                    const float re = __half2float(__ushort_as_half((unsigned short)(lws[e2] & 0xffffu)));
                    const float im = __half2float(__ushort_as_half((unsigned short)(lws[e2] >> 16)));
                    ah[r][e2] = pack_bf(re, im, kSel);
                    al[r][e2] = pack_bf(lo_part(re), lo_part(im), kSel);
                }
            }
            // one RHS tile at a time: its two source forms, then its products
#pragma unroll
            for (int t = 0; t < NT; t++) {
                const int u = PAIR ? t / 2 : t;
                u32x4 bh[2], bl[2];
#pragma unroll
                for (int e2 = 0; e2 < 4; e2++) {
                    if constexpr (HV) {   // the stored dword is (Re, Im) in bf16
                        unsigned w;
                        if constexpr (PAIR) w = (t & 1) ? R.s[e2][u].y : R.s[e2][u].x;
                        else w = R.s[e2][u];
                        bh[0][e2] = w ^ 0x80000000u;                      // (Re x, -Im x)
                        bh[1][e2] = __builtin_amdgcn_alignbit(w, w, 16);  // (Im x, Re x)
                    } else {
                        c32 sv;
                        if constexpr (PAIR) {
                            const float4 w = R.s[e2][u];
                            sv = (t & 1) ? c32{w.z, w.w} : c32{w.x, w.y};
                        } else {
                            const float2 w = R.s[e2][u];
                            sv = c32{w.x, w.y};
                        }
                        const float lr = lo_part(sv.re), li = lo_part(sv.im);
                        bh[0][e2] = pack_bf(sv.re, sv.im, kSel) ^ 0x80000000u;
                        bh[1][e2] = pack_bf(sv.im, sv.re, kSel);
                        bl[0][e2] = pack_bf(lr, li, kSel) ^ 0x80000000u;
                        bl[1][e2] = pack_bf(li, lr, kSel);
                    }
                }
#pragma unroll
                for (int r = 0; r < RT; r++) {
                    const bf16x8 ahv = __builtin_bit_cast(bf16x8, ah[r]), alv = __builtin_bit_cast(bf16x8, al[r]);
#pragma unroll
                    for (int p = 0; p < 2; p++) {
                        const bf16x8 bhv = __builtin_bit_cast(bf16x8, bh[p]);
                        acc[r][t][p] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahv, bhv, acc[r][t][p], 0, 0, 0);
                        if constexpr (!HV) {
                            const bf16x8 blv = __builtin_bit_cast(bf16x8, bl[p]);
                            acc[r][t][p] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahv, blv, acc[r][t][p], 0, 0, 0);
                        }
                        acc[r][t][p] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alv, bhv, acc[r][t][p], 0, 0, 0);
                    }
                }
            }
        };
        // the epilogue's operands (b for EPI_RESID, aux for EPI_DOTS, aux and b for
        // EPI_CG) at this lane's results, loaded while the last chunks are multiplied
        // (PAIR layout)
        using EpoV = typename std::conditional<HV, uint2, float4>::type;
        constexpr int NE = EPI == EPI_CG ? 2 : (EPI == EPI_CHEB ? 3 : 1);
        // the FP32-vector Chebyshev step loads its three operands in the epilogue itself:
        // held from the last chunks on, they would cost the kernel its second wave per SIMD
        // (with BF16 vectors the prefetch pays: 0.8 ms per 32-RHS batch at L=512)
        constexpr bool PRE = !(EPI == EPI_CHEB && !HV);
        EpoV epo[PAIR && PRE ? RT : 1][PAIR && PRE ? 4 : 1][PAIR && PRE ? NL : 1][NE];
        auto load_epo = [&]() {
            if constexpr (PAIR && EPI != EPI_NONE && PRE) {
                if (out_ok) {
#pragma unroll
                    for (int r = 0; r < RT; r++)
#pragma unroll
                        for (int v = 0; v < 4; v++)
#pragma unroll
                            for (int u = 0; u < NL; u++) {
                                const size_t o = ((size_t)b * K + 16 * r + 4 * hi + v) * N + 32 * u + 2 * lo;
                                if constexpr (HV) {
                                    if (EPI != EPI_CHEB || hxacc != 1)
                                        epo[r][v][u][0] = *reinterpret_cast<const uint2*>(auxh + o);
                                    if constexpr (EPI == EPI_CG || EPI == EPI_CHEB)
                                        epo[r][v][u][1] = *reinterpret_cast<const uint2*>(bh2 + o);
                                    if constexpr (EPI == EPI_CHEB)   // the previous direction at this site
                                        epo[r][v][u][2] = *reinterpret_cast<const uint2*>(srch + o);
                                } else {
                                    if (EPI != EPI_CHEB || need_x)
                                        epo[r][v][u][0] = *reinterpret_cast<const float4*>((EPI == EPI_RESID ? bvec : aux) + o);
                                    if constexpr (EPI == EPI_CHEB) {
                                        epo[r][v][u][1] = *reinterpret_cast<const float4*>(bvec + o);
                                        if (need_d) epo[r][v][u][2] = *reinterpret_cast<const float4*>(src + o);
                                    }
                                }
                            }
                }
            }
        };
        if constexpr (NT <= 2) {
            // chunks m + 1 and m + 2 in flight while chunk m is multiplied (fully
            // unrolled, so the ring indices are compile-time)
            Raw raw[3];
            load_raw(raw[0], 0);
            if (NCH > 1) load_raw(raw[1], 1);
            if (NCH == 1) load_epo();
#pragma unroll
            for (int m = 0; m < NCH; m++) {
                if (m + 2 < NCH) load_raw(raw[(m + 2) % 3], m + 2);
                else if (m + 2 == NCH) load_epo();   // the ring slot of chunk m + 2 is free
                __builtin_amdgcn_sched_barrier(0);
                mma_chunk(raw[m % 3]);
            }
        } else {   // N = 64: registers allow one chunk ahead
            Raw raw[2];
            load_raw(raw[0], 0);
#pragma unroll 2
            for (int m = 0; m < NCH; m++) {
                if (m + 1 < NCH) load_raw(raw[(m + 1) & 1], m + 1);
                __builtin_amdgcn_sched_barrier(0);
                mma_chunk(raw[m & 1]);
            }
            load_epo();
        }
        if (out_ok) {
#pragma unroll
            for (int r = 0; r < RT; r++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int ki = 16 * r + 4 * hi + v;
                    const size_t row = ((size_t)b * K + ki) * N;
#pragma unroll
                    for (int u = 0; u < NL; u++) {
                        if constexpr (PAIR && HV && EPI == EPI_CHEB) {
                            // r -= q, d_new = c1 d + c2 r, x += d_new, from the unrounded q
                            const size_t idx = row + 32 * u + 2 * lo;
                            const uint2 xw = epo[r][v][u][0], rw = epo[r][v][u][1], dw = epo[r][v][u][2];
                            const c32 q0{acc[r][2 * u][0][v], acc[r][2 * u][1][v]};
                            const c32 q1{acc[r][2 * u + 1][0][v], acc[r][2 * u + 1][1][v]};
                            const c32 r0 = unpack_h(rw.x) - q0, r1 = unpack_h(rw.y) - q1;
                            const c32 d0 = cc1 * unpack_h(dw.x) + cc2 * r0;
                            const c32 d1 = cc1 * unpack_h(dw.y) + cc2 * r1;
                            *reinterpret_cast<uint2*>(rwh + idx) = uint2{pack_h(r0), pack_h(r1)};
                            *reinterpret_cast<uint2*>(dsth + idx) = uint2{pack_h(d0), pack_h(d1)};
                            if (hxacc != 1) {   // x += d_new (and d when the last step left x alone)
                                c32 x0 = unpack_h(xw.x) + d0, x1 = unpack_h(xw.y) + d1;
                                if (hxacc == 2) {
                                    x0 = x0 + unpack_h(dw.x);
                                    x1 = x1 + unpack_h(dw.y);
                                }
                                *reinterpret_cast<uint2*>(xwh + idx) = uint2{pack_h(x0), pack_h(x1)};
                            }
                        } else if constexpr (PAIR && !HV && EPI == EPI_CHEB) {
                            // FP32 vectors: r -= s q, then (mode) d_new = c1 d + c2 r, x += d_new
                            const size_t idx = row + 32 * u + 2 * lo;
                            const float4 rw = *reinterpret_cast<const float4*>(bvec + idx);
                            const float sc = ccs;
                            const c32 r0 = c32{rw.x, rw.y} - sc * c32{acc[r][2 * u][0][v], acc[r][2 * u][1][v]};
                            const c32 r1 =
                                c32{rw.z, rw.w} - sc * c32{acc[r][2 * u + 1][0][v], acc[r][2 * u + 1][1][v]};
                            if (cmode != 2) *reinterpret_cast<float4*>(rwf + idx) = float4{r0.re, r0.im, r1.re, r1.im};
                            if (need_d) {
                                const float4 xw = need_x ? *reinterpret_cast<const float4*>(aux + idx)
                                                         : float4{0.0f, 0.0f, 0.0f, 0.0f};
                                const float4 dw = *reinterpret_cast<const float4*>(src + idx);
                                const c32 p0 = sc * c32{dw.x, dw.y}, p1 = sc * c32{dw.z, dw.w};
                                c32 x0 = cst == 1 ? p0 : (cst == 2 ? c32{xw.x, xw.y} + p0 : c32{xw.x, xw.y});
                                c32 x1 = cst == 1 ? p1 : (cst == 2 ? c32{xw.z, xw.w} + p1 : c32{xw.z, xw.w});
                                if (fxac == 2) {   // x was left alone by the step before
                                    x0 = x0 + p0;
                                    x1 = x1 + p1;
                                    if (fxad) {
                                        const float4 a = *reinterpret_cast<const float4*>(fxad + idx);
                                        x0 = x0 + fxas * c32{a.x, a.y};
                                        x1 = x1 + fxas * c32{a.z, a.w};
                                    }
                                }
                                const c32 d0 = cc1 * p0 + cc2 * r0;
                                const c32 d1 = cc1 * p1 + cc2 * r1;
                                if (cmode == 0) *reinterpret_cast<float4*>(dst + idx) = float4{d0.re, d0.im, d1.re, d1.im};
                                const c32 y0 = cmode == 1 ? x0 : x0 + d0, y1 = cmode == 1 ? x1 : x1 + d1;
                                if (fxac != 1) *reinterpret_cast<float4*>(xwf + idx) = float4{y0.re, y0.im, y1.re, y1.im};
                            }
                        } else if constexpr (!HV && EPI == EPI_CHEB) {
                            const size_t idx = row + 16 * u + lo;
                            const c32 rn = bvec[idx] - ccs * c32{acc[r][u][0][v], acc[r][u][1][v]};
                            if (cmode != 2) rwf[idx] = rn;
                            if (need_d) {
                                const c32 dp = ccs * src[idx];
                                c32 xp = cst == 1 ? dp : (cst == 2 ? aux[idx] + dp : aux[idx]);
                                if (fxac == 2) xp = xp + dp + (fxad ? fxas * fxad[idx] : c32{0.0f, 0.0f});
                                const c32 dn = cc1 * dp + cc2 * rn;
                                if (cmode == 0) dst[idx] = dn;
                                if (fxac != 1) xwf[idx] = cmode == 1 ? xp : xp + dn;
                            }
                        } else if constexpr (HV && EPI == EPI_CHEB) {
                            const size_t idx = row + 16 * u + lo;
                            const c32 q{acc[r][u][0][v], acc[r][u][1][v]};
                            const c32 rn = unpack_h(bh2[idx]) - q;
                            const c32 dp = unpack_h(srch[idx]);
                            const c32 dn = cc1 * dp + cc2 * rn;
                            rwh[idx] = pack_h(rn);
                            dsth[idx] = pack_h(dn);
                            if (hxacc != 1) xwh[idx] = pack_h(unpack_h(auxh[idx]) + dn + (hxacc == 2 ? dp : c32{0.0f, 0.0f}));
                        } else if constexpr (PAIR && HV) {
                            const size_t idx = row + 32 * u + 2 * lo;
                            const unsigned w0 = pack_h(c32{acc[r][2 * u][0][v], acc[r][2 * u][1][v]});
                            const unsigned w1 = pack_h(c32{acc[r][2 * u + 1][0][v], acc[r][2 * u + 1][1][v]});
                            *reinterpret_cast<uint2*>(dsth + idx) = uint2{w0, w1};
                            if (RED) {   // dots of the stored (rounded) values
                                const c32 o0 = unpack_h(w0), o1 = unpack_h(w1);
                                const uint2 w = epo[r][v][u][0];
                                const c32 p0 = cdot(o0, unpack_h(w.x)), p1 = cdot(o1, unpack_h(w.y));
                                dacc[2 * u][0] += p0.re;
                                dacc[2 * u][1] += p0.im;
                                dacc[2 * u][NC - 1] += norm2(o0);
                                dacc[2 * u + 1][0] += p1.re;
                                dacc[2 * u + 1][1] += p1.im;
                                dacc[2 * u + 1][NC - 1] += norm2(o1);
                                if constexpr (EPI == EPI_CG) {
                                    const uint2 z = epo[r][v][u][NE - 1];
                                    const c32 q0 = cdot(o0, unpack_h(z.x)), q1 = cdot(o1, unpack_h(z.y));
                                    dacc[2 * u][2] += q0.re;
                                    dacc[2 * u][3] += q0.im;
                                    dacc[2 * u + 1][2] += q1.re;
                                    dacc[2 * u + 1][3] += q1.im;
                                }
                            }
                        } else if constexpr (HV) {
                            const size_t idx = row + 16 * u + lo;
                            const unsigned w = pack_h(c32{acc[r][u][0][v], acc[r][u][1][v]});
                            dsth[idx] = w;
                            if (RED) {
                                const c32 o = unpack_h(w);
                                const c32 p = cdot(o, unpack_h(auxh[idx]));
                                dacc[u][0] += p.re;
                                dacc[u][1] += p.im;
                                dacc[u][NC - 1] += norm2(o);
                                if constexpr (EPI == EPI_CG) {
                                    const c32 q = cdot(o, unpack_h(bh2[idx]));
                                    dacc[u][2] += q.re;
                                    dacc[u][3] += q.im;
                                }
                            }
                        } else if constexpr (PAIR) {
                            const size_t idx = row + 32 * u + 2 * lo;
                            c32 o0{acc[r][2 * u][0][v], acc[r][2 * u][1][v]};
                            c32 o1{acc[r][2 * u + 1][0][v], acc[r][2 * u + 1][1][v]};
                            if (EPI == EPI_RESID) {
                                const float4 w = epo[r][v][u][0];
                                o0 = c32{w.x, w.y} - o0;
                                o1 = c32{w.z, w.w} - o1;
                            }
                            *reinterpret_cast<float4*>(dst + idx) = float4{o0.re, o0.im, o1.re, o1.im};
                            if (EPI == EPI_DOTS) {
                                const float4 w = epo[r][v][u][0];
                                const c32 p0 = cdot(o0, c32{w.x, w.y}), p1 = cdot(o1, c32{w.z, w.w});
                                dacc[2 * u][0] += p0.re;
                                dacc[2 * u][1] += p0.im;
                                dacc[2 * u][2] += norm2(o0);
                                dacc[2 * u + 1][0] += p1.re;
                                dacc[2 * u + 1][1] += p1.im;
                                dacc[2 * u + 1][2] += norm2(o1);
                            }
                        } else {
                            const size_t idx = row + 16 * u + lo;
                            c32 o{acc[r][u][0][v], acc[r][u][1][v]};
                            if (EPI == EPI_RESID) o = bvec[idx] - o;
                            dst[idx] = o;
                            if (EPI == EPI_DOTS) {
                                const c32 p = cdot(o, aux[idx]);
                                dacc[u][0] += p.re;
                                dacc[u][1] += p.im;
                                dacc[u][2] += norm2(o);
                            }
                        }
                    }
                }
        }
    }
    if (RED) {
#pragma unroll
        for (int t = 0; t < NT; t++)
#pragma unroll
            for (int c = 0; c < NC; c++) {
                dacc[t][c] += __shfl_xor(dacc[t][c], 16, 64);
                dacc[t][c] += __shfl_xor(dacc[t][c], 32, 64);
            }
        int tj, lj;
        if constexpr (PAIR) {
            tj = 2 * ((lane & 63) >> 5) + (lane & 1);
            lj = (lane & 31) >> 1;
        } else {
            tj = (lane >> 4) & (NT - 1);
            lj = lane & 15;
        }
        double v[NC];
#pragma unroll
        for (int c = 0; c < NC; c++) v[c] = 0.0;
#pragma unroll
        for (int t = 0; t < NT; t++)
#pragma unroll
            for (int c = 0; c < NC; c++) {
                const double x = __shfl(dacc[t][c], lj, 64);
                if (t == tj && lane < N) v[c] = x;
            }
        reduce_finish_b<NC, 64>(v, fin, N, gridDim.x, blockIdx.x);
    }
}

template <typename TL, int K, int NT, bool HV, bool HERM>
static void coarse_bf16_v(const CoarseView<TL>& C, void* dst, const void* src, Epi epi, const void* b,
                          const void* aux, const FinB& fin, int gdots, hipStream_t st) {
    const size_t waves = (size_t)C.nx * C.ny;
    const int nw = kBlock / 64;
    const bool red = epi == EPI_DOTS || epi == EPI_CG;
    const int g = (int)std::min<size_t>((waves + nw - 1) / nw, red ? gdots : 1 << 16);
    if (epi == EPI_NONE)
        k_b_coarse_bf16<TL, EPI_NONE, K, NT, HV, HERM><<<g, kBlock, 0, st>>>(C, dst, src, b, aux, fin);
    else if (epi == EPI_DOTS)
        k_b_coarse_bf16<TL, EPI_DOTS, K, NT, HV, HERM><<<g, kBlock, 0, st>>>(C, dst, src, b, aux, fin);
    else if (epi == EPI_CHEB)
        k_b_coarse_bf16<TL, EPI_CHEB, K, NT, HV, HERM><<<g, kBlock, 0, st>>>(C, dst, src, b, aux, fin);
    else if constexpr (!HV)
        k_b_coarse_bf16<TL, EPI_RESID, K, NT, false, HERM><<<g, kBlock, 0, st>>>(C, dst, src, b, aux, fin);
    else if (epi == EPI_CG)
        k_b_coarse_bf16<TL, EPI_CG, K, NT, true, HERM><<<g, kBlock, 0, st>>>(C, dst, src, b, aux, fin);
}
// The tuner's choice for this shape, else the defaults: the Hermitian reads whenever
// the links allow them, kRedGridCoarse workgroups with fused dot products.
template <typename TL, int K, int NT, bool HV = false>
static void coarse_bf16(const CoarseView<TL>& C, void* dst, const void* src, Epi epi, const void* b, const void* aux,
                        const FinB& fin, hipStream_t st) {
    const int choice = tune_lookup(TuneKey{TUNE_COARSE_BF16, C.nx, C.ny, K, 16 * NT, (HV ? 4 : 0) + (int)epi});
    const int gdots = choice < 0 ? kRedGridCoarse : kTuneGrids[choice >> 1];
    if constexpr (K % 16 == 0) {
        if (C.herm && (choice < 0 || (choice & 1))) {
            coarse_bf16_v<TL, K, NT, HV, true>(C, dst, src, epi, b, aux, fin, gdots, st);
            return;
        }
    }
    coarse_bf16_v<TL, K, NT, HV, false>(C, dst, src, epi, b, aux, fin, gdots, st);
}

template <typename T, typename TL, int K, int NT>
static void coarse_mfma(const CoarseView<TL>& C, cplx<T>* dst, const cplx<T>* src, Epi epi, const cplx<T>* b,
                        const cplx<T>* aux, const FinB& fin, hipStream_t st) {
    const size_t waves = (size_t)C.nx * C.ny;
    const int nw = kBlock / 64;
    const int g = (int)std::min<size_t>((waves + nw - 1) / nw, epi == EPI_DOTS ? kRedGridCoarse : 1 << 16);
    if (epi == EPI_NONE)
        k_b_coarse_mfma<T, TL, EPI_NONE, K, NT><<<g, kBlock, 0, st>>>(C, dst, src, b, aux, fin);
    else if (epi == EPI_RESID)
        k_b_coarse_mfma<T, TL, EPI_RESID, K, NT><<<g, kBlock, 0, st>>>(C, dst, src, b, aux, fin);
    else
        k_b_coarse_mfma<T, TL, EPI_DOTS, K, NT><<<g, kBlock, 0, st>>>(C, dst, src, b, aux, fin);
}

bool b_coarse_cheb_fits(int k, int N, bool f16_links_f32) {
    return f16_links_f32 && (k == 8 || k == 16 || k == 32) && (N == 16 || N == 32 || N == 64);
}

template <typename T, typename TL>
void launch_b_coarse_apply(const CoarseView<TL>& C, cplx<T>* dst, const cplx<T>* src, Epi epi, const cplx<T>* b,
                           const cplx<T>* aux, const FinB& fin, int N, hipStream_t st) {
    SMG_REQUIRE(epi == EPI_NONE || epi == EPI_RESID || epi == EPI_DOTS || epi == EPI_CHEB, "coarse apply: epilogue");
    // the Chebyshev step (r, x in b, aux; the direction in src): FP16 links on the matrix cores
    SMG_REQUIRE(epi != EPI_CHEB || (b_coarse_cheb_fits(C.k, N, std::is_same<T, float>::value &&
                                                                   std::is_same<TL, ch16>::value) &&
                                    cheb_ptrs_ok(dst, src, b, aux, fin)),
                "coarse apply: the Chebyshev epilogue needs FP16 links, FP32 vectors, k in {8,16,32}, "
                "N in {16,32,64} and r, x, d, d_new (and new r) that do not overlap as given");
    if constexpr (std::is_same<T, float>::value && std::is_same<TL, ch16>::value) {
        if ((C.k == 8 || C.k == 16 || C.k == 32) && (N == 16 || N == 32 || N == 64)) {
            // FP16 links: bf16x3 matrix-core path
            if (C.k == 8) {
                if (N == 16) coarse_bf16<TL, 8, 1>(C, dst, src, epi, b, aux, fin, st);
                else if (N == 32) coarse_bf16<TL, 8, 2>(C, dst, src, epi, b, aux, fin, st);
                else coarse_bf16<TL, 8, 4>(C, dst, src, epi, b, aux, fin, st);
            } else if (C.k == 16) {
                if (N == 16) coarse_bf16<TL, 16, 1>(C, dst, src, epi, b, aux, fin, st);
                else if (N == 32) coarse_bf16<TL, 16, 2>(C, dst, src, epi, b, aux, fin, st);
                else coarse_bf16<TL, 16, 4>(C, dst, src, epi, b, aux, fin, st);
            } else {
                if (N == 16) coarse_bf16<TL, 32, 1>(C, dst, src, epi, b, aux, fin, st);
                else if (N == 32) coarse_bf16<TL, 32, 2>(C, dst, src, epi, b, aux, fin, st);
                else coarse_bf16<TL, 32, 4>(C, dst, src, epi, b, aux, fin, st);
            }
            SMG_LAUNCH_CHECK();
            return;
        }
    }
    if ((C.k == 8 || C.k == 16 || C.k == 32) && (N == 16 || N == 32 || N == 64)) {
        // matrix-core path
        if (C.k == 8) {
            if (N == 16) coarse_mfma<T, TL, 8, 1>(C, dst, src, epi, b, aux, fin, st);
            else if (N == 32) coarse_mfma<T, TL, 8, 2>(C, dst, src, epi, b, aux, fin, st);
            else coarse_mfma<T, TL, 8, 4>(C, dst, src, epi, b, aux, fin, st);
        } else if (C.k == 16) {
            if (N == 16) coarse_mfma<T, TL, 16, 1>(C, dst, src, epi, b, aux, fin, st);
            else if (N == 32) coarse_mfma<T, TL, 16, 2>(C, dst, src, epi, b, aux, fin, st);
            else coarse_mfma<T, TL, 16, 4>(C, dst, src, epi, b, aux, fin, st);
        } else {
            if (N == 16) coarse_mfma<T, TL, 32, 1>(C, dst, src, epi, b, aux, fin, st);
            else if (N == 32) coarse_mfma<T, TL, 32, 2>(C, dst, src, epi, b, aux, fin, st);
            else coarse_mfma<T, TL, 32, 4>(C, dst, src, epi, b, aux, fin, st);
        }
        SMG_LAUNCH_CHECK();
        return;
    }
    const size_t items = (size_t)C.nx * C.ny * C.k * N;
    if (epi == EPI_NONE)
        k_b_coarse<T, TL, EPI_NONE><<<b_grid(items, 1 << 16), kBlock, 0, st>>>(C, dst, src, b, aux, fin, N);
    else if (epi == EPI_RESID)
        k_b_coarse<T, TL, EPI_RESID><<<b_grid(items, 1 << 16), kBlock, 0, st>>>(C, dst, src, b, aux, fin, N);
    else
        k_b_coarse<T, TL, EPI_DOTS><<<b_grid(items, kRedGridCoarse), kBlock, 0, st>>>(C, dst, src, b, aux, fin, N);
    SMG_LAUNCH_CHECK();
}

// The coarse apply on BF16 vectors (FP16 links): epi EPI_NONE, EPI_DOTS, EPI_CG or EPI_CHEB
void launch_b_coarse_apply_h(const CoarseView<ch16>& C, cb16* dst, const cb16* src, Epi epi, const cb16* aux,
                             const FinB& fin, int N, hipStream_t st, const cb16* b) {
    SMG_REQUIRE((C.k == 8 || C.k == 16 || C.k == 32) && (N == 16 || N == 32 || N == 64) && epi != EPI_RESID,
                "BF16-vector coarse apply: k in {8,16,32}, N in {16,32,64}, no residual epilogue");
    SMG_REQUIRE(epi != EPI_CG || (aux && b), "BF16-vector coarse apply: the CG epilogue takes aux and b");
    // the Chebyshev step writes r (b) and x (aux) in place and the new direction to dst,
    // while other workgroups read the old direction (src) at their neighbours
    SMG_REQUIRE(epi != EPI_CHEB || (aux && b && dst != src && aux != src && b != src && dst != aux && dst != b &&
                                    aux != b && fin.cheb_start == 0 && !fin.cheb_rout && fin.cheb_mode == 0 &&
                                    fin.cheb_xacc >= 0 && fin.cheb_xacc <= 2 && !fin.cheb_xadd),
                "BF16-vector coarse apply: the Chebyshev epilogue takes distinct r (b), x (aux), d (src) and "
                "d_new (dst)");
    auto go = [&](auto kc, auto ntc) {
        constexpr int K = decltype(kc)::value, NT = decltype(ntc)::value;
        coarse_bf16<ch16, K, NT, true>(C, dst, src, epi, b, aux, fin, st);
    };
    auto by_n = [&](auto kc) {
        if (N == 16) go(kc, std::integral_constant<int, 1>{});
        else if (N == 32) go(kc, std::integral_constant<int, 2>{});
        else go(kc, std::integral_constant<int, 4>{});
    };
    if (C.k == 8) by_n(std::integral_constant<int, 8>{});
    else if (C.k == 16) by_n(std::integral_constant<int, 16>{});
    else by_n(std::integral_constant<int, 32>{});
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
        const int k0 = A.chiral ? agg_chi_dof(A, c) * (A.k / 2) : 0, k1 = A.chiral ? k0 + A.k / 2 : A.k;
        cplx<T> acc = {(T)0, (T)0};
        for (int kk = k0; kk < k1; kk++) acc = acc + cc[(size_t)kk * N] * v[(size_t)kk * bsize];
        fine[t] = accumulate ? fine[t] + acc : acc;
    }
}
// The same with the number of coarse dof summed per fine dof (k, or k/2 with chiral
// aggregation) known at compile time: the KH loads of a work-item issue together
// instead of one dependent round trip per term.
template <typename T, int KH, int P>
__global__ __launch_bounds__(kBlock) void k_b_prolong_k(AggView A, const cplx<T>* __restrict__ bv,
                                                        const cplx<T>* __restrict__ coarse,
                                                        cplx<T>* __restrict__ fine, int accumulate, int N) {
    const int lN = lg2(N);
    const size_t runs = ((size_t)A.nx * A.ny * A.nd * N) / P;
    const int bsize = A.bx * A.by * A.nd;
    for (size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x; t < runs; t += (size_t)gridDim.x * kBlock) {
        const size_t i = t * P;
        const size_t f = i >> lN;
        const int j = (int)(i & (N - 1));
        const int s = (int)(f / A.nd), c = (int)(f % A.nd);
        const int x = s % A.nx, y = s / A.nx;
        const int b = x / A.bx + A.nbx * (y / A.by);
        const int loc = (x % A.bx) + A.bx * (y % A.by);
        const int k0 = A.chiral ? agg_chi_dof(A, c) * KH : 0;
        const cplx<T>* v = bv + ((size_t)b * A.k + k0) * bsize + loc * A.nd + c;
        const cplx<T>* cc = coarse + ((size_t)b * A.k + k0) * N + j;
        RhsRun<T, P> cv[KH];
        cplx<T> vv[KH];
#pragma unroll
        for (int q = 0; q < KH; q++) {
            cv[q].load(cc + (size_t)q * N);
            vv[q] = v[(size_t)q * bsize];
        }
        RhsRun<T, P> out;
#pragma unroll
        for (int u = 0; u < P; u++) {
            cplx<T> acc = {(T)0, (T)0};
#pragma unroll
            for (int q = 0; q < KH; q++) acc = acc + cv[q].v[u] * vv[q];
            out.v[u] = acc;
        }
        if (accumulate) {
            RhsRun<T, P> old;
            old.load(fine + i);
#pragma unroll
            for (int u = 0; u < P; u++) out.v[u] = old.v[u] + out.v[u];
        }
        out.store(fine + i);
    }
}
// ---- transfers with a workgroup per aggregate: the aggregate's coarse block (k x N),
// block vectors (k x bsize) and, for the restriction, fine block (bsize x N) are staged
// in LDS once with 16-byte copies, and every output is formed from LDS in the same
// order of terms as the kernels above (the same sums).  Each coarse entry and block
// vector is then read from HBM once per launch instead of once per output that uses it.
template <typename T>
__device__ __forceinline__ void stage16(cplx<T>* __restrict__ dst, const cplx<T>* __restrict__ src, int n) {
    const int n16 = n * (int)sizeof(cplx<T>) / 16;
    const uint4* s = reinterpret_cast<const uint4*>(src);
    uint4* d = reinterpret_cast<uint4*>(dst);
    for (int i = threadIdx.x; i < n16; i += kBlock) d[i] = s[i];
}
// fine(s, c)[j] (+)= sum_kk v_{b,kk}(loc, c) coarse[b*k + kk][j], kk over c's chirality
template <typename T>
__global__ __launch_bounds__(kBlock) void k_b_prolong_lds(AggView A, const cplx<T>* __restrict__ bv,
                                                          const cplx<T>* __restrict__ coarse,
                                                          cplx<T>* __restrict__ fine, int accumulate, int N) {
    extern __shared__ uint4 smem_pl[];
    const int k = A.k, bsize = A.bx * A.by * A.nd, lN = lg2(N);
    cplx<T>* cs = reinterpret_cast<cplx<T>*>(smem_pl);   // [k][N]
    cplx<T>* vs = cs + k * N;                             // [k][bsize]
    const int b = blockIdx.x;
    stage16(cs, coarse + (size_t)b * k * N, k * N);
    stage16(vs, bv + (size_t)b * k * bsize, k * bsize);
    __syncthreads();
    const int bxi = b % A.nbx, byi = b / A.nbx;
    const int kh = A.chiral ? k / 2 : k;
    for (int o = threadIdx.x; o < (bsize << lN); o += kBlock) {
        const int e = o >> lN, j = o & (N - 1);
        const int loc = e / A.nd, c = e % A.nd;
        const int k0 = A.chiral ? agg_chi_dof(A, c) * kh : 0;
        cplx<T> acc = {(T)0, (T)0};
        for (int q = 0; q < kh; q++) acc = acc + cs[(k0 + q) * N + j] * vs[(k0 + q) * bsize + e];
        const int sx = bxi * A.bx + loc % A.bx, sy = byi * A.by + loc / A.bx;
        cplx<T>* out = fine + (((size_t)(sx + A.nx * sy) * A.nd + c) << lN) + j;
        *out = accumulate ? *out + acc : acc;
    }
}
// coarse[b*k + kk][j] = sum over the aggregate's (loc, c) of conj(v_{b,kk}(loc, c)) fine(s, c)[j]
template <typename T>
__global__ __launch_bounds__(kBlock) void k_b_restrict_lds(AggView A, const cplx<T>* __restrict__ bv,
                                                           const cplx<T>* __restrict__ fine,
                                                           cplx<T>* __restrict__ coarse, int N) {
    extern __shared__ uint4 smem_rs[];
    const int k = A.k, bs = A.bx * A.by, bsize = bs * A.nd, lN = lg2(N);
    cplx<T>* fs = reinterpret_cast<cplx<T>*>(smem_rs);   // [bsize][N]
    cplx<T>* vs = fs + (bsize << lN);                      // [k][bsize]
    const int b = blockIdx.x;
    const int bxi = b % A.nbx, byi = b / A.nbx;
    // a site's nd x N entries are contiguous: stage site by site
    const int per16 = (A.nd << lN) * (int)sizeof(cplx<T>) / 16;
    for (int i = threadIdx.x; i < bs * per16; i += kBlock) {
        const int loc = i / per16, w = i % per16;
        const int sx = bxi * A.bx + loc % A.bx, sy = byi * A.by + loc / A.bx;
        const uint4* s = reinterpret_cast<const uint4*>(fine + ((size_t)(sx + A.nx * sy) * A.nd << lN));
        reinterpret_cast<uint4*>(fs)[loc * per16 + w] = s[w];
    }
    stage16(vs, bv + (size_t)b * k * bsize, k * bsize);
    __syncthreads();
    const int nc = A.chiral ? A.nd / 2 : A.nd;
    for (int o = threadIdx.x; o < (k << lN); o += kBlock) {
        const int kk = o >> lN, j = o & (N - 1);
        const int c0 = A.chiral ? agg_chi_vec(A, kk) * nc : 0;
        const cplx<T>* v = vs + kk * bsize;
        cplx<T> acc = {(T)0, (T)0};
        for (int loc = 0; loc < bs; loc++)
            for (int c = c0; c < c0 + nc; c++) acc = acc + cdot(v[loc * A.nd + c], fs[((loc * A.nd + c) << lN) + j]);
        coarse[(((size_t)b * k + kk) << lN) + j] = acc;
    }
}
// ---- transfers on the matrix cores (FP32 vectors, N a multiple of 16): one wave per
// 16 x 16 tile of one chirality half of one aggregate, v_mfma_f32_16x16x4_f32 with each
// complex product as four real ones (FP32 products and FP32 sums, as the kernels above
// form them, in another order).  Entry e (0 .. kc-1) of half h is dof h*nc + e % nc of
// the aggregate's site e / nc (nh = 2 halves with chiral aggregation, else 1).
typedef float f32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ size_t half_entry_site(const AggView& A, int b, int loc) {
    const int sx = (b % A.nbx) * A.bx + loc % A.bx, sy = (b / A.nbx) * A.by + loc / A.bx;
    return (size_t)sx + (size_t)A.nx * sy;
}
// coarse[b*k + kk][j] = sum over half h's KC entries of conj(V[kk][e]) fine[e][j], kk in
// half h: A = the 16 block vectors of the tile (rows) by 4 entries, B = 4 entries by 16 RHS
template <int KC>
__global__ __launch_bounds__(kBlock) void k_b_restrict_mfma(AggView A, const c32* __restrict__ bv,
                                                             const c32* __restrict__ fine, c32* __restrict__ coarse,
                                                             int N) {
    const int nh = A.chiral ? 2 : 1, kh = A.k / nh, nc = A.nd / nh, nmt = kh >> 4, njc = N >> 4;
    const int bsize = A.bx * A.by * A.nd;
    size_t w = (size_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
    if (w >= (size_t)A.nbx * A.nby * nh * nmt * njc) return;
    const int jc = (int)(w % njc);
    w /= njc;
    const int mt = (int)(w % nmt);
    w /= nmt;
    const int h = (int)(w % nh), b = (int)(w / nh);
    const int lane = threadIdx.x & 63, m = lane & 15, q = lane >> 4;
    const int j = jc * 16 + m, kk0 = h * kh + mt * 16;
    const c32* vrow = bv + ((size_t)b * A.k + kk0 + m) * bsize;
    c32 va[KC / 4], fa[KC / 4];
#pragma unroll
    for (int s = 0; s < KC / 4; s++) {
        const int e = 4 * s + q, loc = e / nc, c = h * nc + e % nc;
        va[s] = vrow[loc * A.nd + c];
        fa[s] = fine[(half_entry_site(A, b, loc) * A.nd + c) * N + j];
    }
    f32x4 cr = {0.0f, 0.0f, 0.0f, 0.0f}, ci = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int s = 0; s < KC / 4; s++) {   // conj(v) f = (vr fr + vi fi) + i (vr fi - vi fr)
        cr = __builtin_amdgcn_mfma_f32_16x16x4f32(va[s].re, fa[s].re, cr, 0, 0, 0);
        cr = __builtin_amdgcn_mfma_f32_16x16x4f32(va[s].im, fa[s].im, cr, 0, 0, 0);
        ci = __builtin_amdgcn_mfma_f32_16x16x4f32(va[s].re, fa[s].im, ci, 0, 0, 0);
        ci = __builtin_amdgcn_mfma_f32_16x16x4f32(-va[s].im, fa[s].re, ci, 0, 0, 0);
    }
    // lane (m, q) holds rows 4q .. 4q+3 of column m
#pragma unroll
    for (int i = 0; i < 4; i++) coarse[((size_t)b * A.k + kk0 + 4 * q + i) * N + j] = c32{cr[i], ci[i]};
}
// fine[e][j] (+)= sum over half h's KH block vectors of V[kk][e] coarse[kk][j]: A = 16
// entries (rows) by 4 block vectors, B = 4 block vectors by 16 RHS
// accumulate 0: fine = P c; 1: fine += P c; 2: fine = s base + P c
template <int KH>
__global__ __launch_bounds__(kBlock) void k_b_prolong_mfma(AggView A, const c32* __restrict__ bv,
                                                            const c32* __restrict__ coarse, c32* __restrict__ fine,
                                                            int accumulate, int N, const c32* __restrict__ base,
                                                            float bs) {
    const int nh = A.chiral ? 2 : 1, nc = A.nd / nh, kc = A.bx * A.by * nc, nmt = kc >> 4, njc = N >> 4;
    const int bsize = A.bx * A.by * A.nd;
    size_t w = (size_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
    if (w >= (size_t)A.nbx * A.nby * nh * nmt * njc) return;
    const int jc = (int)(w % njc);
    w /= njc;
    const int mt = (int)(w % nmt);
    w /= nmt;
    const int h = (int)(w % nh), b = (int)(w / nh);
    const int lane = threadIdx.x & 63, m = lane & 15, q = lane >> 4;
    const int j = jc * 16 + m;
    const int ea = mt * 16 + m;   // this lane's A row: entry ea of the half
    const c32* vcol = bv + ((size_t)b * A.k + h * KH) * bsize + (ea / nc) * A.nd + h * nc + ea % nc;
    const c32* crow = coarse + ((size_t)b * A.k + h * KH) * N + j;
    c32 va[KH / 4], ca[KH / 4];
#pragma unroll
    for (int s = 0; s < KH / 4; s++) {
        va[s] = vcol[(size_t)(4 * s + q) * bsize];
        ca[s] = crow[(size_t)(4 * s + q) * N];
    }
    f32x4 fr = {0.0f, 0.0f, 0.0f, 0.0f}, fi = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int s = 0; s < KH / 4; s++) {   // v c = (vr cr - vi ci) + i (vr ci + vi cr)
        fr = __builtin_amdgcn_mfma_f32_16x16x4f32(va[s].re, ca[s].re, fr, 0, 0, 0);
        fr = __builtin_amdgcn_mfma_f32_16x16x4f32(-va[s].im, ca[s].im, fr, 0, 0, 0);
        fi = __builtin_amdgcn_mfma_f32_16x16x4f32(va[s].re, ca[s].im, fi, 0, 0, 0);
        fi = __builtin_amdgcn_mfma_f32_16x16x4f32(va[s].im, ca[s].re, fi, 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int e = mt * 16 + 4 * q + i, loc = e / nc, c = h * nc + e % nc;
        c32* out = fine + (half_entry_site(A, b, loc) * A.nd + c) * N + j;
        const c32 v{fr[i], fi[i]};
        if (accumulate == 2) *out = bs * base[out - fine] + v;
        else *out = accumulate ? *out + v : v;
    }
}
// whether the matrix-core transfers take this aggregation and N (the tuner's
// TUNE_TRANSFER choice 2): N a multiple of 16, 16 or 32 block vectors per half, and
// 16, 32 or 64 entries per half
bool b_transfer_mfma_fits(const AggView& A, int N) {
    const int nh = A.chiral ? 2 : 1;
    if (A.k % nh || A.nd % nh || N % 16) return false;
    const int kh = A.k / nh, kc = A.bx * A.by * (A.nd / nh);
    return (kh == 16 || kh == 32) && (kc == 16 || kc == 32 || kc == 64);
}
static void restrict_mfma(const AggView& A, const c32* bv, const c32* fine, c32* coarse, int N, hipStream_t st) {
    const int nh = A.chiral ? 2 : 1, kc = A.bx * A.by * (A.nd / nh);
    const size_t tiles = (size_t)A.nbx * A.nby * nh * ((A.k / nh) >> 4) * (N >> 4);
    const unsigned g = (unsigned)((tiles + kBlock / 64 - 1) / (kBlock / 64));
    if (kc == 16) k_b_restrict_mfma<16><<<g, kBlock, 0, st>>>(A, bv, fine, coarse, N);
    else if (kc == 32) k_b_restrict_mfma<32><<<g, kBlock, 0, st>>>(A, bv, fine, coarse, N);
    else k_b_restrict_mfma<64><<<g, kBlock, 0, st>>>(A, bv, fine, coarse, N);
}
static void prolong_mfma(const AggView& A, const c32* bv, const c32* coarse, c32* fine, int accumulate, int N,
                         hipStream_t st, const c32* base, float bs) {
    const int nh = A.chiral ? 2 : 1, kh = A.k / nh, kc = A.bx * A.by * (A.nd / nh);
    const size_t tiles = (size_t)A.nbx * A.nby * nh * (kc >> 4) * (N >> 4);
    const unsigned g = (unsigned)((tiles + kBlock / 64 - 1) / (kBlock / 64));
    if (kh == 16) k_b_prolong_mfma<16><<<g, kBlock, 0, st>>>(A, bv, coarse, fine, accumulate, N, base, bs);
    else k_b_prolong_mfma<32><<<g, kBlock, 0, st>>>(A, bv, coarse, fine, accumulate, N, base, bs);
}
constexpr size_t kTransferLdsMax = 64 * 1024;
// LDS bytes of the staged transfer, or 0 where it does not apply (sizes that are not
// whole 16-byte runs, or more than kTransferLdsMax)
size_t b_transfer_lds_bytes(const AggView& A, int N, size_t cbytes, bool prolong) {
    const size_t bsize = (size_t)A.bx * A.by * A.nd, k = A.k;
    const size_t n = prolong ? k * N + k * bsize : bsize * N + k * bsize;
    const bool whole = (k * N * cbytes) % 16 == 0 && (k * bsize * cbytes) % 16 == 0 &&
                       ((size_t)A.nd * N * cbytes) % 16 == 0;
    if ((N & (N - 1)) != 0 || !whole || n * cbytes > kTransferLdsMax) return 0;
    return n * cbytes;
}
static int transfer_choice(const AggView& A, int N, size_t cbytes, bool prolong) {
    return tune_lookup(TuneKey{TUNE_TRANSFER, A.nx, A.ny, 128 * A.nd + A.k, N,
                               2 * (cbytes == sizeof(c64) ? 1 : 0) + (prolong ? 1 : 0)});
}

template <typename T>
void launch_b_prolong(const AggView& A, const cplx<T>* bv, const cplx<T>* coarse, cplx<T>* fine, int accumulate,
                      int N, hipStream_t st, const cplx<T>* base, double bs) {
    SMG_REQUIRE(accumulate >= 0 && accumulate <= 2 && (accumulate != 2 || (base && base != fine)),
                "batched prolongation: accumulate 0, 1 or 2 (with a base vector apart from the output)");
    const int choice = transfer_choice(A, N, sizeof(cplx<T>), true);
    if constexpr (std::is_same<T, float>::value) {
        if (choice == 2 && b_transfer_mfma_fits(A, N)) {
            prolong_mfma(A, bv, coarse, fine, accumulate, N, st, base, (float)bs);
            SMG_LAUNCH_CHECK();
            return;
        }
    }
    if (accumulate == 2) {   // the other kernels: fine = s base first, then fine += P c
        launch_b_axpby_real<T>(fine, base, bs, 0.0, (size_t)A.nx * A.ny * A.nd * N, st);
        accumulate = 1;
    }
    if (choice == 1) {
        const size_t lds = b_transfer_lds_bytes(A, N, sizeof(cplx<T>), true);
        if (lds) {
            k_b_prolong_lds<T><<<A.nbx * A.nby, kBlock, lds, st>>>(A, bv, coarse, fine, accumulate, N);
            SMG_LAUNCH_CHECK();
            return;
        }
    }
    const size_t items = (size_t)A.nx * A.ny * A.nd * N;
    const int g = b_grid(items, 1 << 16);
    const int kh = A.chiral ? A.k / 2 : A.k;
    if (kh != 4 && kh != 8 && kh != 16 && kh != 32) {
        k_b_prolong<T><<<g, kBlock, 0, st>>>(A, bv, coarse, fine, accumulate, N);
    } else {
        run_dispatch<T>(N, [&](auto pc) {
            constexpr int P = decltype(pc)::value;
            const int gp = b_grid(items / P, 1 << 16);
            if (kh == 4) k_b_prolong_k<T, 4, P><<<gp, kBlock, 0, st>>>(A, bv, coarse, fine, accumulate, N);
            else if (kh == 8) k_b_prolong_k<T, 8, P><<<gp, kBlock, 0, st>>>(A, bv, coarse, fine, accumulate, N);
            else if (kh == 16) k_b_prolong_k<T, 16, P><<<gp, kBlock, 0, st>>>(A, bv, coarse, fine, accumulate, N);
            else k_b_prolong_k<T, 32, P><<<gp, kBlock, 0, st>>>(A, bv, coarse, fine, accumulate, N);
        });
    }
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
        const int c0 = A.chiral ? agg_chi_vec(A, kk) * (A.nd / 2) : 0, c1 = A.chiral ? c0 + A.nd / 2 : A.nd;
        cplx<T> acc = {(T)0, (T)0};
        for (int loc = 0; loc < bsites; loc++) {
            const int s = (bxi * A.bx + loc % A.bx) + A.nx * (byi * A.by + loc / A.bx);
            const cplx<T>* fp = fine + ((size_t)s * A.nd) * N + j;
            cplx<T> a = cdot(v[loc * A.nd + c0], fp[(size_t)c0 * N]);
            for (int c = c0 + 1; c < c1; c++) a = a + cdot(v[loc * A.nd + c], fp[(size_t)c * N]);
            acc = acc + a;
        }
        coarse[t] = acc;
    }
}
// The same with the aggregate's sites BS and the dof per site summed NC (nd, or
// nd/2 with chiral aggregation) known at compile time; the loads are issued in
// groups of up to 16 terms.
template <typename T, int BS, int NC, int P>
__global__ __launch_bounds__(kBlock) void k_b_restrict_k(AggView A, const cplx<T>* __restrict__ bv,
                                                         const cplx<T>* __restrict__ fine,
                                                         cplx<T>* __restrict__ coarse, int N) {
    constexpr int NT = BS * NC, G = NT < 16 ? NT : 16;
    static_assert(NT % G == 0, "terms in whole groups");
    const int lN = lg2(N);
    const size_t runs = ((size_t)A.nbx * A.nby * A.k * N) / P;
    const int bsize = BS * A.nd;
    for (size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x; t < runs; t += (size_t)gridDim.x * kBlock) {
        const size_t i = t * P;
        const size_t row = i >> lN;   // b*k + kk
        const int j = (int)(i & (N - 1));
        const int b = (int)(row / A.k), kk = (int)(row % A.k);
        const int bxi = b % A.nbx, byi = b / A.nbx;
        const cplx<T>* v = bv + ((size_t)b * A.k + kk) * bsize;
        const int c0 = A.chiral ? agg_chi_vec(A, kk) * NC : 0;
        cplx<T> acc[P];
#pragma unroll
        for (int u = 0; u < P; u++) acc[u] = cplx<T>{(T)0, (T)0};
#pragma unroll
        for (int g0 = 0; g0 < NT; g0 += G) {
            RhsRun<T, P> fv[G];
            cplx<T> vv[G];
#pragma unroll
            for (int q = 0; q < G; q++) {
                const int loc = (g0 + q) / NC, c = c0 + (g0 + q) % NC;
                const int sx = bxi * A.bx + loc % A.bx, sy = byi * A.by + loc / A.bx;
                fv[q].load(fine + ((size_t)(sx + A.nx * sy) * A.nd + c) * N + j);
                vv[q] = v[loc * A.nd + c];
            }
#pragma unroll
            for (int q = 0; q < G; q++)
#pragma unroll
                for (int u = 0; u < P; u++) acc[u] = acc[u] + cdot(vv[q], fv[q].v[u]);
        }
        RhsRun<T, P> out;
#pragma unroll
        for (int u = 0; u < P; u++) out.v[u] = acc[u];
        out.store(coarse + i);
    }
}
template <typename T, int BS, int P>
static bool restrict_nc(const AggView& A, int nc, int g, const cplx<T>* bv, const cplx<T>* fine, cplx<T>* coarse,
                        int N, hipStream_t st) {
    switch (nc) {
        case 1: k_b_restrict_k<T, BS, 1, P><<<g, kBlock, 0, st>>>(A, bv, fine, coarse, N); return true;
        case 2: k_b_restrict_k<T, BS, 2, P><<<g, kBlock, 0, st>>>(A, bv, fine, coarse, N); return true;
        case 4: k_b_restrict_k<T, BS, 4, P><<<g, kBlock, 0, st>>>(A, bv, fine, coarse, N); return true;
        case 8: k_b_restrict_k<T, BS, 8, P><<<g, kBlock, 0, st>>>(A, bv, fine, coarse, N); return true;
        case 16: k_b_restrict_k<T, BS, 16, P><<<g, kBlock, 0, st>>>(A, bv, fine, coarse, N); return true;
        case 32: k_b_restrict_k<T, BS, 32, P><<<g, kBlock, 0, st>>>(A, bv, fine, coarse, N); return true;
        default: return false;
    }
}
template <typename T>
void launch_b_restrict(const AggView& A, const cplx<T>* bv, const cplx<T>* fine, cplx<T>* coarse, int N,
                       hipStream_t st) {
    const int choice = transfer_choice(A, N, sizeof(cplx<T>), false);
    if constexpr (std::is_same<T, float>::value) {
        if (choice == 2 && b_transfer_mfma_fits(A, N)) {
            restrict_mfma(A, bv, fine, coarse, N, st);
            SMG_LAUNCH_CHECK();
            return;
        }
    }
    if (choice == 1) {
        const size_t lds = b_transfer_lds_bytes(A, N, sizeof(cplx<T>), false);
        if (lds) {
            k_b_restrict_lds<T><<<A.nbx * A.nby, kBlock, lds, st>>>(A, bv, fine, coarse, N);
            SMG_LAUNCH_CHECK();
            return;
        }
    }
    const size_t items = (size_t)A.nbx * A.nby * A.k * N;
    const int g = b_grid(items, 1 << 16);
    const int bs = A.bx * A.by, nc = A.chiral ? A.nd / 2 : A.nd;
    bool done = false;
    run_dispatch<T>(N, [&](auto pc) {
        constexpr int P = decltype(pc)::value;
        const int gp = b_grid(items / P, 1 << 16);
        if (bs == 4) done = restrict_nc<T, 4, P>(A, nc, gp, bv, fine, coarse, N, st);
        else if (bs == 16) done = restrict_nc<T, 16, P>(A, nc, gp, bv, fine, coarse, N, st);
    });
    if (!done) k_b_restrict<T><<<g, kBlock, 0, st>>>(A, bv, fine, coarse, N);
    SMG_LAUNCH_CHECK();
}

// Y = M X, M row-major n x n (the dense coarsest inverse); one work-item per (row, j).
// Used when the matrix-core path does not apply (N < 16 or n % 8 != 0).
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

// Matrix-core dense product: wave w computes one tile of 8 complex rows x N RHS
// over column slice w / ntile (of SK equal slices); slices land in part[slice]
// and are summed in slice order by k_b_dense_sum.
template <typename T, int NT>
__global__ __launch_bounds__(kBlock) void k_b_dense_mfma(const cplx<T>* __restrict__ M,
                                                         const cplx<T>* __restrict__ X,
                                                         cplx<T>* __restrict__ part, int n, int SK) {
    constexpr int N = 16 * NT;
    using MT = MfmaTile<T>;
    using acc_t = typename MT::acc;
    const int lane = threadIdx.x & 63, lo = lane & 15, hi = lane >> 4;
    const int ntile = n / 8;
    const int w = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
    if (w >= ntile * SK) return;   // wave-uniform; no workgroup barrier follows
    const int tile = w % ntile, slice = w / ntile;
    const int kc = n / SK, k0 = slice * kc;
    const int a_c = MT::row_c(lo);
    const cplx<T>* mp = M + (size_t)(8 * tile + MT::row_ki(lo)) * n + k0 + (hi >> 1);
    const T* xp = reinterpret_cast<const T*>(X) + ((size_t)(k0 + (hi >> 1)) * N + lo) * 2 + (hi & 1);
    acc_t acc[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) acc[t] = acc_t{0, 0, 0, 0};
    for (int k = 0; k < kc; k += 2) {
        const T a = real_form(mp[k], a_c, hi & 1);
#pragma unroll
        for (int t = 0; t < NT; t++) acc[t] = MT::mma(a, xp[((size_t)k * N + 16 * t) * 2], acc[t]);
    }
    cplx<T>* out = part + (size_t)slice * n * N;
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
        for (int h = 0; h < 2; h++)
            out[(size_t)(8 * tile + 2 * hi + h) * N + 16 * t + lo] = cplx<T>{acc[t][2 * h], acc[t][2 * h + 1]};
}
template <typename T>
__global__ __launch_bounds__(kBlock) void k_b_dense_sum(const cplx<T>* __restrict__ part, cplx<T>* __restrict__ Y,
                                                        size_t items, int SK) {
    for (size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x; t < items; t += (size_t)gridDim.x * kBlock) {
        cplx<T> a = part[t];
        for (int s = 1; s < SK; s++) a = a + part[(size_t)s * items + t];
        Y[t] = a;
    }
}

// column slices of the matrix-core dense product: about 2048+ waves, kc = n/SK even
static int dense_slices(size_t n) {
    int sk = 1;
    while (sk < 64 && (n / 8) * sk < 2048 && n % (size_t)(4 * sk) == 0) sk *= 2;
    return sk;
}
size_t b_dense_work(size_t n, int N) {
    return (N >= 16 && N <= 64 && n % 8 == 0) ? (size_t)dense_slices(n) * n * N : 0;
}

template <typename T>
void launch_b_dense_apply(const cplx<T>* M, const cplx<T>* x, cplx<T>* y, size_t n, int N, cplx<T>* work,
                          hipStream_t st) {
    const size_t need = b_dense_work(n, N);
    if (need > 0 && work != nullptr) {
        SMG_REQUIRE(n < (size_t)1 << 30, "dense coarsest: n too large");
        const int sk = dense_slices(n);
        const size_t waves = (n / 8) * sk;
        const int g = (int)((waves + 3) / 4);
        if (N == 16) k_b_dense_mfma<T, 1><<<g, kBlock, 0, st>>>(M, x, work, (int)n, sk);
        else if (N == 32) k_b_dense_mfma<T, 2><<<g, kBlock, 0, st>>>(M, x, work, (int)n, sk);
        else k_b_dense_mfma<T, 4><<<g, kBlock, 0, st>>>(M, x, work, (int)n, sk);
        SMG_LAUNCH_CHECK();
        k_b_dense_sum<T><<<b_grid(n * N, 1 << 16), kBlock, 0, st>>>(work, y, n * N, sk);
        SMG_LAUNCH_CHECK();
        return;
    }
    k_b_dense<T><<<b_grid(n * N, 1 << 16), kBlock, 0, st>>>(M, x, y, n, N);
    SMG_LAUNCH_CHECK();
}

// ----------------------------------------------------------------- BLAS-1
// The vector kernels below take runs of P right-hand sides per work-item (P = 2:
// 16-byte accesses in FP32), chosen by run_dispatch.
template <typename T, int P>
__global__ __launch_bounds__(kBlock) void k_b_norm2(const cplx<T>* __restrict__ a, size_t runs, FinB fin, int N) {
    double acc[1][P] = {};
    for (size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x; t < runs; t += (size_t)gridDim.x * kBlock) {
        RhsRun<T, P> x;
        x.load(a + t * P);
#pragma unroll
        for (int q = 0; q < P; q++) acc[0][q] += (double)norm2(x.v[q]);
    }
    double v[1];
    runs_to_lanes<1, P>(acc, v, N);
    reduce_finish_b<1, 64>(v, fin, N, gridDim.x, blockIdx.x);
}
template <typename T>
void launch_b_norm2(const cplx<T>* a, size_t n, const FinB& fin, int N, hipStream_t st) {
    run_dispatch<T>(N, [&](auto pc) {
        constexpr int P = decltype(pc)::value;
        k_b_norm2<T, P><<<b_grid(n * N / P, kRedGridVec), kBlock, 0, st>>>(a, n * N / P, fin, N);
    });
    SMG_LAUNCH_CHECK();
}

template <typename T, int P>
__global__ __launch_bounds__(kBlock) void k_b_dots(const cplx<T>* __restrict__ a, const cplx<T>* __restrict__ b,
                                                   size_t runs, FinB fin, int N) {
    double acc[3][P] = {};
    for (size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x; t < runs; t += (size_t)gridDim.x * kBlock) {
        RhsRun<T, P> x, y;
        x.load(a + t * P);
        y.load(b + t * P);
#pragma unroll
        for (int q = 0; q < P; q++) {
            const cplx<T> d = cdot(x.v[q], y.v[q]);
            acc[0][q] += (double)d.re;
            acc[1][q] += (double)d.im;
            acc[2][q] += (double)norm2(x.v[q]);
        }
    }
    double v[3];
    runs_to_lanes<3, P>(acc, v, N);
    reduce_finish_b<3, 64>(v, fin, N, gridDim.x, blockIdx.x);
}
template <typename T>
void launch_b_dots(const cplx<T>* a, const cplx<T>* b, size_t n, const FinB& fin, int N, hipStream_t st) {
    run_dispatch<T>(N, [&](auto pc) {
        constexpr int P = decltype(pc)::value;
        k_b_dots<T, P><<<b_grid(n * N / P, kRedGridVec), kBlock, 0, st>>>(a, b, n * N / P, fin, N);
    });
    SMG_LAUNCH_CHECK();
}

// <V_i, w> for i < NV per right-hand side, runs of P right-hand sides per work-item
// (16-byte accesses in FP32); NV is a compile-time count so that only its 2 NV sums
// are held
template <typename T, int NV, int P>
__global__ __launch_bounds__(kBlock) void k_b_multi_dot(VecPtrs<T> V, const cplx<T>* __restrict__ w, size_t runs,
                                                        FinB fin, int N) {
    double acc[2 * NV][P] = {};
    for (size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x; t < runs; t += (size_t)gridDim.x * kBlock) {
        RhsRun<T, P> we;
        we.load(w + t * P);
#pragma unroll
        for (int i = 0; i < NV; i++) {
            RhsRun<T, P> ve;
            ve.load(V.v[i] + t * P);
#pragma unroll
            for (int q = 0; q < P; q++) {
                const cplx<T> d = cdot(ve.v[q], we.v[q]);
                acc[2 * i][q] += (double)d.re;
                acc[2 * i + 1][q] += (double)d.im;
            }
        }
    }
    double v[2 * NV];
    runs_to_lanes<2 * NV, P>(acc, v, N);
    reduce_finish_b<2 * NV, 32>(v, fin, N, gridDim.x, blockIdx.x);
}
template <typename T, int NV>
static void multi_dot_nv(const VecPtrs<T>& V, int nv, const cplx<T>* w, size_t n, const FinB& fin, int N,
                         hipStream_t st) {
    if (nv == NV) {
        run_dispatch<T>(N, [&](auto pc) {
            constexpr int P = decltype(pc)::value;
            k_b_multi_dot<T, NV, P><<<b_grid(n * N / P, kRedGridMulti), kBlock, 0, st>>>(V, w, n * N / P, fin, N);
        });
        return;
    }
    if constexpr (NV > 1) multi_dot_nv<T, NV - 1>(V, nv, w, n, fin, N, st);
}
template <typename T>
void launch_b_multi_dot(const VecPtrs<T>& V, int nv, const cplx<T>* w, size_t n, const FinB& fin, int N,
                        hipStream_t st) {
    SMG_REQUIRE(nv >= 1 && nv <= 16 && N <= 32, "batched multi_dot: 1..16 vectors, N <= 32");
    SMG_REQUIRE(fin.ostride >= 2 * nv, "batched multi_dot: the output stride holds 2 nv sums per RHS");
    multi_dot_nv<T, 16>(V, nv, w, n, fin, N, st);
    SMG_LAUNCH_CHECK();
}

template <typename T>
struct ApapB {
    const double* a[16];
};
// With DOTS the new Ap's sums {<Ap, r>, |Ap|^2} per RHS are reduced into fin as well
template <typename T, int P, bool DOTS>
__global__ __launch_bounds__(kBlock) void k_b_gcr_orth(cplx<T>* __restrict__ p, cplx<T>* __restrict__ Ap,
                                                       VecPtrs<T> Pv, VecPtrs<T> APv, int nv,
                                                       const double* __restrict__ dots, ApapB<T> apap, size_t runs,
                                                       int N, const cplx<T>* __restrict__ r, FinB fin) {
    double acc[3][P] = {};
    __shared__ cplx<T> beta[16 * 64];
    for (int q = threadIdx.x; q < nv * N; q += kBlock) {
        const int i = q / N, j = q % N;
        const double d = apap.a[i][j];
        const double* dd = dots + (size_t)j * 32 + 2 * i;
        beta[i * N + j] = d > 0.0 ? cplx<T>{(T)(dd[0] / d), (T)(dd[1] / d)} : cplx<T>{(T)0, (T)0};
    }
    __syncthreads();
    for (size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x; t < runs; t += (size_t)gridDim.x * kBlock) {
        const size_t e = t * P;
        const int j = (int)(e & (N - 1));
        RhsRun<T, P> pe, ae;
        pe.load(p + e);
        ae.load(Ap + e);
        for (int i = 0; i < nv; i++) {
            RhsRun<T, P> pi, ai;
            pi.load(Pv.v[i] + e);
            ai.load(APv.v[i] + e);
#pragma unroll
            for (int q = 0; q < P; q++) {
                const cplx<T> bb = beta[i * N + j + q];
                pe.v[q] = pe.v[q] - bb * pi.v[q];
                ae.v[q] = ae.v[q] - bb * ai.v[q];
            }
        }
        pe.store(p + e);
        ae.store(Ap + e);
        if constexpr (DOTS) {
            RhsRun<T, P> re;
            re.load(r + e);
#pragma unroll
            for (int q = 0; q < P; q++) {
                const cplx<T> d = cdot(ae.v[q], re.v[q]);
                acc[0][q] += (double)d.re;
                acc[1][q] += (double)d.im;
                acc[2][q] += (double)norm2(ae.v[q]);
            }
        }
    }
    if constexpr (DOTS) {
        double v[3];
        runs_to_lanes<3, P>(acc, v, N);
        reduce_finish_b<3, 64>(v, fin, N, gridDim.x, blockIdx.x);
    }
}
template <typename T>
void launch_b_gcr_orth(cplx<T>* p, cplx<T>* Ap, const VecPtrs<T>& P, const VecPtrs<T>& AP, int nv,
                       const double* dots, const double* const* apap, size_t n, int N, hipStream_t st,
                       const cplx<T>* r, const FinB* fin) {
    SMG_REQUIRE(nv >= 1 && nv <= 16 && N <= 64, "batched GCR orthogonalisation: 1..16 directions, N <= 64");
    ApapB<T> a;
    for (int i = 0; i < nv; i++) a.a[i] = apap[i];
    run_dispatch<T>(N, [&](auto pc) {
        constexpr int RP = decltype(pc)::value;
        const size_t runs = n * N / RP;
        if (fin)
            k_b_gcr_orth<T, RP, true><<<b_grid(runs, kRedGridVec), kBlock, 0, st>>>(p, Ap, P, AP, nv, dots, a, runs,
                                                                                    N, r, *fin);
        else
            k_b_gcr_orth<T, RP, false><<<b_grid(runs, 1 << 14), kBlock, 0, st>>>(p, Ap, P, AP, nv, dots, a, runs, N,
                                                                                 r, FinB());
    });
    SMG_LAUNCH_CHECK();
}

// per-RHS complex coefficients s[2j], s[2j+1] of the run starting at RHS j
template <typename T, int P>
__device__ __forceinline__ void run_coef(const double* __restrict__ s, int j, cplx<T> (&a)[P]) {
#pragma unroll
    for (int q = 0; q < P; q++) a[q] = cplx<T>{(T)s[2 * (j + q)], (T)s[2 * (j + q) + 1]};
}

// MODE 0: x += w r, r -= w Ar; 1 (the first sweep from x = 0, r = b): x = w b,
// r = b - w Ar; 2 (the last sweep when the residual is not used again): x += w r
template <typename T, int P, int MODE>
__global__ __launch_bounds__(kBlock) void k_b_mr_update(cplx<T>* __restrict__ x, cplx<T>* __restrict__ r,
                                                        const cplx<T>* __restrict__ Ar, const cplx<T>* __restrict__ b,
                                                        const double* __restrict__ s, size_t runs, int N) {
    for (size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x; t < runs; t += (size_t)gridDim.x * kBlock) {
        const size_t i = t * P;
        cplx<T> w[P];
        run_coef<T, P>(s, (int)(i & (N - 1)), w);
        RhsRun<T, P> xv, rv, av;
        if (MODE != 1) xv.load(x + i);
        rv.load((MODE == 1 ? b : r) + i);
        if (MODE != 2) av.load(Ar + i);
#pragma unroll
        for (int q = 0; q < P; q++) {
            xv.v[q] = MODE == 1 ? w[q] * rv.v[q] : xv.v[q] + w[q] * rv.v[q];
            if (MODE != 2) rv.v[q] = rv.v[q] - w[q] * av.v[q];
        }
        xv.store(x + i);
        if (MODE != 2) rv.store(r + i);
    }
}
template <typename T>
void launch_b_mr_update(cplx<T>* x, cplx<T>* r, const cplx<T>* Ar, const double* s, size_t n, int N, hipStream_t st,
                        int mode, const cplx<T>* b) {
    SMG_REQUIRE(mode >= 0 && mode <= 2 && (mode != 1 || b), "MR update: mode 0, 1 (with b) or 2");
    run_dispatch<T>(N, [&](auto pc) {
        constexpr int P = decltype(pc)::value;
        const int g = b_grid(n * N / P, 1 << 16);
        if (mode == 0) k_b_mr_update<T, P, 0><<<g, kBlock, 0, st>>>(x, r, Ar, b, s, n * N / P, N);
        else if (mode == 1) k_b_mr_update<T, P, 1><<<g, kBlock, 0, st>>>(x, r, Ar, b, s, n * N / P, N);
        else k_b_mr_update<T, P, 2><<<g, kBlock, 0, st>>>(x, r, Ar, b, s, n * N / P, N);
    });
    SMG_LAUNCH_CHECK();
}

template <typename T, int P>
__global__ __launch_bounds__(kBlock) void k_b_update(cplx<T>* __restrict__ x, cplx<T>* __restrict__ r,
                                                     const cplx<T>* __restrict__ p, const cplx<T>* __restrict__ Ap,
                                                     const double* __restrict__ s, size_t runs, FinB fin, int N) {
    if (fin.skip && *fin.skip) return;   // CG finished for every RHS
    double acc[1][P] = {};
    for (size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x; t < runs; t += (size_t)gridDim.x * kBlock) {
        const size_t i = t * P;
        cplx<T> a[P];
        run_coef<T, P>(s, (int)(i & (N - 1)), a);
        RhsRun<T, P> xv, rv, pv, av;
        rv.load(r + i);
        av.load(Ap + i);
        if (x) {   // without x, the caller forms it later from the directions
            xv.load(x + i);
            pv.load(p + i);
        }
#pragma unroll
        for (int q = 0; q < P; q++) {
            if (x) xv.v[q] = xv.v[q] + a[q] * pv.v[q];
            rv.v[q] = rv.v[q] - a[q] * av.v[q];
            acc[0][q] += (double)norm2(rv.v[q]);
        }
        if (x) xv.store(x + i);
        rv.store(r + i);
    }
    double v[1];
    runs_to_lanes<1, P>(acc, v, N);
    reduce_finish_b<1, 64>(v, fin, N, gridDim.x, blockIdx.x);
}
template <typename T>
void launch_b_update(cplx<T>* x, cplx<T>* r, const cplx<T>* p, const cplx<T>* Ap, const double* s, size_t n,
                     const FinB& fin, int N, hipStream_t st) {
    run_dispatch<T>(N, [&](auto pc) {
        constexpr int P = decltype(pc)::value;
        k_b_update<T, P><<<b_grid(n * N / P, kRedGridVec), kBlock, 0, st>>>(x, r, p, Ap, s, n * N / P, fin, N);
    });
    SMG_LAUNCH_CHECK();
}

// x (+)= sum_{i < m} a_i,j P_i (a_i,j complex at coef.a[i][2j..]): the batched GCR's
// deferred iterate update
template <typename T>
struct CoefB {
    const double* a[16];
};
template <typename T, int P>
__global__ __launch_bounds__(kBlock) void k_b_combine(cplx<T>* __restrict__ x, VecPtrs<T> Pv, CoefB<T> coef, int m,
                                                      int accumulate, size_t runs, int N) {
    for (size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x; t < runs; t += (size_t)gridDim.x * kBlock) {
        const size_t e = t * P;
        const int j = (int)(e & (N - 1));
        RhsRun<T, P> xv;
        if (accumulate) xv.load(x + e);
        else
#pragma unroll
            for (int q = 0; q < P; q++) xv.v[q] = cplx<T>{(T)0, (T)0};
        for (int i = 0; i < m; i++) {
            cplx<T> a[P];
            run_coef<T, P>(coef.a[i], j, a);
            RhsRun<T, P> pv;
            pv.load(Pv.v[i] + e);
#pragma unroll
            for (int q = 0; q < P; q++) xv.v[q] = xv.v[q] + a[q] * pv.v[q];
        }
        xv.store(x + e);
    }
}
template <typename T>
void launch_b_combine(cplx<T>* x, const VecPtrs<T>& P, const double* const* coef, int m, int accumulate, size_t n,
                      int N, hipStream_t st) {
    SMG_REQUIRE(m >= 1 && m <= 16, "batched combination: 1..16 vectors");
    CoefB<T> c;
    for (int i = 0; i < m; i++) c.a[i] = coef[i];
    run_dispatch<T>(N, [&](auto pc) {
        constexpr int RP = decltype(pc)::value;
        k_b_combine<T, RP><<<b_grid(n * N / RP, 1 << 14), kBlock, 0, st>>>(x, P, c, m, accumulate, n * N / RP, N);
    });
    SMG_LAUNCH_CHECK();
}

template <typename T, int P>
__global__ void k_b_xpay(cplx<T>* __restrict__ p, const cplx<T>* __restrict__ z, const double* __restrict__ s,
                         size_t runs, int N, const int* __restrict__ skip) {
    if (skip && *skip) return;
    for (size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x; t < runs; t += (size_t)gridDim.x * kBlock) {
        const size_t i = t * P;
        cplx<T> b[P];
        run_coef<T, P>(s, (int)(i & (N - 1)), b);
        RhsRun<T, P> pv, zv;
        pv.load(p + i);
        zv.load(z + i);
#pragma unroll
        for (int q = 0; q < P; q++) pv.v[q] = zv.v[q] + b[q] * pv.v[q];
        pv.store(p + i);
    }
}
template <typename T>
void launch_b_xpay(cplx<T>* p, const cplx<T>* z, const double* s, size_t n, int N, hipStream_t st,
                   const int* skip) {
    run_dispatch<T>(N, [&](auto pc) {
        constexpr int P = decltype(pc)::value;
        k_b_xpay<T, P><<<b_grid(n * N / P, 1 << 16), kBlock, 0, st>>>(p, z, s, n * N / P, N, skip);
    });
    SMG_LAUNCH_CHECK();
}

// per-RHS CG program state: aux_j[5] = tol^2, flag_j = 0; the solve's done and
// late flags cleared
__global__ void k_b_cg_init(double* aux, int astride, int N, double tol2, int* done, int* late) {
    const int j = threadIdx.x;
    if (j < N) {
        aux[(size_t)j * astride + 5] = tol2;
        *reinterpret_cast<int*>(aux + (size_t)j * astride + 8) = 0;
    }
    if (j == 0 && done) *done = 0;
    if (j == 0 && late) *late = 0;
}
void launch_b_cg_init(double* aux, int astride, int N, double tol2, int* done, int* late, hipStream_t st) {
    k_b_cg_init<<<1, 64, 0, st>>>(aux, astride, N, tol2, done, late);
    SMG_LAUNCH_CHECK();
}

// The two vector passes of a CG iteration after the apply (src/coarse_op.h:520-540):
// k_b_cg_r: r -= a_j Ap with |r|^2 finished by fin (FIN_CG_BETA, which sets
// *fin.done once every RHS has converged); k_b_cg_xp: x += a_j p, p = r + b_j p, the
// solution update taken in the pass that reads p anyway (8 vector passes per
// iteration instead of 9).  An RHS that has converged has a_j = 0 from then on.
// Once the solve is done, k_b_cg_r of the next iteration finds *fin.skip set,
// returns at entry and sets *late, which stops k_b_cg_xp: the iteration in which
// the last RHS converged still adds its step to x.
template <typename T, int P>
__global__ __launch_bounds__(kBlock) void k_b_cg_r(cplx<T>* __restrict__ r, const cplx<T>* __restrict__ Ap,
                                                   const double* __restrict__ s, size_t runs, FinB fin, int N,
                                                   int* __restrict__ late) {
    if (fin.skip && *fin.skip) {
        if (late && blockIdx.x == 0 && threadIdx.x == 0) *late = 1;
        return;
    }
    double acc[1][P] = {};
    for (size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x; t < runs; t += (size_t)gridDim.x * kBlock) {
        const size_t i = t * P;
        cplx<T> a[P];
        run_coef<T, P>(s, (int)(i & (N - 1)), a);
        RhsRun<T, P> rv, av;
        rv.load(r + i);
        av.load(Ap + i);
#pragma unroll
        for (int q = 0; q < P; q++) {
            rv.v[q] = rv.v[q] - a[q] * av.v[q];
            acc[0][q] += (double)norm2(rv.v[q]);
        }
        rv.store(r + i);
    }
    double v[1];
    runs_to_lanes<1, P>(acc, v, N);
    reduce_finish_b<1, 64>(v, fin, N, gridDim.x, blockIdx.x);
}
template <typename T, int P>
__global__ __launch_bounds__(kBlock) void k_b_cg_xp(cplx<T>* __restrict__ x, cplx<T>* __restrict__ p,
                                                    const cplx<T>* __restrict__ r, const double* __restrict__ sa,
                                                    const double* __restrict__ sb, size_t runs, int N,
                                                    const int* __restrict__ late) {
    if (late && *late) return;
    for (size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x; t < runs; t += (size_t)gridDim.x * kBlock) {
        const size_t i = t * P;
        cplx<T> a[P], b[P];
        run_coef<T, P>(sa, (int)(i & (N - 1)), a);
        run_coef<T, P>(sb, (int)(i & (N - 1)), b);
        RhsRun<T, P> xv, pv, rv;
        xv.load(x + i);
        pv.load(p + i);
        rv.load(r + i);
#pragma unroll
        for (int q = 0; q < P; q++) {
            xv.v[q] = xv.v[q] + a[q] * pv.v[q];
            pv.v[q] = rv.v[q] + b[q] * pv.v[q];
        }
        xv.store(x + i);
        pv.store(p + i);
    }
}
template <typename T>
void launch_b_cg_r(cplx<T>* r, const cplx<T>* Ap, const double* a, size_t n, const FinB& fin, int N, int* late,
                   hipStream_t st) {
    run_dispatch<T>(N, [&](auto pc) {
        constexpr int P = decltype(pc)::value;
        k_b_cg_r<T, P><<<b_grid(n * N / P, kRedGridVec), kBlock, 0, st>>>(r, Ap, a, n * N / P, fin, N, late);
    });
    SMG_LAUNCH_CHECK();
}
template <typename T>
void launch_b_cg_xp(cplx<T>* x, cplx<T>* p, const cplx<T>* r, const double* a, const double* b, size_t n, int N,
                    const int* late, hipStream_t st) {
    run_dispatch<T>(N, [&](auto pc) {
        constexpr int P = decltype(pc)::value;
        k_b_cg_xp<T, P><<<b_grid(n * N / P, 1 << 16), kBlock, 0, st>>>(x, p, r, a, b, n * N / P, N, late);
    });
    SMG_LAUNCH_CHECK();
}

// ----------------------------------------------------------------- BF16-vector CG
// The coarsest-level CG of the FP32 cycle with r, p, Ap and x stored as BF16 (the
// solve runs to a loose tolerance, 0.1 by default, which BF16's 8-bit mantissa
// resolves; FP32 arithmetic throughout): each work-item takes a run of 4 right-hand
// sides (one 16-byte access), N >= 16.  Semantics as k_b_cg_r / k_b_cg_xp.
struct HRun {
    c32 v[4];
    __device__ __forceinline__ void load(const cb16* p) {
        const uint4 w = *reinterpret_cast<const uint4*>(p);
        v[0] = unpack_h(w.x);
        v[1] = unpack_h(w.y);
        v[2] = unpack_h(w.z);
        v[3] = unpack_h(w.w);
    }
    __device__ __forceinline__ uint4 store(cb16* p) {   // returns the stored dwords
        const uint4 w{pack_h(v[0]), pack_h(v[1]), pack_h(v[2]), pack_h(v[3])};
        *reinterpret_cast<uint4*>(p) = w;
        return w;
    }
};
// r = p = bf16(b), x = 0, |r|^2 (of the stored r) finished by fin (FIN_CG_INIT)
__global__ __launch_bounds__(kBlock) void k_h_cg_init(const c32* __restrict__ b, cb16* __restrict__ r,
                                                      cb16* __restrict__ p, cb16* __restrict__ x, size_t runs,
                                                      FinB fin, int N) {
    double acc[1][4] = {};
    for (size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x; t < runs; t += (size_t)gridDim.x * kBlock) {
        const size_t i = t * 4;
        const float4 b0 = *reinterpret_cast<const float4*>(b + i), b1 = *reinterpret_cast<const float4*>(b + i + 2);
        HRun rv;
        rv.v[0] = c32{b0.x, b0.y};
        rv.v[1] = c32{b0.z, b0.w};
        rv.v[2] = c32{b1.x, b1.y};
        rv.v[3] = c32{b1.z, b1.w};
        const uint4 w = rv.store(r + i);
        *reinterpret_cast<uint4*>(p + i) = w;
        *reinterpret_cast<uint4*>(x + i) = uint4{0u, 0u, 0u, 0u};
        const unsigned ws[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
        for (int q = 0; q < 4; q++) acc[0][q] += (double)norm2(unpack_h(ws[q]));
    }
    double v[1];
    runs_to_lanes<1, 4>(acc, v, N);
    reduce_finish_b<1, 64>(v, fin, N, gridDim.x, blockIdx.x);
}
// The first Chebyshev step from x = 0: r = bf16(b), d = bf16(b / theta), x = d
__global__ __launch_bounds__(kBlock) void k_h_cheb_init(const c32* __restrict__ b, cb16* __restrict__ r,
                                                        cb16* __restrict__ d, cb16* __restrict__ x, size_t runs,
                                                        float inv_theta) {
    for (size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x; t < runs; t += (size_t)gridDim.x * kBlock) {
        const size_t i = t * 4;
        const float4 b0 = *reinterpret_cast<const float4*>(b + i), b1 = *reinterpret_cast<const float4*>(b + i + 2);
        HRun rv, dv;
        rv.v[0] = c32{b0.x, b0.y};
        rv.v[1] = c32{b0.z, b0.w};
        rv.v[2] = c32{b1.x, b1.y};
        rv.v[3] = c32{b1.z, b1.w};
        rv.store(r + i);
#pragma unroll
        for (int k = 0; k < 4; k++) dv.v[k] = inv_theta * rv.v[k];
        const uint4 w = dv.store(d + i);
        *reinterpret_cast<uint4*>(x + i) = w;
    }
}
// one CG step after the apply q = A p (EPI_CG, FIN_CG_STEP): r -= a q, x += a p,
// p = r + b p with r as stored (rounded to BF16), a_j at a[2j] and b_j at aux[16j + 6]
__global__ __launch_bounds__(kBlock) void k_h_cg_step(cb16* __restrict__ x, cb16* __restrict__ p,
                                                      cb16* __restrict__ r, const cb16* __restrict__ q,
                                                      const double* __restrict__ sa, const double* __restrict__ aux,
                                                      size_t runs, int N, const int* __restrict__ late) {
    if (late && *late) return;
    for (size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x; t < runs; t += (size_t)gridDim.x * kBlock) {
        const size_t i = t * 4;
        const int j0 = (int)(i & (N - 1));
        c32 a[4], b[4];
        run_coef<float, 4>(sa, j0, a);
#pragma unroll
        for (int k = 0; k < 4; k++) b[k] = c32{(float)aux[16 * (j0 + k) + 6], (float)aux[16 * (j0 + k) + 7]};
        HRun xv, pv, rv, qv;
        rv.load(r + i);
        qv.load(q + i);
        xv.load(x + i);
        pv.load(p + i);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            rv.v[k] = rv.v[k] - a[k] * qv.v[k];
            xv.v[k] = xv.v[k] + a[k] * pv.v[k];
        }
        const uint4 w = rv.store(r + i);
        const unsigned ws[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
        for (int k = 0; k < 4; k++) pv.v[k] = unpack_h(ws[k]) + b[k] * pv.v[k];
        xv.store(x + i);
        pv.store(p + i);
    }
}
__global__ __launch_bounds__(kBlock) void k_h_to_f32(c32* __restrict__ out, const cb16* __restrict__ in,
                                                     size_t runs) {
    const size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= runs) return;
    HRun v;
    v.load(in + 4 * t);
    *reinterpret_cast<float4*>(out + 4 * t) = float4{v.v[0].re, v.v[0].im, v.v[1].re, v.v[1].im};
    *reinterpret_cast<float4*>(out + 4 * t + 2) = float4{v.v[2].re, v.v[2].im, v.v[3].re, v.v[3].im};
}
__global__ __launch_bounds__(kBlock) void k_h_from_f32(cb16* __restrict__ out, const c32* __restrict__ in,
                                                       size_t runs) {
    const size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= runs) return;
    const float4 a = *reinterpret_cast<const float4*>(in + 4 * t), b = *reinterpret_cast<const float4*>(in + 4 * t + 2);
    *reinterpret_cast<uint4*>(out + 4 * t) =
        uint4{pack_h(c32{a.x, a.y}), pack_h(c32{a.z, a.w}), pack_h(c32{b.x, b.y}), pack_h(c32{b.z, b.w})};
}
void launch_h_from_f32(cb16* out, const c32* in, size_t n, int N, hipStream_t st) {
    SMG_REQUIRE(N % 4 == 0, "BF16 conversion: N a multiple of 4");
    const size_t runs = n * N / 4;
    k_h_from_f32<<<(int)((runs + kBlock - 1) / kBlock), kBlock, 0, st>>>(out, in, runs);
    SMG_LAUNCH_CHECK();
}
void launch_h_cg_init(const c32* b, cb16* r, cb16* p, cb16* x, size_t n, const FinB& fin, int N, hipStream_t st) {
    SMG_REQUIRE(N >= 16 && N <= 64 && (N & (N - 1)) == 0, "BF16 CG: N in {16, 32, 64}");
    k_h_cg_init<<<b_grid(n * N / 4, kRedGridVec), kBlock, 0, st>>>(b, r, p, x, n * N / 4, fin, N);
    SMG_LAUNCH_CHECK();
}
void launch_h_cheb_init(const c32* b, cb16* r, cb16* d, cb16* x, size_t n, int N, double theta, hipStream_t st) {
    SMG_REQUIRE(N >= 16 && N <= 64 && (N & (N - 1)) == 0, "BF16 Chebyshev: N in {16, 32, 64}");
    SMG_REQUIRE(theta > 0.0, "BF16 Chebyshev: the interval's centre must be positive");
    k_h_cheb_init<<<b_grid(n * N / 4, 1 << 16), kBlock, 0, st>>>(b, r, d, x, n * N / 4, (float)(1.0 / theta));
    SMG_LAUNCH_CHECK();
}
void launch_h_cg_step(cb16* x, cb16* p, cb16* r, const cb16* q, const double* a, const double* aux, size_t n, int N,
                      const int* late, hipStream_t st) {
    SMG_REQUIRE(N >= 16 && N <= 64 && (N & (N - 1)) == 0, "BF16 CG: N in {16, 32, 64}");
    // the tuner's workgroup count for this shape, else the whole vector in one sweep
    const int choice = tune_lookup(TuneKey{TUNE_CG_UPDATE, (int)n, N, 0, 0, 0});
    const int cap = choice < 0 ? 1 << 16 : kTuneGrids[choice];
    k_h_cg_step<<<b_grid(n * N / 4, cap), kBlock, 0, st>>>(x, p, r, q, a, aux, n * N / 4, N, late);
    SMG_LAUNCH_CHECK();
}
void launch_h_to_f32(c32* out, const cb16* in, size_t n, int N, hipStream_t st) {
    SMG_REQUIRE(N % 4 == 0, "BF16 conversion: N a multiple of 4");
    const size_t runs = n * N / 4;
    k_h_to_f32<<<(int)((runs + kBlock - 1) / kBlock), kBlock, 0, st>>>(out, in, runs);
    SMG_LAUNCH_CHECK();
}

template <typename T>
__global__ void k_b_axpby_real(cplx<T>* __restrict__ y, const cplx<T>* __restrict__ x, T a, T c, size_t items) {
    for (size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x; t < items; t += (size_t)gridDim.x * kBlock)
        y[t] = c == (T)0 ? a * x[t] : a * x[t] + c * y[t];   // c = 0 does not read y
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
                                      hipStream_t, const cplx<T>*, double);                                      \
    template void launch_b_restrict<T>(const AggView&, const cplx<T>*, const cplx<T>*, cplx<T>*, int, hipStream_t); \
    template void launch_b_dense_apply<T>(const cplx<T>*, const cplx<T>*, cplx<T>*, size_t, int, cplx<T>*, \
                                           hipStream_t);                                                   \
    template void launch_b_norm2<T>(const cplx<T>*, size_t, const FinB&, int, hipStream_t);                     \
    template void launch_b_dots<T>(const cplx<T>*, const cplx<T>*, size_t, const FinB&, int, hipStream_t);      \
    template void launch_b_multi_dot<T>(const VecPtrs<T>&, int, const cplx<T>*, size_t, const FinB&, int,        \
                                        hipStream_t);                                                            \
    template void launch_b_gcr_orth<T>(cplx<T>*, cplx<T>*, const VecPtrs<T>&, const VecPtrs<T>&, int,            \
                                       const double*, const double* const*, size_t, int, hipStream_t,            \
                                       const cplx<T>*, const FinB*);                                             \
    template void launch_b_mr_update<T>(cplx<T>*, cplx<T>*, const cplx<T>*, const double*, size_t, int,          \
                                        hipStream_t, int, const cplx<T>*);                                       \
    template void launch_b_combine<T>(cplx<T>*, const VecPtrs<T>&, const double* const*, int, int, size_t, int,   \
                                      hipStream_t);                                                              \
    template void launch_b_update<T>(cplx<T>*, cplx<T>*, const cplx<T>*, const cplx<T>*, const double*, size_t,  \
                                     const FinB&, int, hipStream_t);                                             \
    template void launch_b_axpby_real<T>(cplx<T>*, const cplx<T>*, double, double, size_t, hipStream_t);    \
    template void launch_b_xpay<T>(cplx<T>*, const cplx<T>*, const double*, size_t, int, hipStream_t, const int*); \
    template void launch_b_cg_r<T>(cplx<T>*, const cplx<T>*, const double*, size_t, const FinB&, int, int*,        \
                                   hipStream_t);                                                                    \
    template void launch_b_cg_xp<T>(cplx<T>*, cplx<T>*, const cplx<T>*, const double*, const double*, size_t, int, \
                                    const int*, hipStream_t);
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
