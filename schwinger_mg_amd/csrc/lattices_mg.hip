// lattices_mg.hip -- the multigrid cycle and GCR over many lattices of different sizes,
// one launch per kernel for all of them.
//
// Every level of every lattice's hierarchy is a segment of that level's concatenated
// vector (SegDesc): lattice g's nx x nx sites of nd dof at complex offset off, its
// coarse links (levels >= 1) and its block vectors for the transfer to the next level.
// A launch gives lattice g the workgroups wg0_g .. wg0_g + nwg_g - 1, which stride over
// the lattice's work-items, so reductions finish per lattice (finish_seg) and every
// lattice keeps its own smoother, CG and GCR scalars.  The transfers and the coarse
// operator are the single-lattice ones (aggregation maps of src/prolongator.h:35-108,
// the 9-point stencil of src/coarse_op.h:282-297) at the lattice's offsets; FP64
// throughout.
#include "smg_lattices.h"

namespace smg {

namespace {
// this workgroup's lattice, its first work-item and the stride of its loop
struct SegRange {
    int g;
    long long e0, step;
};
__device__ __forceinline__ SegRange seg_range(const SegDesc* __restrict__ d, int G) {
    const int g = my_lattice(d, G);
    return SegRange{g, (long long)((int)blockIdx.x - d[g].wg0) * kBlock + threadIdx.x,
                    (long long)d[g].nwg * kBlock};
}
}  // namespace

// c[b*k + j] = sum over the aggregate's sites of sum_c conj(v_{b,j}(loc, c)) f(s, c)
// (src/prolongator.h:94-108), the site's dof summed first as in the single-lattice
// restrict; chiral: block vector j vanishes off its chirality
__global__ __launch_bounds__(kBlock) void k_seg_restrict(const SegDesc* __restrict__ fd,
                                                         const SegDesc* __restrict__ cd, int G,
                                                         const c64* __restrict__ bv, const c64* __restrict__ fine,
                                                         c64* __restrict__ coarse, int bx, int by, int chiral) {
    const SegRange R = seg_range(cd, G);
    const SegDesc C = cd[R.g], F = fd[R.g];
    const int k = C.nd, nd = F.nd, bsize = bx * by * nd;
    for (long long e = R.e0; e < C.n; e += R.step) {
        const int b = (int)(e / k), j = (int)(e % k);
        const int bxi = b % C.nx, byi = b / C.nx;
        const int c0 = chiral ? (j >= k / 2 ? nd / 2 : 0) : 0, c1 = chiral ? c0 + nd / 2 : nd;
        const c64* v = bv + F.bv_off + ((size_t)b * k + j) * bsize;
        const c64* f = fine + F.off;
        c64 acc{0.0, 0.0};
        for (int loc = 0; loc < bx * by; loc++) {
            const int lx = loc % bx, ly = loc / bx;
            const size_t s = (size_t)(bxi * bx + lx) + (size_t)F.nx * (byi * by + ly);
            c64 a = cdot(v[loc * nd + c0], f[s * nd + c0]);
            for (int c = c0 + 1; c < c1; c++) a = a + cdot(v[loc * nd + c], f[s * nd + c]);
            acc = acc + a;
        }
        coarse[C.off + e] = acc;
    }
}

// f(s, c) (+)= sum_j v_{b,j}(loc, c) c[b*k + j] (src/prolongator.h:77-90); chiral: only
// the coarse dof of this dof's chirality contribute
__global__ __launch_bounds__(kBlock) void k_seg_prolong(const SegDesc* __restrict__ fd,
                                                        const SegDesc* __restrict__ cd, int G,
                                                        const c64* __restrict__ bv, const c64* __restrict__ coarse,
                                                        c64* __restrict__ fine, int accumulate, int bx, int by,
                                                        int chiral) {
    const SegRange R = seg_range(fd, G);
    const SegDesc F = fd[R.g], C = cd[R.g];
    const int nd = F.nd, k = C.nd, bsize = bx * by * nd;
    for (long long e = R.e0; e < F.n; e += R.step) {
        const int s = (int)(e / nd), c = (int)(e % nd);
        const int x = s % F.nx, y = s / F.nx;
        const int b = x / bx + C.nx * (y / by), loc = x % bx + bx * (y % by);
        const c64* v = bv + F.bv_off + (size_t)b * k * bsize + (size_t)loc * nd + c;
        const c64* cc = coarse + C.off + (size_t)b * k;
        const int j0 = chiral ? (c >= nd / 2 ? k / 2 : 0) : 0, j1 = chiral ? j0 + k / 2 : k;
        c64 acc{0.0, 0.0};
        for (int j = j0; j < j1; j++) acc = acc + cc[j] * v[(size_t)j * bsize];
        c64* o = fine + F.off + e;
        *o = accumulate ? *o + acc : acc;
    }
}

// dst[b, ki] = sum_d sum_kj L[b][d][ki][kj] src[nbr(b, d), kj] (src/coarse_op.h:282-297)
// with a run of kSegRun (= 4, or K if smaller) consecutive link entries (b, ki, kj..) per
// lane: one 16-byte load per direction for FP16 links, K / kSegRun lanes per row, whose
// sums over kj end in a shuffle tree inside the row's lanes.  The workgroups of a lattice
// stride over its nx*nx*K*K/kSegRun lanes in steps of a multiple of kBlock (so of the
// row's lanes): every row stays inside one wave and one step, and the loop bound is
// uniform per workgroup.  EPI_RESID: dst = bvec - A src; EPI_DOTS: {<dst, aux>, |dst|^2}
// per lattice into fin.  TL: the links' storage (c64, or ch16: FP16 links widened on
// load, FP64 arithmetic).
template <int EPI, int K, typename TL>
__global__ __launch_bounds__(kBlock) void k_seg_coarse(const SegDesc* __restrict__ ld, int G,
                                                       const TL* __restrict__ links, const c64* __restrict__ src,
                                                       c64* __restrict__ dst, const c64* __restrict__ bvec,
                                                       const c64* __restrict__ aux, LatFin fin) {
    constexpr int V = seg_coarse_run(K), RL = K / V, KK = K * K, RK = K * RL;
    static_assert(kBlock % RL == 0 && RL <= 64, "a row's lanes must sit inside one wave");
    if (EPI == EPI_DOTS && fin.skip && *fin.skip) return;   // grid-uniform
    const int g = my_lattice(ld, G);
    const SegDesc C = ld[g];
    const int n = C.nx, nb = n * n;
    double acc[3] = {0.0, 0.0, 0.0};
    for (long long t0 = (long long)((int)blockIdx.x - C.wg0) * kBlock; t0 < C.n; t0 += (long long)C.nwg * kBlock) {
        const long long t = t0 + threadIdx.x;
        const int b = (int)(t / RK), e = (int)(t % RK), ki = e / RL, kj0 = (e % RL) * V;
        c64 sum{0.0, 0.0};
        if (b < nb) {
            const int bx = b % n, by = b / n;
            const int xm = bx == 0 ? n - 1 : bx - 1, xp = bx + 1 == n ? 0 : bx + 1;
            const int ym = by == 0 ? n - 1 : by - 1, yp = by + 1 == n ? 0 : by + 1;
            const int nbr[9] = {b, xp + by * n, xm + by * n, bx + yp * n, bx + ym * n,
                                xp + yp * n, xp + ym * n, xm + yp * n, xm + ym * n};
            const TL* lk = links + C.link_off + (size_t)b * 9 * KK + ki * K + kj0;
            const c64* s = src + C.off + kj0;
#pragma unroll
            for (int d = 0; d < 9; d++) {
                TL lv[V];
                if constexpr (std::is_same<TL, ch16>::value && V == 4) {
                    const uint4 w = *reinterpret_cast<const uint4*>(lk + d * KK);
                    const unsigned ws[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
                    for (int v = 0; v < 4; v++) lv[v] = __builtin_bit_cast(ch16, ws[v]);
                } else {
#pragma unroll
                    for (int v = 0; v < V; v++) lv[v] = lk[d * KK + v];
                }
                const c64* sv = s + (size_t)nbr[d] * K;
#pragma unroll
                for (int v = 0; v < V; v++) sum = sum + cvt<double>(load_c(&lv[v])) * sv[v];
            }
        }
#pragma unroll
        for (int off = RL / 2; off > 0; off >>= 1) {
            sum.re += __shfl_xor(sum.re, off, 64);
            sum.im += __shfl_xor(sum.im, off, 64);
        }
        if (b < nb && kj0 == 0) {
            const size_t o = C.off + (size_t)b * K + ki;
            c64 out = sum;
            if (EPI == EPI_RESID) out = bvec[o] - out;
            dst[o] = out;
            if (EPI == EPI_DOTS) {
                const c64 p = cdot(out, aux[o]);
                acc[0] += p.re;
                acc[1] += p.im;
                acc[2] += norm2(out);
            }
        }
    }
    if (EPI == EPI_DOTS) finish_seg<3>(acc, fin, ld, G, g);
}

// vector passes with per-lattice scalars (complex a_g at s[2g], s[2g + 1])
// MR step (src/smoother.cpp:34-40): x += w r, r -= w Ar
__global__ __launch_bounds__(kBlock) void k_seg_mr_update(const SegDesc* __restrict__ d, int G, c64* __restrict__ x,
                                                          c64* __restrict__ r, const c64* __restrict__ Ar,
                                                          const double* __restrict__ w) {
    const SegRange R = seg_range(d, G);
    const SegDesc D = d[R.g];
    const c64 a{w[2 * R.g], w[2 * R.g + 1]};
    for (long long e = R.e0; e < D.n; e += R.step) {
        const size_t i = D.off + e;
        const c64 rv = r[i];
        x[i] = x[i] + a * rv;
        r[i] = rv - a * Ar[i];
    }
}
// CG (src/coarse_op.h:508-541): x = 0, r = p = b, {|r|^2}
__global__ __launch_bounds__(kBlock) void k_seg_cg_init(const SegDesc* __restrict__ d, int G,
                                                        const c64* __restrict__ b, c64* __restrict__ x,
                                                        c64* __restrict__ r, c64* __restrict__ p, LatFin fin) {
    const SegRange R = seg_range(d, G);
    const SegDesc D = d[R.g];
    double acc[1] = {0.0};
    for (long long e = R.e0; e < D.n; e += R.step) {
        const size_t i = D.off + e;
        const c64 bv = b[i];
        r[i] = bv;
        p[i] = bv;
        x[i] = c64{0.0, 0.0};
        acc[0] += norm2(bv);
    }
    finish_seg<1>(acc, fin, d, G, R.g);
}
// r -= alpha Ap, {|r|^2}; once every lattice has stopped, marks *late and does nothing
__global__ __launch_bounds__(kBlock) void k_seg_cg_r(const SegDesc* __restrict__ d, int G, c64* __restrict__ r,
                                                     const c64* __restrict__ Ap, const double* __restrict__ alpha,
                                                     LatFin fin, int* __restrict__ late) {
    if (fin.skip && *fin.skip) {
        if (late && blockIdx.x == 0 && threadIdx.x == 0) *late = 1;
        return;
    }
    const SegRange R = seg_range(d, G);
    const SegDesc D = d[R.g];
    const c64 a{alpha[2 * R.g], alpha[2 * R.g + 1]};
    double acc[1] = {0.0};
    for (long long e = R.e0; e < D.n; e += R.step) {
        const size_t i = D.off + e;
        const c64 rv = r[i] - a * Ap[i];
        r[i] = rv;
        acc[0] += norm2(rv);
    }
    finish_seg<1>(acc, fin, d, G, R.g);
}
// x += alpha p; p = r + beta p on lattices whose CG goes on (stop flag of cg_aux clear;
// a stopped lattice's alpha is zero, so its x stays as it is)
__global__ __launch_bounds__(kBlock) void k_seg_cg_xp(const SegDesc* __restrict__ d, int G, c64* __restrict__ x,
                                                      c64* __restrict__ p, const c64* __restrict__ r,
                                                      const double* __restrict__ alpha,
                                                      const double* __restrict__ beta,
                                                      const double* __restrict__ cg_aux, const int* __restrict__ late) {
    if (late && *late) return;
    const SegRange R = seg_range(d, G);
    const SegDesc D = d[R.g];
    const c64 a{alpha[2 * R.g], alpha[2 * R.g + 1]}, bt{beta[2 * R.g], beta[2 * R.g + 1]};
    const bool go = *reinterpret_cast<const int*>(cg_aux + (size_t)R.g * 16 + 8) == 0;
    for (long long e = R.e0; e < D.n; e += R.step) {
        const size_t i = D.off + e;
        const c64 pv = p[i];
        x[i] = x[i] + a * pv;
        if (go) p[i] = r[i] + bt * pv;
    }
}
// GCR passes: {<V_i, w>} for i < NV (NC = 2 NV)
template <int NV>
__global__ __launch_bounds__(kBlock) void k_seg_multi_dot(const SegDesc* __restrict__ d, int G, SegVecs V,
                                                          const c64* __restrict__ w, LatFin fin) {
    const SegRange R = seg_range(d, G);
    const SegDesc D = d[R.g];
    double acc[2 * NV];
#pragma unroll
    for (int c = 0; c < 2 * NV; c++) acc[c] = 0.0;
    for (long long e = R.e0; e < D.n; e += R.step) {
        const size_t i = D.off + e;
        const c64 wv = w[i];
#pragma unroll
        for (int v = 0; v < NV; v++) {
            const c64 p = cdot(V.p[v][i], wv);
            acc[2 * v] += p.re;
            acc[2 * v + 1] += p.im;
        }
    }
    finish_seg<2 * NV>(acc, fin, d, G, R.g);
}
// z -= sum_i beta_i P_i, Az -= sum_i beta_i AP_i with beta_i = <AP_i, Az> / |AP_i|^2 on
// lattice g (classical Gram-Schmidt against the kept directions, src/solvers.cpp:260-266;
// dots[g][i] from the unmodified Az)
__global__ __launch_bounds__(kBlock) void k_seg_gcr_orth(const SegDesc* __restrict__ d, int G, c64* __restrict__ z,
                                                         c64* __restrict__ Az, SegVecs P, SegVecs AP, int nv,
                                                         const double* __restrict__ dots, SegRows apap) {
    const SegRange R = seg_range(d, G);
    const SegDesc D = d[R.g];
    c64 bt[kSegWindow];
    for (int v = 0; v < nv; v++) {
        const double den = apap.p[v][R.g];
        const c64 dv{dots[(size_t)R.g * 2 * kSegWindow + 2 * v], dots[(size_t)R.g * 2 * kSegWindow + 2 * v + 1]};
        // a lattice whose direction collapsed (|AP|^2 < 1e-30, zero step) projects nothing
        bt[v] = den >= 1e-30 ? c64{dv.re / den, dv.im / den} : c64{0.0, 0.0};
    }
    for (long long e = R.e0; e < D.n; e += R.step) {
        const size_t i = D.off + e;
        c64 zv = z[i], av = Az[i];
        for (int v = 0; v < nv; v++) {
            zv = zv - bt[v] * P.p[v][i];
            av = av - bt[v] * AP.p[v][i];
        }
        z[i] = zv;
        Az[i] = av;
    }
}
// {Re<a, b>, Im<a, b>, |a|^2} per lattice
__global__ __launch_bounds__(kBlock) void k_seg_dots(const SegDesc* __restrict__ d, int G, const c64* __restrict__ a,
                                                     const c64* __restrict__ b, LatFin fin) {
    const SegRange R = seg_range(d, G);
    const SegDesc D = d[R.g];
    double acc[3] = {0.0, 0.0, 0.0};
    for (long long e = R.e0; e < D.n; e += R.step) {
        const size_t i = D.off + e;
        const c64 av = a[i], p = cdot(av, b[i]);
        acc[0] += p.re;
        acc[1] += p.im;
        acc[2] += norm2(av);
    }
    finish_seg<3>(acc, fin, d, G, R.g);
}
// x += a_g z, r -= a_g Az on lattices with active[g] != 0; {|r|^2} per lattice
__global__ __launch_bounds__(kBlock) void k_seg_update(const SegDesc* __restrict__ d, int G, c64* __restrict__ x,
                                                       c64* __restrict__ r, const c64* __restrict__ z,
                                                       const c64* __restrict__ Az, const double* __restrict__ alpha,
                                                       const int* __restrict__ active, LatFin fin) {
    const SegRange R = seg_range(d, G);
    const SegDesc D = d[R.g];
    const bool on = active[R.g] != 0;
    const c64 a{alpha[2 * R.g], alpha[2 * R.g + 1]};
    double acc[1] = {0.0};
    for (long long e = R.e0; e < D.n; e += R.step) {
        const size_t i = D.off + e;
        c64 rv = r[i];
        if (on) {
            x[i] = x[i] + a * z[i];
            rv = rv - a * Az[i];
            r[i] = rv;
        }
        acc[0] += norm2(rv);
    }
    finish_seg<1>(acc, fin, d, G, R.g);
}

// ------------------------------------------------------------------ launchers
void launch_seg_restrict(const SegDesc* fd, const SegDesc* cd, int G, int nwg_c, const c64* bv, const c64* fine,
                         c64* coarse, int bx, int by, int chiral, hipStream_t st) {
    k_seg_restrict<<<nwg_c, kBlock, 0, st>>>(fd, cd, G, bv, fine, coarse, bx, by, chiral);
    SMG_LAUNCH_CHECK();
}
void launch_seg_prolong(const SegDesc* fd, const SegDesc* cd, int G, int nwg_f, const c64* bv, const c64* coarse,
                        c64* fine, int accumulate, int bx, int by, int chiral, hipStream_t st) {
    k_seg_prolong<<<nwg_f, kBlock, 0, st>>>(fd, cd, G, bv, coarse, fine, accumulate, bx, by, chiral);
    SMG_LAUNCH_CHECK();
}
template <int EPI, typename TL>
static void seg_coarse_k(const SegDesc* ld, int G, int nwg, int K, const TL* links, const c64* src, c64* dst,
                         const c64* bvec, const c64* aux, const LatFin& fin, hipStream_t st) {
    switch (K) {
#define SMG_SK(k)                                                                                      \
    case k:                                                                                            \
        k_seg_coarse<EPI, k, TL><<<nwg, kBlock, 0, st>>>(ld, G, links, src, dst, bvec, aux, fin);      \
        break;
        SMG_SK(2) SMG_SK(4) SMG_SK(8) SMG_SK(16) SMG_SK(32) SMG_SK(64)
#undef SMG_SK
    default:
        throw Error(-1, "many-lattice coarse operator: k must be a power of two in 2..64");
    }
}
template <typename TL>
void launch_seg_coarse(const SegDesc* lanes, int G, int nwg_lanes, int K, const TL* links, const c64* src, c64* dst,
                       Epi epi, const c64* bvec, const c64* aux, const LatFin& fin, hipStream_t st) {
    if (epi == EPI_NONE) seg_coarse_k<EPI_NONE>(lanes, G, nwg_lanes, K, links, src, dst, bvec, aux, fin, st);
    else if (epi == EPI_RESID) seg_coarse_k<EPI_RESID>(lanes, G, nwg_lanes, K, links, src, dst, bvec, aux, fin, st);
    else seg_coarse_k<EPI_DOTS>(lanes, G, nwg_lanes, K, links, src, dst, bvec, aux, fin, st);
    SMG_LAUNCH_CHECK();
}
template void launch_seg_coarse<c64>(const SegDesc*, int, int, int, const c64*, const c64*, c64*, Epi, const c64*,
                                     const c64*, const LatFin&, hipStream_t);
template void launch_seg_coarse<ch16>(const SegDesc*, int, int, int, const ch16*, const c64*, c64*, Epi, const c64*,
                                      const c64*, const LatFin&, hipStream_t);
void launch_seg_mr_update(const SegDesc* d, int G, int nwg, c64* x, c64* r, const c64* Ar, const double* w,
                          hipStream_t st) {
    k_seg_mr_update<<<nwg, kBlock, 0, st>>>(d, G, x, r, Ar, w);
    SMG_LAUNCH_CHECK();
}
void launch_seg_cg_init(const SegDesc* d, int G, int nwg, const c64* b, c64* x, c64* r, c64* p, const LatFin& fin,
                        hipStream_t st) {
    k_seg_cg_init<<<nwg, kBlock, 0, st>>>(d, G, b, x, r, p, fin);
    SMG_LAUNCH_CHECK();
}
void launch_seg_cg_r(const SegDesc* d, int G, int nwg, c64* r, const c64* Ap, const double* alpha, const LatFin& fin,
                     int* late, hipStream_t st) {
    k_seg_cg_r<<<nwg, kBlock, 0, st>>>(d, G, r, Ap, alpha, fin, late);
    SMG_LAUNCH_CHECK();
}
void launch_seg_cg_xp(const SegDesc* d, int G, int nwg, c64* x, c64* p, const c64* r, const double* alpha,
                      const double* beta, const double* cg_aux, const int* late, hipStream_t st) {
    k_seg_cg_xp<<<nwg, kBlock, 0, st>>>(d, G, x, p, r, alpha, beta, cg_aux, late);
    SMG_LAUNCH_CHECK();
}
void launch_seg_multi_dot(const SegDesc* d, int G, int nwg, const SegVecs& V, int nv, const c64* w, const LatFin& fin,
                          hipStream_t st) {
    SMG_REQUIRE(nv >= 1 && nv <= kSegWindow, "multi-dot over 1..kSegWindow vectors");
    switch (nv) {
#define SMG_MD(n)                                                        \
    case n:                                                              \
        k_seg_multi_dot<n><<<nwg, kBlock, 0, st>>>(d, G, V, w, fin);     \
        break;
        SMG_MD(1) SMG_MD(2) SMG_MD(3) SMG_MD(4) SMG_MD(5) SMG_MD(6) SMG_MD(7) SMG_MD(8)
#undef SMG_MD
    default:
        break;
    }
    SMG_LAUNCH_CHECK();
}
void launch_seg_gcr_orth(const SegDesc* d, int G, int nwg, c64* z, c64* Az, const SegVecs& P, const SegVecs& AP,
                         int nv, const double* dots, const SegRows& apap, hipStream_t st) {
    SMG_REQUIRE(nv >= 0 && nv <= kSegWindow, "projection on 0..kSegWindow directions");
    k_seg_gcr_orth<<<nwg, kBlock, 0, st>>>(d, G, z, Az, P, AP, nv, dots, apap);
    SMG_LAUNCH_CHECK();
}
void launch_seg_dots(const SegDesc* d, int G, int nwg, const c64* a, const c64* b, const LatFin& fin,
                     hipStream_t st) {
    k_seg_dots<<<nwg, kBlock, 0, st>>>(d, G, a, b, fin);
    SMG_LAUNCH_CHECK();
}
void launch_seg_update(const SegDesc* d, int G, int nwg, c64* x, c64* r, const c64* z, const c64* Az,
                       const double* alpha, const int* active, const LatFin& fin, hipStream_t st) {
    k_seg_update<<<nwg, kBlock, 0, st>>>(d, G, x, r, z, Az, alpha, active, fin);
    SMG_LAUNCH_CHECK();
}

}  // namespace smg
