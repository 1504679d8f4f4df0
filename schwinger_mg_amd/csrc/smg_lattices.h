// smg_lattices.h -- launchers of the many-lattice kernels (lattices.hip).
#pragma once
#include <algorithm>
#include <vector>

#include "smg_kernels.h"

namespace smg {

// one lattice of a batch: FP64 links U[mu*V + s], the scaled clover diagonal C (or
// null), 2r + m, r, mu_t; its first site in the concatenated vectors, and its first
// workgroup and workgroup count in a launch (whose work-items stride over its sites)
struct LatDesc {
    const c64* U;
    const double* C;
    double diag, r, mu;
    int L;
    int wg0, nwg;
    long long site0;
};

// Workgroups of one launch over many lattices: lattice g would take ceil(n_g / kBlock);
// when they total more than the budget, each takes its share of the budget (at least
// one), so the whole launch runs about `budget` workgroups and every reducing launch
// pays the partial-sum hand-off of that many workgroups, not of one per kBlock items.
constexpr int kLatBudget = 1024;
inline std::vector<int> lattice_workgroups(const std::vector<long long>& n, int budget = kLatBudget) {
    std::vector<long long> nat(n.size());
    long long tot = 0;
    for (size_t g = 0; g < n.size(); g++) {
        nat[g] = std::max(1LL, (n[g] + kBlock - 1) / kBlock);
        tot += nat[g];
    }
    std::vector<int> w(n.size());
    for (size_t g = 0; g < n.size(); g++)
        w[g] = (int)(tot <= budget ? nat[g] : std::max(1LL, std::min(nat[g], (nat[g] * budget + tot / 2) / tot)));
    return w;
}

// a per-lattice reduction request: partial rows (one per workgroup) in part, the
// completion counter, and the scalar program run per lattice t on out + t*ostride,
// aux + t*astride; done / skip as in FinB
struct LatFin {
    int op = FIN_NONE;
    double* out = nullptr;
    int ostride = 2;
    double* aux = nullptr;
    int astride = 16;
    double* part = nullptr;
    unsigned* counter = nullptr;
    int* done = nullptr;
    const int* skip = nullptr;
};

// One lattice's share of one level of the hierarchies over many lattices: nx x nx
// sites of nd dof at complex offset off in the level's concatenated vector, the
// lattice's coarse links at link_off (levels >= 1; [site][9][nd][nd]) and its block
// vectors for the transfer to the next level at bv_off.  A launch over a table of
// these runs nwg workgroups, wg0 .. wg0 + nwg - 1, for lattice g, which stride over
// the lattice's n work-items (its vector entries, or, for the coarse operator, its
// link entries); nwg is capped at kSegCap so every lattice's reduction finishes over
// few partial rows.
struct SegDesc {
    long long off;
    long long link_off;
    long long bv_off;
    long long n;
    int nx, nd, wg0, nwg;
};
// consecutive link entries of a row per lane in the many-lattice coarse operator (its
// lanes table holds nx*nx*nd*nd / seg_coarse_run(nd) lanes per lattice)
__host__ __device__ constexpr int seg_coarse_run(int K) { return K >= 4 ? 4 : K; }
constexpr int kSegWindow = 8;     // GCR directions kept by the many-lattice solve
constexpr int kSegMaxLat = 128;   // lattices of one many-lattice hierarchy
constexpr int kSegCap = 128;      // workgroups per lattice of one launch
struct SegVecs {
    const c64* p[kSegWindow];
};
struct SegRows {                  // per window vector: its |AP|^2 per lattice
    const double* p[kSegWindow];
};

// The segment of this workgroup in a table of segments (lattices, or one level of
// their hierarchies) whose element t owns workgroups wg0_t .. wg0_{t+1} - 1: a binary
// search, broadcast via LDS.  Desc has an int member wg0.
#ifdef __HIPCC__
template <typename Desc>
__device__ __forceinline__ int my_lattice(const Desc* __restrict__ d, int G) {
    __shared__ int g_sh;
    if (threadIdx.x == 0) {
        int lo = 0, hi = G - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (d[mid].wg0 <= (int)blockIdx.x) lo = mid;
            else hi = mid - 1;
        }
        g_sh = lo;
    }
    __syncthreads();
    return g_sh;
}

// End of a reducing kernel: every thread holds its partial v[NC] for this workgroup's
// lattice g; all threads of every workgroup call this.  The workgroups of lattice g
// (wg0_g .. wg0_{g+1} - 1) arrive at that lattice's counter (f.counter + 1 + g); the
// last of them sums the lattice's rows with one wave (lane i adds rows i, i + 64, ...
// in order, then a fixed shuffle tree), so the sum is the same on every run with the
// same table, and runs fin_program with out + g*ostride and aux + g*astride.  The
// lattices finish in parallel.  With done set, the lattices' finishing workgroups then
// arrive at f.counter, and the last sets *done = 1 once every lattice's flag is set.
template <int NC, typename Desc>
__device__ __forceinline__ void finish_seg(double (&v)[NC], const LatFin& f, const Desc* __restrict__ d, int G,
                                           int g) {
    block_sum<NC>(v);
    __shared__ int last;
    const int b0 = d[g].wg0, b1 = (g + 1 < G) ? d[g + 1].wg0 : (int)gridDim.x;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int c = 0; c < NC; c++) part_store(f.part + (size_t)blockIdx.x * NC + c, v[c]);
        stores_complete();
        last = (counter_arrive(f.counter + 1 + g) == (unsigned)(b1 - b0) - 1u);
    }
    __syncthreads();
    if (!last) return;
    if (threadIdx.x < 64) {   // the first wave, uniformly
        const int lane = threadIdx.x;
        double w[NC];
#pragma unroll
        for (int c = 0; c < NC; c++) w[c] = 0.0;
        for (int b = b0 + lane; b < b1; b += 64)
#pragma unroll
            for (int c = 0; c < NC; c++) w[c] += part_load(f.part + (size_t)b * NC + c);
#pragma unroll
        for (int c = 0; c < NC; c++) w[c] = wave_sum(w[c]);
        if (lane == 0) {
            fin_program(f.op, w, NC, f.out ? f.out + (size_t)g * f.ostride : nullptr,
                        f.aux ? f.aux + (size_t)g * f.astride : nullptr);
            counter_reset(f.counter + 1 + g);
            if (f.done) {
                stores_complete();   // this lattice's flag reaches memory before the arrival
                if (counter_arrive(f.counter) == (unsigned)G - 1u) {
                    int all = 1;
                    for (int t = 0; t < G; t++)
                        all &= __hip_atomic_load(reinterpret_cast<const int*>(f.aux + (size_t)t * f.astride + 8),
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
                    *f.done = all;
                    counter_reset(f.counter);
                }
            }
        }
    }
}
#endif

// nwg = total workgroups (sum over lattices of ceil(V_g / kBlock)); d in device memory
void launch_ml_dslash(const LatDesc* d, int G, int nwg, c64* dst, const c64* src, double musign, int g5in, int g5out,
                      Epi epi, const c64* aux, const LatFin& fin, hipStream_t st);
void launch_ml_cg_state(double* aux, int G, double tol2, int* done, int* late, hipStream_t st);
void launch_ml_cg_init(const LatDesc* d, int G, int nwg, const c64* b, c64* x, c64* r, c64* p, const LatFin& fin,
                       hipStream_t st);
void launch_ml_cg_r(const LatDesc* d, int G, int nwg, c64* r, const c64* Ap, const double* alpha, const LatFin& fin,
                    int* late, hipStream_t st);
void launch_ml_cg_xp(const LatDesc* d, int G, int nwg, c64* x, c64* p, const c64* r, const double* alpha,
                     const double* beta, const int* late, hipStream_t st);
void launch_ml_norm2(const LatDesc* d, int G, int nwg, const c64* v, const LatFin& fin, hipStream_t st);

// multigrid over many lattices (lattices_mg.hip).  fd / cd: a finer level's and the
// next coarser level's tables (one work-item per vector entry); transfers over bx x by
// aggregates; the coarse operator runs on a table of one work-item per run of
// seg_coarse_run(K) link entries of a row (K = nd of that level, a power of two <= 64;
// links FP64 or FP16, vectors and arithmetic FP64).  nwg: the workgroups of the table the launch runs on.
void launch_seg_restrict(const SegDesc* fd, const SegDesc* cd, int G, int nwg_c, const c64* bv, const c64* fine,
                         c64* coarse, int bx, int by, int chiral, hipStream_t st);
void launch_seg_prolong(const SegDesc* fd, const SegDesc* cd, int G, int nwg_f, const c64* bv, const c64* coarse,
                        c64* fine, int accumulate, int bx, int by, int chiral, hipStream_t st);
template <typename TL>
void launch_seg_coarse(const SegDesc* lanes, int G, int nwg_lanes, int K, const TL* links, const c64* src, c64* dst,
                       Epi epi, const c64* bvec, const c64* aux, const LatFin& fin, hipStream_t st);
void launch_seg_mr_update(const SegDesc* d, int G, int nwg, c64* x, c64* r, const c64* Ar, const double* w,
                          hipStream_t st);
void launch_seg_cg_init(const SegDesc* d, int G, int nwg, const c64* b, c64* x, c64* r, c64* p, const LatFin& fin,
                        hipStream_t st);
void launch_seg_cg_r(const SegDesc* d, int G, int nwg, c64* r, const c64* Ap, const double* alpha, const LatFin& fin,
                     int* late, hipStream_t st);
void launch_seg_cg_xp(const SegDesc* d, int G, int nwg, c64* x, c64* p, const c64* r, const double* alpha,
                      const double* beta, const double* cg_aux, const int* late, hipStream_t st);
void launch_seg_multi_dot(const SegDesc* d, int G, int nwg, const SegVecs& V, int nv, const c64* w, const LatFin& fin,
                          hipStream_t st);
void launch_seg_gcr_orth(const SegDesc* d, int G, int nwg, c64* z, c64* Az, const SegVecs& P, const SegVecs& AP,
                         int nv, const double* dots, const SegRows& apap, hipStream_t st);
void launch_seg_dots(const SegDesc* d, int G, int nwg, const c64* a, const c64* b, const LatFin& fin,
                     hipStream_t st);
void launch_seg_update(const SegDesc* d, int G, int nwg, c64* x, c64* r, const c64* z, const c64* Az,
                       const double* alpha, const int* active, const LatFin& fin, hipStream_t st);

}  // namespace smg
