// runtime_lat.h -- multigrid-preconditioned GCR over many lattices of different sizes,
// every kernel one launch for all lattices (the kernels: lattices.hip, lattices_mg.hip).
#pragma once
#include <vector>

#include "runtime.h"
#include "smg_lattices.h"

namespace smg {

struct LatticeMG {
    int G = 0, nlev = 0;
    MGConfig cfg;
    const LatDesc* fine = nullptr;   // the lattices' table (device), one work-item per site
    int nwg_sites = 0;
    // per level: the table of one work-item per vector entry, and (levels >= 1) of one
    // per run of a row's link entries (the coarse operator), with their launch sizes; the
    // level's concatenated length
    std::vector<std::vector<SegDesc>> hent;
    std::vector<DevMem> ent, lanes;
    std::vector<int> nwg_ent, nwg_lanes;
    std::vector<size_t> len;
    std::vector<int> kdim;                     // dof per site of every level
    std::vector<Field<double>> bv;             // per transfer: the block vectors, lattice after lattice
    std::vector<Field<double>> links;          // per coarse level l >= 1 (index l - 1), lattice after lattice
    std::vector<DevMem> links16;               // the same in FP16 (cfg.coarse_half: the cycle reads these)
    // cycle vectors per level (ts, e2s: the second coarse solve of a W-cycle), the
    // coarsest CG's, the fine operator's intermediate
    std::vector<Field<double>> xs, bs, rs, Ars, ts, e2s;
    Field<double> cp, cAp, t0;
    DevMem sc;   // per-lattice scalars (the layout is in runtime_lat.cpp)
    // the cycle captured as a hipGraph (cfg.use_graph) on fixed input/output buffers
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    Field<double> gin, gout;

    // dirac: the lattices' operators in table order; starts: every lattice's near-null
    // start vectors (host FP64), lattice after lattice, each in the layout of MG's
    // (every level's start vectors, level 0 first).  Each lattice's hierarchy is built
    // by the single-lattice setup and copied into the concatenated storage.
    LatticeMG(const std::vector<const Dirac*>& dirac, const LatDesc* fine_table, int nwg_sites, const MGConfig& c,
              const c64* starts);
    ~LatticeMG();
    LatticeMG(const LatticeMG&) = delete;
    LatticeMG& operator=(const LatticeMG&) = delete;
    // one V-cycle on every lattice: x = M b (level-0 concatenated vectors, device)
    void precondition(const c64* b, c64* x, hipStream_t st);
    // D^dag D on every lattice: dst = A src (EPI_RESID: dst = bvec - A src), with
    // {<dst, aux>, |dst|^2} per lattice into fin for EPI_DOTS
    void apply(int l, c64* dst, const c64* src, Epi epi, const c64* bvec, const c64* aux, const LatFin& fin,
               hipStream_t st);
    // GCR(trunc) preconditioned by the cycle, every lattice from x = 0 until
    // |r_g| / |b_g| < tol or maxit iterations; iters[g] and the FP64 true residuals
    // res[g] per lattice; returns the iterations run (the largest iters[g])
    int gcr(const c64* b, c64* x, int maxit, double tol, int trunc, int* iters, double* res, hipStream_t st);
    void reset_graph();

   private:
    void run(int l, const c64* b, c64* x, hipStream_t st);
    void smooth(int l, c64* x, const c64* b, int n, bool zero_init, hipStream_t st);
    void coarsest(const c64* b, c64* x, hipStream_t st);
    LatFin fin(int op, double* out, int ostride, double* aux, int astride) const;
};

}  // namespace smg
