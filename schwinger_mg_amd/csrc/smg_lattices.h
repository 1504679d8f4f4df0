// smg_lattices.h -- launchers of the many-lattice kernels (lattices.hip).
#pragma once
#include "smg_kernels.h"

namespace smg {

// one lattice of a batch: FP64 links U[mu*V + s], the scaled clover diagonal C (or
// null), 2r + m, r, mu_t; its first site in the concatenated vectors and its first
// workgroup in a launch (one work-item per site, kBlock per workgroup)
struct LatDesc {
    const c64* U;
    const double* C;
    double diag, r, mu;
    int L;
    int wg0;
    long long site0;
};

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

}  // namespace smg
