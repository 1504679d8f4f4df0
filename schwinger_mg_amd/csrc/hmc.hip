// hmc.hip -- elementwise kernels of the HMC molecular dynamics on the device:
// the link update U <- U exp(i s pi), the momentum update pi <- pi + s (F_g + F_f),
// and the clover term C = 0.5 c_sw F01 (FP64 and FP32) recomputed after a gauge
// change.  HBM-bound, one work-item per link or site.
#include "smg_kernels.h"

namespace smg {

// gauge.U[mu][s] *= exp(i scale pi[mu][s]) (src/hmc_trajectory.cpp:352-357), in FP64,
// with the FP32 copy of the links rewritten in the same pass
__global__ __launch_bounds__(kBlock) void k_link_update(c64* __restrict__ U, c32* __restrict__ U32,
                                                        const double* __restrict__ pi, double scale, size_t n) {
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        double sn, cs;
        sincos(scale * pi[i], &sn, &cs);
        const c64 u = U[i];
        const c64 w = {u.re * cs - u.im * sn, u.re * sn + u.im * cs};
        U[i] = w;
        U32[i] = c32{(float)w.re, (float)w.im};
    }
}
void launch_link_update(c64* U, c32* U32, const double* pi, double scale, size_t n, hipStream_t st) {
    k_link_update<<<grid_for(n), kBlock, 0, st>>>(U, U32, pi, scale, n);
    SMG_LAUNCH_CHECK();
}

// mom.pi[mu][s] += scale * (gf[mu][s] + ff[mu][s]) (src/hmc_trajectory.cpp:345-350)
__global__ __launch_bounds__(kBlock) void k_mom_update(double* __restrict__ pi, const double* __restrict__ gf,
                                                       const double* __restrict__ ff, double scale, size_t n) {
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock)
        pi[i] += scale * (gf[i] + ff[i]);
}
void launch_mom_update(double* pi, const double* gf, const double* ff, double scale, size_t n, hipStream_t st) {
    k_mom_update<<<grid_for(n), kBlock, 0, st>>>(pi, gf, ff, scale, n);
    SMG_LAUNCH_CHECK();
}

// C = 0.5 c_sw F01 (src/dirac.h:102), kept in FP64 and FP32
__global__ __launch_bounds__(kBlock) void k_clover_scale(const double* __restrict__ F01, double h,
                                                         double* __restrict__ C64, float* __restrict__ C32, int V) {
    for (int s = blockIdx.x * kBlock + threadIdx.x; s < V; s += gridDim.x * kBlock) {
        const double c = h * F01[s];
        C64[s] = c;
        C32[s] = (float)c;
    }
}
void launch_clover_scale(const double* F01, double csw, double* C64, float* C32, int V, hipStream_t st) {
    k_clover_scale<<<grid_for((size_t)V), kBlock, 0, st>>>(F01, 0.5 * csw, C64, C32, V);
    SMG_LAUNCH_CHECK();
}

}  // namespace smg
