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

}  // namespace smg
