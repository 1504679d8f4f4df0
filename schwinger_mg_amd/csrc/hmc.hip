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

// ----------------------------------------------------------------- U(1) heatbath
// Checkerboard heatbath of the Wilson action S = -beta sum_x Re P(x): the links of one
// direction mu and one parity p are independent given the rest (their staples hold
// links of the other parity or direction); each is drawn from exp(beta Re(U A)), A its
// staple sum, as U = e^{i phi} conj(A)/|A| with phi from the von Mises density
// exp(kappa cos phi), kappa = beta |A| (Best & Fisher, Appl. Statist. 28 (1979) 152).
// Random numbers: Philox4x32-10 (Salmon et al., SC11), key = the 64-bit seed, counter
// (link index mu*V + x, 4*sweep + 2*mu + p, attempt, 0): the draws are a function of
// the link and the sweep only, so any launch shape gives the same field.
__device__ __forceinline__ void philox4x32(uint4& c, uint2 k) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        const unsigned lo0 = 0xD2511F53u * c.x, hi0 = __umulhi(0xD2511F53u, c.x);
        const unsigned lo1 = 0xCD9E8D57u * c.z, hi1 = __umulhi(0xCD9E8D57u, c.z);
        c = uint4{hi1 ^ c.y ^ k.x, lo1, hi0 ^ c.w ^ k.y, lo0};
        k.x += 0x9E3779B9u;
        k.y += 0xBB67AE85u;
    }
}
__device__ __forceinline__ double unit_open(unsigned x) { return ((double)x + 0.5) * (1.0 / 4294967296.0); }
// a von Mises(kappa) angle, or NaN after 1000 rejections (the link is then kept)
__device__ double von_mises(double kappa, unsigned c0, unsigned c1, uint2 key) {
    const double pi = 3.14159265358979323846;
    if (kappa < 1e-8) {
        uint4 c{c0, c1, 0u, 0u};
        philox4x32(c, key);
        return pi * (2.0 * unit_open(c.x) - 1.0);
    }
    const double tau = 1.0 + sqrt(1.0 + 4.0 * kappa * kappa);
    const double rho = (tau - sqrt(2.0 * tau)) / (2.0 * kappa);
    const double rr = (1.0 + rho * rho) / (2.0 * rho);
    for (unsigned a = 0; a < 1000u; a++) {
        uint4 c{c0, c1, a, 0u};
        philox4x32(c, key);
        const double u1 = unit_open(c.x), u2 = unit_open(c.y), u3 = unit_open(c.z);
        const double z = cos(pi * u1);
        const double f = (1.0 + rr * z) / (rr + z);
        const double cc = kappa * (rr - f);
        if (cc * (2.0 - cc) - u2 > 0.0 || log(cc / u2) + 1.0 - cc >= 0.0) {
            const double ph = acos(fmax(-1.0, fmin(1.0, f)));
            return u3 > 0.5 ? ph : -ph;
        }
    }
    return __longlong_as_double(0x7ff8000000000000ll);
}
__global__ __launch_bounds__(kBlock) void k_heatbath(c64* __restrict__ U, int L, double beta, int mu, int par,
                                                     unsigned sub, uint2 key) {
    const int V = L * L, half = V / 2;
    const int h = blockIdx.x * kBlock + threadIdx.x;
    if (h >= half) return;
    // the h-th site of parity par in lexicographic order (L even: L/2 per row)
    const int y = h / (L / 2);
    const int x = 2 * (h % (L / 2)) + ((y + par) & 1);
    const int s = x + L * y;
    const int xp = (x + 1) % L + L * y, xm = (x - 1 + L) % L + L * y;
    const int yp = x + L * ((y + 1) % L), ym = x + L * ((y - 1 + L) % L);
    const c64* U0 = U;
    const c64* U1 = U + V;
    c64 A;
    if (mu == 0) {
        const int xp_ym = (x + 1) % L + L * ((y - 1 + L) % L);
        A = U1[xp] * conj(U0[yp]) * conj(U1[s]) + conj(U0[ym]) * conj(U1[xp_ym]) * U1[ym];
    } else {
        const int xm_yp = (x - 1 + L) % L + L * ((y + 1) % L);
        A = conj(U0[s]) * conj(U1[xp]) * U0[yp] + U0[xm] * conj(U0[xm_yp]) * conj(U1[xm]);
    }
    const double a = sqrt(A.re * A.re + A.im * A.im);
    const double ph = von_mises(beta * a, (unsigned)(mu * V + s), sub, key);
    if (isnan(ph)) return;
    const c64 dir = a > 0.0 ? c64{A.re / a, -A.im / a} : c64{1.0, 0.0};
    double sn, cs;
    sincos(ph, &sn, &cs);
    U[(size_t)mu * V + s] = c64{cs, sn} * dir;
}
void launch_heatbath(c64* U, int L, double beta, int n_sweeps, unsigned long long seed, unsigned first_sweep,
                     hipStream_t st) {
    SMG_REQUIRE(L >= 2 && L % 2 == 0, "heatbath: L must be even");
    const uint2 key{(unsigned)seed, (unsigned)(seed >> 32)};
    const int half = L * L / 2;
    for (int sw = 0; sw < n_sweeps; sw++)
        for (int mu = 0; mu < 2; mu++)
            for (int par = 0; par < 2; par++) {
                const unsigned sub = 4u * (first_sweep + (unsigned)sw) + 2u * (unsigned)mu + (unsigned)par;
                k_heatbath<<<(half + kBlock - 1) / kBlock, kBlock, 0, st>>>(U, L, beta, mu, par, sub, key);
                SMG_LAUNCH_CHECK();
            }
}

}  // namespace smg
