// dirac.hip -- Wilson-clover(-twisted) stencil kernels for gfx950.
//
// Layouts are the reference's (src/lattice.h:19, src/dirac.h:99, src/gauge.h:12):
// spinor psi[2*s + spin], s = x + L*y; links U[mu*V + s].  One work-item per
// (half-)lattice site; a wave64 covers 64 consecutive x of one row, so the +-x
// neighbours are in the same cache lines and the +-y neighbours are the rows the
// neighbouring waves stream (L2 reuse).  Every kernel is HBM-bound.
#include "smg_kernels.h"

namespace smg {

template <typename T>
__device__ __forceinline__ void ld_sp(const cplx<T>* __restrict__ p, int s, int g5, cplx<T>& a, cplx<T>& b) {
    a = p[2 * s];
    b = p[2 * s + 1];
    if (g5) b = -b;
}

// The Wilson hopping sum of src/dirac.h:58-89 with all phase scales 1:
//   h = -1/2 sum_mu [ (r - sigma_mu) U_mu(s) psi(s+mu) + (r + sigma_mu) U_mu(s-mu)^* psi(s-mu) ]
// with sigma_1 = sigma_x, sigma_2 = sigma_y.
template <typename T>
__device__ __forceinline__ void wilson_hop(T r, cplx<T> u0, cplx<T> u0b, cplx<T> u1, cplx<T> u1b, cplx<T> xf0,
                                           cplx<T> xf1, cplx<T> xb0, cplx<T> xb1, cplx<T> yf0, cplx<T> yf1,
                                           cplx<T> yb0, cplx<T> yb1, cplx<T>& h0, cplx<T>& h1) {
    const cplx<T> up0 = u0 * xf0, up1 = u0 * xf1;
    const cplx<T> ud = conj(u0b);
    const cplx<T> q0 = ud * xb0, q1 = ud * xb1;
    const cplx<T> vp0 = u1 * yf0, vp1 = u1 * yf1;
    const cplx<T> wd = conj(u1b);
    const cplx<T> w0 = wd * yb0, w1 = wd * yb1;
    const T hr = (T)0.5 * r;
    // spin 0: r(up0+q0+vp0+w0) - up1 + q1 + i(vp1 - w1)
    // spin 1: r(up1+q1+vp1+w1) - up0 + q0 - i(vp0 - w0)
    cplx<T> s0 = (up0 + q0) + (vp0 + w0);
    cplx<T> s1 = (up1 + q1) + (vp1 + w1);
    cplx<T> t0 = (q1 - up1) + mul_i(vp1 - w1);
    cplx<T> t1 = (q0 - up0) - mul_i(vp0 - w0);
    h0 = -(hr * s0) - (T)0.5 * t0;
    h1 = -(hr * s1) - (T)0.5 * t1;
}

// ----------------------------------------------------------------- full lattice
// (D psi)(s) at site s = x + L*y (src/dirac.h:93-117): the clover(-twisted) diagonal and
// the hopping sum; g5in: psi's lower spin negated on load
template <typename T>
__device__ __forceinline__ void dslash_site(const DiracView<T>& D, const cplx<T>* __restrict__ src, T mu, int g5in,
                                            int s, int x, int y, cplx<T>& o0, cplx<T>& o1) {
    const int L = D.L, V = L * L;
    const int sxp = (x + 1 == L) ? s - x : s + 1;
    const int sxm = (x == 0) ? s + L - 1 : s - 1;
    const int syp = (y + 1 == L) ? x : s + L;
    const int sym = (y == 0) ? s + V - L : s - L;
    cplx<T> s0, s1, xf0, xf1, xb0, xb1, yf0, yf1, yb0, yb1;
    ld_sp(src, s, g5in, s0, s1);
    ld_sp(src, sxp, g5in, xf0, xf1);
    ld_sp(src, sxm, g5in, xb0, xb1);
    ld_sp(src, syp, g5in, yf0, yf1);
    ld_sp(src, sym, g5in, yb0, yb1);
    const cplx<T>* U0 = D.U;
    const cplx<T>* U1 = D.U + V;
    cplx<T> h0, h1;
    wilson_hop(D.r, U0[s], U0[sxm], U1[s], U1[sym], xf0, xf1, xb0, xb1, yf0, yf1, yb0, yb1, h0, h1);
    const T C = D.C ? D.C[s] : (T)0;
    const cplx<T> e0 = mk<T>(D.diag + C, mu), e1 = mk<T>(D.diag - C, -mu);
    o0 = e0 * s0 + h0;
    o1 = e1 * s1 + h1;
}

template <typename T, int EPI>
__global__ __launch_bounds__(kBlock) void k_dslash(DiracView<T> D, cplx<T>* __restrict__ dst,
                                                   const cplx<T>* __restrict__ src, T mu, int g5in, int g5out,
                                                   const cplx<T>* __restrict__ bvec,
                                                   const cplx<T>* __restrict__ aux, Fin fin, size_t vstride) {
    const int L = D.L, V = L * L;
    const int s = blockIdx.x * kBlock + threadIdx.x;
    const size_t off = (size_t)blockIdx.y * vstride;
    src += off;
    dst += off;
    double acc[3] = {0.0, 0.0, 0.0};
    if (s < V) {
        cplx<T> o0, o1;
        dslash_site(D, src, mu, g5in, s, s % L, s / L, o0, o1);
        if (g5out) o1 = -o1;
        if (EPI == EPI_RESID) {
            const cplx<T>* bb = bvec + off;
            o0 = bb[2 * s] - o0;
            o1 = bb[2 * s + 1] - o1;
        }
        dst[2 * s] = o0;
        dst[2 * s + 1] = o1;
        if (EPI == EPI_DOTS) {
            const cplx<T>* a = aux + off;
            const cplx<T> a0 = a[2 * s], a1 = a[2 * s + 1];
            const cplx<T> d0 = cdot(o0, a0), d1 = cdot(o1, a1);
            acc[0] = (double)d0.re + (double)d1.re;
            acc[1] = (double)d0.im + (double)d1.im;
            acc[2] = (double)norm2(o0) + (double)norm2(o1);
        }
    }
    if (EPI == EPI_DOTS) reduce_finish<3>(acc, fin, gridDim.x * gridDim.y, blockIdx.y * gridDim.x + blockIdx.x);
}

template <typename T>
void launch_dslash(const DiracView<T>& D, cplx<T>* dst, const cplx<T>* src, T mu, int g5in, int g5out, Epi epi,
                   const cplx<T>* b, const cplx<T>* aux, const Fin& fin, int nrhs, size_t vstride, hipStream_t st) {
    const int V = D.L * D.L;
    dim3 grid(grid_for(V), nrhs);
    switch (epi) {
    case EPI_NONE:
        k_dslash<T, EPI_NONE><<<grid, kBlock, 0, st>>>(D, dst, src, mu, g5in, g5out, b, aux, fin, vstride);
        break;
    case EPI_RESID:
        k_dslash<T, EPI_RESID><<<grid, kBlock, 0, st>>>(D, dst, src, mu, g5in, g5out, b, aux, fin, vstride);
        break;
    default:
        SMG_REQUIRE(nrhs == 1, "the reducing epilogue takes one right-hand side");
        k_dslash<T, EPI_DOTS><<<grid, kBlock, 0, st>>>(D, dst, src, mu, g5in, g5out, b, aux, fin, vstride);
        break;
    }
    SMG_LAUNCH_CHECK();
}

// ----------------------------------------------------------------- even-odd
// half-lattice index i of parity p -> site (x, y); eo index of (x, y) = y*L/2 + x/2
// (the closed form of src/lattice.h:28-41 for even L).
template <typename T, int MODE, int EPI>
__global__ __launch_bounds__(kBlock) void k_eo_hop(DiracView<T> D, cplx<T>* __restrict__ dst,
                                                   const cplx<T>* __restrict__ src, int p, T mu, int g5src,
                                                   int g5aux, int g5out, const cplx<T>* __restrict__ aux,
                                                   const cplx<T>* __restrict__ bvec,
                                                   const cplx<T>* __restrict__ aux2, Fin fin) {
    const int L = D.L, V = L * L, Lh = L >> 1, Vh = V >> 1;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double acc[3] = {0.0, 0.0, 0.0};
    if (i < Vh) {
        const int y = i / Lh;
        const int x = 2 * (i - y * Lh) + ((y + p) & 1);
        const int s = x + L * y;
        const int xp = (x + 1 == L) ? 0 : x + 1, xm = (x == 0) ? L - 1 : x - 1;
        const int yp = (y + 1 == L) ? 0 : y + 1, ym = (y == 0) ? L - 1 : y - 1;
        cplx<T> xf0, xf1, xb0, xb1, yf0, yf1, yb0, yb1;
        ld_sp(src, y * Lh + (xp >> 1), g5src, xf0, xf1);
        ld_sp(src, y * Lh + (xm >> 1), g5src, xb0, xb1);
        ld_sp(src, yp * Lh + (x >> 1), g5src, yf0, yf1);
        ld_sp(src, ym * Lh + (x >> 1), g5src, yb0, yb1);
        const cplx<T>* U0 = D.U;
        const cplx<T>* U1 = D.U + V;
        cplx<T> h0, h1;
        wilson_hop(D.r, U0[s], U0[xm + L * y], U1[s], U1[x + L * ym], xf0, xf1, xb0, xb1, yf0, yf1, yb0, yb1,
                   h0, h1);
        cplx<T> o0, o1;
        if (MODE == 0) {
            o0 = h0;
            o1 = h1;
        } else {
            const T C = D.C ? D.C[s] : (T)0;
            const cplx<T> e0 = mk<T>(D.diag + C, mu), e1 = mk<T>(D.diag - C, -mu);
            if (MODE == 1) {
                o0 = cdiv(h0, e0);
                o1 = cdiv(h1, e1);
            } else {
                cplx<T> a0, a1;
                ld_sp(aux, i, g5aux, a0, a1);
                o0 = e0 * a0 - h0;
                o1 = e1 * a1 - h1;
            }
        }
        if (g5out) o1 = -o1;
        if (EPI == EPI_RESID) {
            o0 = bvec[2 * i] - o0;
            o1 = bvec[2 * i + 1] - o1;
        }
        dst[2 * i] = o0;
        dst[2 * i + 1] = o1;
        if (EPI == EPI_DOTS) {
            const cplx<T> a0 = aux2[2 * i], a1 = aux2[2 * i + 1];
            const cplx<T> d0 = cdot(o0, a0), d1 = cdot(o1, a1);
            acc[0] = (double)d0.re + (double)d1.re;
            acc[1] = (double)d0.im + (double)d1.im;
            acc[2] = (double)norm2(o0) + (double)norm2(o1);
        }
    }
    if (EPI == EPI_DOTS) reduce_finish<3>(acc, fin, gridDim.x, blockIdx.x);
}

template <typename T>
void launch_eo_hop(const DiracView<T>& D, cplx<T>* dst, const cplx<T>* src, int p, int mode, T mu, int g5in,
                   int g5out, const cplx<T>* aux, Epi epi, const cplx<T>* b, const cplx<T>* aux2, const Fin& fin,
                   hipStream_t st) {
    SMG_REQUIRE(D.L % 2 == 0, "even-odd operators need an even L");
    const int Vh = D.L * D.L / 2;
    const int g = grid_for(Vh);
    // In mode 2 the g5 flag applies to aux (the diagonal term); src is the even tmp.
    const int g5src = (mode == 2) ? 0 : g5in, g5aux = (mode == 2) ? g5in : 0;
#define SMG_EO(M, E)                                                                                  \
    k_eo_hop<T, M, E><<<g, kBlock, 0, st>>>(D, dst, src, p, mu, g5src, g5aux, g5out, aux, b, aux2, fin)
    if (mode == 0) SMG_EO(0, EPI_NONE);
    else if (mode == 1) SMG_EO(1, EPI_NONE);
    else if (epi == EPI_NONE) SMG_EO(2, EPI_NONE);
    else if (epi == EPI_RESID) SMG_EO(2, EPI_RESID);
    else SMG_EO(2, EPI_DOTS);
#undef SMG_EO
    SMG_LAUNCH_CHECK();
}

template <typename T>
__global__ void k_eo_diag(DiracView<T> D, cplx<T>* __restrict__ dst, const cplx<T>* __restrict__ src, int p,
                          int inv, T mu) {
    const int L = D.L, Lh = L >> 1, Vh = L * L / 2;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= Vh) return;
    const int y = i / Lh, x = 2 * (i - y * Lh) + ((y + p) & 1), s = x + L * y;
    const T C = D.C ? D.C[s] : (T)0;
    const cplx<T> e0 = mk<T>(D.diag + C, mu), e1 = mk<T>(D.diag - C, -mu);
    if (inv) {
        dst[2 * i] = cdiv(src[2 * i], e0);
        dst[2 * i + 1] = cdiv(src[2 * i + 1], e1);
    } else {
        dst[2 * i] = e0 * src[2 * i];
        dst[2 * i + 1] = e1 * src[2 * i + 1];
    }
}
template <typename T>
void launch_eo_diag(const DiracView<T>& D, cplx<T>* dst, const cplx<T>* src, int p, int inv, T mu, hipStream_t st) {
    k_eo_diag<T><<<grid_for(D.L * D.L / 2), kBlock, 0, st>>>(D, dst, src, p, inv, mu);
    SMG_LAUNCH_CHECK();
}

template <typename T>
__global__ void k_eo_pack(int L, cplx<T>* __restrict__ half, cplx<T>* __restrict__ full, int p, int to_half) {
    const int Lh = L >> 1, Vh = L * L / 2;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= Vh) return;
    const int y = i / Lh, x = 2 * (i - y * Lh) + ((y + p) & 1), s = x + L * y;
    if (to_half) {
        half[2 * i] = full[2 * s];
        half[2 * i + 1] = full[2 * s + 1];
    } else {
        full[2 * s] = half[2 * i];
        full[2 * s + 1] = half[2 * i + 1];
    }
}
template <typename T>
void launch_eo_gather(int L, cplx<T>* half, const cplx<T>* full, int p, hipStream_t st) {
    k_eo_pack<T><<<grid_for(L * L / 2), kBlock, 0, st>>>(L, half, const_cast<cplx<T>*>(full), p, 1);
    SMG_LAUNCH_CHECK();
}
template <typename T>
void launch_eo_scatter(int L, cplx<T>* full, const cplx<T>* half, int p, hipStream_t st) {
    k_eo_pack<T><<<grid_for(L * L / 2), kBlock, 0, st>>>(L, const_cast<cplx<T>*>(half), full, p, 0);
    SMG_LAUNCH_CHECK();
}

// ----------------------------------------------------------------- gauge-only kernels (FP64)
struct Nb {
    int s, xp, xm, yp, ym;
};
__device__ __forceinline__ int nb_site(int L, int x, int y) { return ((x + L) % L) + L * ((y + L) % L); }

// src/dirac.h:26-52
__global__ void k_clover(int L, const c64* __restrict__ U, double* __restrict__ F01) {
    const int V = L * L;
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= V) return;
    const int x = s % L, y = s / L;
    const c64* U0 = U;
    const c64* U1 = U + V;
    const int sxp = nb_site(L, x + 1, y), sxm = nb_site(L, x - 1, y);
    const int syp = nb_site(L, x, y + 1), sym = nb_site(L, x, y - 1);
    const int sxmyp = nb_site(L, x - 1, y + 1), sxmym = nb_site(L, x - 1, y - 1), sxpym = nb_site(L, x + 1, y - 1);
    c64 P_ur = U0[s] * U1[sxp] * conj(U0[syp]) * conj(U1[s]);
    c64 P_ul = U1[s] * conj(U0[sxmyp]) * conj(U1[sxm]) * U0[sxm];
    c64 P_ll = conj(U0[sxm]) * conj(U1[sxmym]) * U0[sxmym] * U1[sym];
    c64 P_lr = conj(U1[sym]) * U0[sym] * U1[sxpym] * conj(U0[s]);
    F01[s] = (P_ur.im + P_ul.im + P_ll.im + P_lr.im) / 4.0;
}
void launch_clover(int L, const c64* U, double* F01, hipStream_t st) {
    k_clover<<<grid_for(L * L), kBlock, 0, st>>>(L, U, F01);
    SMG_LAUNCH_CHECK();
}

// src/dirac.h:159-174
__global__ void k_delta_dirac(DiracView<double> D, c64* __restrict__ dst, const c64* __restrict__ src,
                              const double* __restrict__ pi, double dt) {
    const int L = D.L, V = L * L;
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= V) return;
    const int x = s % L, y = s / L;
    const int sxp = nb_site(L, x + 1, y), sxm = nb_site(L, x - 1, y);
    const int syp = nb_site(L, x, y + 1), sym = nb_site(L, x, y - 1);
    double sn, cs;
    sincos(dt * pi[s], &sn, &cs);
    const c64 fx = {cs - 1.0, sn};
    sincos(-dt * pi[sxm], &sn, &cs);
    const c64 bx = {cs - 1.0, sn};
    sincos(dt * pi[V + s], &sn, &cs);
    const c64 fy = {cs - 1.0, sn};
    sincos(-dt * pi[V + sym], &sn, &cs);
    const c64 by = {cs - 1.0, sn};
    // the backward scale multiplies U^dag (src/dirac.h:71-76), i.e. conj(conj(scale) U)
    const c64* U0 = D.U;
    const c64* U1 = D.U + V;
    c64 h0, h1;
    wilson_hop(D.r, fx * U0[s], conj(bx) * U0[sxm], fy * U1[s], conj(by) * U1[sym], src[2 * sxp], src[2 * sxp + 1],
               src[2 * sxm], src[2 * sxm + 1], src[2 * syp], src[2 * syp + 1], src[2 * sym], src[2 * sym + 1], h0,
               h1);
    dst[2 * s] = h0;
    dst[2 * s + 1] = h1;
}
void launch_delta_dirac(const DiracView<double>& D, c64* dst, const c64* src, const double* pi, double dt,
                        hipStream_t st) {
    k_delta_dirac<<<grid_for(D.L * D.L), kBlock, 0, st>>>(D, dst, src, pi, dt);
    SMG_LAUNCH_CHECK();
}

// src/dirac.h:348-366: partial sums over even sites
__global__ __launch_bounds__(kBlock) void k_logdet(DiracView<double> D, double mu, Fin fin) {
    const int L = D.L, Lh = L >> 1, Vh = L * L / 2;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double v[1] = {0.0};
    if (i < Vh) {
        const int y = i / Lh, x = 2 * (i - y * Lh) + (y & 1), s = x + L * y;
        const double C = D.C ? D.C[s] : 0.0;
        const double d0 = D.diag + C, d1 = D.diag - C;
        v[0] = 0.5 * log(d0 * d0 + mu * mu) + 0.5 * log(d1 * d1 + mu * mu);
    }
    reduce_finish<1>(v, fin, gridDim.x, blockIdx.x);
}
void launch_logdet(const DiracView<double>& D, double mu, const Fin& fin, hipStream_t st) {
    k_logdet<<<grid_for(D.L * D.L / 2), kBlock, 0, st>>>(D, mu, fin);
    SMG_LAUNCH_CHECK();
}

__global__ __launch_bounds__(kBlock) void k_plaq_sum(int L, const c64* __restrict__ U, Fin fin) {
    const int V = L * L;
    const int s = blockIdx.x * kBlock + threadIdx.x;
    double v[1] = {0.0};
    if (s < V) {
        const int x = s % L, y = s / L;
        const c64 P = U[s] * U[V + nb_site(L, x + 1, y)] * conj(U[nb_site(L, x, y + 1)]) * conj(U[V + s]);
        v[0] = P.re;
    }
    reduce_finish<1>(v, fin, gridDim.x, blockIdx.x);
}
void launch_plaq_sum(int L, const c64* U, const Fin& fin, hipStream_t st) {
    k_plaq_sum<<<grid_for(L * L), kBlock, 0, st>>>(L, U, fin);
    SMG_LAUNCH_CHECK();
}

// ----------------------------------------------------------------- forces (FP64)
// Hopping force of src/hmc_forces.cpp:141-189 (= hopping_force_bilinear, :540-569):
// F_mu(s) = Re(i c_bwd - i c_fwd),
//   c_fwd = chi(s)^dag (r - sigma_mu) U_mu(s) X(s+mu),
//   c_bwd = chi(s+mu)^dag (r + sigma_mu) U_mu(s)^* X(s).
// chi at s (c0, c1), at s + x (cx0, cx1) and at s + y (cy0, cy1)
__device__ __forceinline__ void hop_force_vals(const DiracView<double>& D, const c64* __restrict__ X, int s, int sfx,
                                               int sfy, c64 c0, c64 c1, c64 cx0, c64 cx1, c64 cy0, c64 cy1,
                                               double& f0, double& f1) {
    const int V = D.L * D.L;
    const double r = D.r;
    const c64 x0 = X[2 * s], x1 = X[2 * s + 1];
    {
        const c64 u = D.U[s], ud = conj(u);
        const c64 ux0 = u * X[2 * sfx], ux1 = u * X[2 * sfx + 1];
        const c64 fw0 = r * ux0 - ux1, fw1 = r * ux1 - ux0;
        const c64 cf = cdot(c0, fw0) + cdot(c1, fw1);
        const c64 udx0 = ud * x0, udx1 = ud * x1;
        const c64 bw0 = r * udx0 + udx1, bw1 = udx0 + r * udx1;
        const c64 cb = cdot(cx0, bw0) + cdot(cx1, bw1);
        f0 = -(cb.im - cf.im);   // Re(i z) = -Im z
    }
    {
        const c64 u = D.U[V + s], ud = conj(u);
        const c64 ux0 = u * X[2 * sfy], ux1 = u * X[2 * sfy + 1];
        const c64 fw0 = r * ux0 + mul_i(ux1), fw1 = r * ux1 - mul_i(ux0);
        const c64 cf = cdot(c0, fw0) + cdot(c1, fw1);
        const c64 udx0 = ud * x0, udx1 = ud * x1;
        const c64 bw0 = r * udx0 - mul_i(udx1), bw1 = mul_i(udx0) + r * udx1;
        const c64 cb = cdot(cy0, bw0) + cdot(cy1, bw1);
        f1 = -(cb.im - cf.im);
    }
}
__device__ __forceinline__ void hop_force_site(const DiracView<double>& D, const c64* __restrict__ chi,
                                               const c64* __restrict__ X, int s, int sfx, int sfy, double& f0,
                                               double& f1) {
    hop_force_vals(D, X, s, sfx, sfy, chi[2 * s], chi[2 * s + 1], chi[2 * sfx], chi[2 * sfx + 1], chi[2 * sfy],
                   chi[2 * sfy + 1], f0, f1);
}

__global__ void k_hopping_force(DiracView<double> D, const c64* __restrict__ chi, const c64* __restrict__ X,
                                double scale, int accumulate, double* __restrict__ F) {
    const int L = D.L, V = L * L;
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= V) return;
    const int x = s % L, y = s / L;
    double f0, f1;
    hop_force_site(D, chi, X, s, nb_site(L, x + 1, y), nb_site(L, x, y + 1), f0, f1);
    if (accumulate) {
        F[s] += scale * f0;
        F[V + s] += scale * f1;
    } else {
        F[s] = scale * f0;
        F[V + s] = scale * f1;
    }
}
void launch_hopping_force(const DiracView<double>& D, const c64* chi, const c64* X, double scale, int accumulate,
                          double* F, hipStream_t st) {
    k_hopping_force<<<grid_for(D.L * D.L), kBlock, 0, st>>>(D, chi, X, scale, accumulate, F);
    SMG_LAUNCH_CHECK();
}

// Clover-derivative insertion, src/hmc_forces.cpp:501-534.  W(site) returns the weight.
template <typename WF>
__device__ __forceinline__ void clover_insert_site(const DiracView<double>& D, int s, int x, int y, WF W, double& f0,
                                                   double& f1) {
    const int L = D.L, V = L * L;
    const c64* U0 = D.U;
    const c64* U1 = D.U + V;
    const int sxp = nb_site(L, x + 1, y), sxm = nb_site(L, x - 1, y);
    const int syp = nb_site(L, x, y + 1), sym = nb_site(L, x, y - 1);
    const int sxpym = nb_site(L, x + 1, y - 1), sxmyp = nb_site(L, x - 1, y + 1), sxpyp = nb_site(L, x + 1, y + 1);
    const double ws = W(s), wxp = W(sxp), wyp = W(syp), wxpyp = W(sxpyp);
    const double re_fwd = (U0[s] * U1[sxp] * conj(U0[syp]) * conj(U1[s])).re;
    const double re_bwd = (U0[sym] * U1[sxpym] * conj(U0[s]) * conj(U1[sym])).re;
    const double w_fwd = ws + wxp + wxpyp + wyp;
    const double w_bwd = W(sym) + W(sxpym) + wxp + ws;
    f0 = re_fwd * w_fwd - re_bwd * w_bwd;
    const double re_fwd1 = (U0[sxm] * U1[s] * conj(U0[sxmyp]) * conj(U1[sxm])).re;
    const double w_fwd1 = W(sxm) + ws + wyp + W(sxmyp);
    f1 = re_fwd1 * w_fwd1 - re_fwd * w_fwd;
}

__global__ void k_clover_insert(DiracView<double> D, const double* __restrict__ w, double factor,
                                double* __restrict__ F) {
    const int L = D.L, V = L * L;
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= V) return;
    double f0, f1;
    clover_insert_site(D, s, s % L, s / L, [&](int t) { return w[t]; }, f0, f1);
    F[s] += factor * f0;
    F[V + s] += factor * f1;
}
void launch_clover_insert(const DiracView<double>& D, const double* w, double factor, double* F, hipStream_t st) {
    k_clover_insert<<<grid_for(D.L * D.L), kBlock, 0, st>>>(D, w, factor, F);
    SMG_LAUNCH_CHECK();
}

__device__ __forceinline__ double sig3w(const c64* __restrict__ chi, const c64* __restrict__ X, int t) {
    const c64 a = cdot(chi[2 * t], X[2 * t]), b = cdot(chi[2 * t + 1], X[2 * t + 1]);
    return a.re - b.re;
}
__global__ void k_sigma3_weight(int V, const c64* __restrict__ chi, const c64* __restrict__ X, double* w) {
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s < V) w[s] = sig3w(chi, X, s);
}
void launch_sigma3_weight(int V, const c64* chi, const c64* X, double* w, hipStream_t st) {
    k_sigma3_weight<<<grid_for(V), kBlock, 0, st>>>(V, chi, X, w);
    SMG_LAUNCH_CHECK();
}

// Fused force backward (hopping + clover), the sigma_3 weight recomputed from
// chi and X at the eight clover sites instead of being stored.
__global__ __launch_bounds__(kBlock) void k_force_backward(DiracView<double> D, double csw,
                                                           const c64* __restrict__ chi, const c64* __restrict__ X,
                                                           double* __restrict__ F) {
    const int L = D.L, V = L * L;
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= V) return;
    const int x = s % L, y = s / L;
    double f0, f1;
    hop_force_site(D, chi, X, s, nb_site(L, x + 1, y), nb_site(L, x, y + 1), f0, f1);
    if (csw != 0.0) {
        double g0, g1;
        clover_insert_site(D, s, x, y, [&](int t) { return sig3w(chi, X, t); }, g0, g1);
        f0 += (csw / 4.0) * g0;
        f1 += (csw / 4.0) * g1;
    }
    F[s] = f0;
    F[V + s] = f1;
}
void launch_force_backward(const DiracView<double>& D, double csw, const c64* chi, const c64* X, double* F,
                           hipStream_t st) {
    k_force_backward<<<grid_for(D.L * D.L), kBlock, 0, st>>>(D, csw, chi, X, F);
    SMG_LAUNCH_CHECK();
}

// The whole fermion force from X alone: chi = D X is recomputed, never stored.  A
// workgroup owns a kForceTile x kForceTile tile of sites; it first evaluates chi on
// the tile and a one-site halo into LDS (the stencil of dslash_site, periodic wrap;
// lattices narrower than the tile wrap onto themselves, which gives the same values),
// then each work-item forms its site's hopping force and clover insertion from the
// LDS chi, whose eight clover sites and two hopping neighbours all lie in the halo.
constexpr int kForceTile = 16;
static_assert(kForceTile * kForceTile == kBlock, "one work-item per tile site");
__global__ __launch_bounds__(kBlock) void k_force_fused(DiracView<double> D, double csw, double mu,
                                                        const c64* __restrict__ X, double* __restrict__ F) {
    constexpr int TH = kForceTile + 2;
    __shared__ c64 chi_s[TH * TH][2];
    const int L = D.L, V = L * L;
    const int ntx = (L + kForceTile - 1) / kForceTile;
    const int tx0 = ((int)blockIdx.x % ntx) * kForceTile, ty0 = ((int)blockIdx.x / ntx) * kForceTile;
    for (int i = threadIdx.x; i < TH * TH; i += kBlock) {
        const int gx = (tx0 + i % TH - 1 + L) % L, gy = (ty0 + i / TH - 1 + L) % L;
        c64 o0, o1;
        dslash_site(D, X, mu, 0, gx + L * gy, gx, gy, o0, o1);
        chi_s[i][0] = o0;
        chi_s[i][1] = o1;
    }
    __syncthreads();
    const int lx = threadIdx.x % kForceTile, ly = threadIdx.x / kForceTile;
    const int x = tx0 + lx, y = ty0 + ly;
    if (x >= L || y >= L) return;
    const int s = x + L * y;
    auto li = [&](int dx, int dy) { return (ly + 1 + dy) * TH + (lx + 1 + dx); };
    const int c = li(0, 0), cx = li(1, 0), cy = li(0, 1);
    double f0, f1;
    hop_force_vals(D, X, s, nb_site(L, x + 1, y), nb_site(L, x, y + 1), chi_s[c][0], chi_s[c][1], chi_s[cx][0],
                   chi_s[cx][1], chi_s[cy][0], chi_s[cy][1], f0, f1);
    if (csw != 0.0) {
        // sigma_3 weight at lattice site t, chi from the LDS entry of the same site
        auto W = [&](int t) {
            const int tx = t % L, ty = t / L;
            int dx = tx - x, dy = ty - y;   // the offset in -1..1, undoing the wrap
            if (dx > 1) dx -= L;
            if (dx < -1) dx += L;
            if (dy > 1) dy -= L;
            if (dy < -1) dy += L;
            const int k = li(dx, dy);
            const c64 a = cdot(chi_s[k][0], X[2 * t]), b = cdot(chi_s[k][1], X[2 * t + 1]);
            return a.re - b.re;
        };
        double g0, g1;
        clover_insert_site(D, s, x, y, W, g0, g1);
        f0 += (csw / 4.0) * g0;
        f1 += (csw / 4.0) * g1;
    }
    F[s] = f0;
    F[V + s] = f1;
}
void launch_force_fused(const DiracView<double>& D, double csw, double mu, const c64* X, double* F,
                        hipStream_t st) {
    const int nt = (D.L + kForceTile - 1) / kForceTile;
    k_force_fused<<<nt * nt, kBlock, 0, st>>>(D, csw, mu, X, F);
    SMG_LAUNCH_CHECK();
}

// e/o Schur force weights, src/hmc_forces.cpp:795-809
__global__ void k_eo_force_weights(int L, const c64* __restrict__ chi_o, const c64* __restrict__ psi_o,
                                   const c64* __restrict__ tmp3, const c64* __restrict__ tmp2, double* w) {
    const int Lh = L >> 1, Vh = L * L / 2;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= Vh) return;
    const int y = i / Lh;
    const int xo = 2 * (i - y * Lh) + ((y + 1) & 1), xe = 2 * (i - y * Lh) + (y & 1);
    w[xo + L * y] = sig3w(chi_o, psi_o, i);
    w[xe + L * y] = sig3w(tmp3, tmp2, i);
}
void launch_eo_force_weights(const DiracView<double>& D, const c64* chi_o, const c64* psi_o, const c64* tmp3,
                             const c64* tmp2, double* w, hipStream_t st) {
    k_eo_force_weights<<<grid_for(D.L * D.L / 2), kBlock, 0, st>>>(D.L, chi_o, psi_o, tmp3, tmp2, w);
    SMG_LAUNCH_CHECK();
}

// src/hmc_forces.cpp:593-600
__global__ void k_logdet_weight(DiracView<double> D, double csw, double mu, double* w) {
    const int L = D.L, V = L * L;
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= V) return;
    const int x = s % L, y = s / L;
    if (((x + y) & 1) != 0) {
        w[s] = 0.0;
        return;
    }
    const double C = D.C ? D.C[s] : 0.0;
    const double d0 = D.diag + C, d1 = D.diag - C, mu2 = mu * mu;
    w[s] = 0.5 * csw * (d0 / (d0 * d0 + mu2) - d1 / (d1 * d1 + mu2));
}
void launch_logdet_weight(const DiracView<double>& D, double csw, double mu, double* w, hipStream_t st) {
    k_logdet_weight<<<grid_for(D.L * D.L), kBlock, 0, st>>>(D, csw, mu, w);
    SMG_LAUNCH_CHECK();
}

// src/hmc_forces.cpp:88-111
__global__ void k_gauge_force(int L, const c64* __restrict__ U, double beta, double* __restrict__ F) {
    const int V = L * L;
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= V) return;
    const int x = s % L, y = s / L;
    const c64* U0 = U;
    const c64* U1 = U + V;
    {
        const int sx = nb_site(L, x + 1, y), sy = nb_site(L, x, y + 1), smy = nb_site(L, x, y - 1);
        const int sxmy = nb_site(L, x + 1, y - 1);
        const c64 st = U1[sx] * conj(U0[sy]) * conj(U1[s]) + conj(U1[sxmy]) * conj(U0[smy]) * U1[smy];
        F[s] = -beta * (U0[s] * st).im;
    }
    {
        const int sy = nb_site(L, x, y + 1), sx = nb_site(L, x + 1, y), smx = nb_site(L, x - 1, y);
        const int symx = nb_site(L, x - 1, y + 1);
        const c64 st = U0[sy] * conj(U1[sx]) * conj(U0[s]) + conj(U0[symx]) * conj(U1[smx]) * U0[smx];
        F[V + s] = -beta * (U1[s] * st).im;
    }
}
void launch_gauge_force(int L, const c64* U, double beta, double* F, hipStream_t st) {
    k_gauge_force<<<grid_for(L * L), kBlock, 0, st>>>(L, U, beta, F);
    SMG_LAUNCH_CHECK();
}

// ----------------------------------------------------------------- instantiations
#define SMG_INST(T)                                                                                                 \
    template void launch_dslash<T>(const DiracView<T>&, cplx<T>*, const cplx<T>*, T, int, int, Epi, const cplx<T>*, \
                                   const cplx<T>*, const Fin&, int, size_t, hipStream_t);                            \
    template void launch_eo_hop<T>(const DiracView<T>&, cplx<T>*, const cplx<T>*, int, int, T, int, int,             \
                                   const cplx<T>*, Epi, const cplx<T>*, const cplx<T>*, const Fin&, hipStream_t);    \
    template void launch_eo_diag<T>(const DiracView<T>&, cplx<T>*, const cplx<T>*, int, int, T, hipStream_t);       \
    template void launch_eo_gather<T>(int, cplx<T>*, const cplx<T>*, int, hipStream_t);                             \
    template void launch_eo_scatter<T>(int, cplx<T>*, const cplx<T>*, int, hipStream_t);
SMG_INST(double)
SMG_INST(float)
#undef SMG_INST

}  // namespace smg
