// lattices.hip -- many Wilson-clover lattices of different sizes in one launch.
//
// G lattices (sizes L_g, V_g = L_g^2 sites) are concatenated: their links, clover terms
// and vectors lie one after the other (vector of lattice g at complex offset
// 2 * site0_g, in the reference's psi[2s + spin] layout).  Workgroups are assigned to
// lattices by a table (lattice g owns workgroups wg0_g .. wg0_g + nwg_g - 1, whose
// work-items stride over its sites), so no workgroup straddles two lattices and a
// reduction finishes per lattice: each workgroup stores its partial row, and the last
// workgroup of each lattice sums that lattice's rows in index order and runs its scalar
// program (deterministic, as the single-lattice reductions).  The CG on D^dag D
// then advances every lattice in lockstep with its own scalars and convergence flag,
// one launch per kernel for all of them.  FP64; every kernel is HBM-bound.
#include "smg_lattices.h"

namespace smg {


// dst = g5^g5out D(mu) g5^g5in src on every lattice; EPI_DOTS adds {<dst, aux>, |dst|^2}
// per lattice into fin
template <int EPI>
__global__ __launch_bounds__(kBlock) void k_ml_dslash(const LatDesc* __restrict__ d, int G, c64* __restrict__ dst,
                                                      const c64* __restrict__ src, double musign, int g5in, int g5out,
                                                      const c64* __restrict__ aux, LatFin fin) {
    const int g = my_lattice(d, G);
    const LatDesc D = d[g];
    const int L = D.L, V = L * L;
    const size_t o = 2 * (size_t)D.site0;
    double acc[3] = {0.0, 0.0, 0.0};
    if (fin.skip && *fin.skip) return;   // grid-uniform: every lattice has converged
    for (int s = ((int)blockIdx.x - D.wg0) * kBlock + threadIdx.x; s < V; s += D.nwg * kBlock) {
        const c64* ps = src + o;
        const int x = s % L, y = s / L;
        const int sxp = (x + 1 == L) ? s - x : s + 1;
        const int sxm = (x == 0) ? s + L - 1 : s - 1;
        const int syp = (y + 1 == L) ? x : s + L;
        const int sym = (y == 0) ? s + V - L : s - L;
        auto ld = [&](int t, c64& a, c64& b) {
            a = ps[2 * t];
            b = ps[2 * t + 1];
            if (g5in) b = -b;
        };
        c64 s0, s1, xf0, xf1, xb0, xb1, yf0, yf1, yb0, yb1;
        ld(s, s0, s1);
        ld(sxp, xf0, xf1);
        ld(sxm, xb0, xb1);
        ld(syp, yf0, yf1);
        ld(sym, yb0, yb1);
        const c64* U0 = D.U;
        const c64* U1 = D.U + V;
        const c64 u0 = U0[s], u0b = conj(U0[sxm]), u1 = U1[s], u1b = conj(U1[sym]);
        const c64 up0 = u0 * xf0, up1 = u0 * xf1, q0 = u0b * xb0, q1 = u0b * xb1;
        const c64 vp0 = u1 * yf0, vp1 = u1 * yf1, w0 = u1b * yb0, w1 = u1b * yb1;
        const double hr = 0.5 * D.r;
        const c64 sm0 = (up0 + q0) + (vp0 + w0), sm1 = (up1 + q1) + (vp1 + w1);
        const c64 t0 = (q1 - up1) + mul_i(vp1 - w1), t1 = (q0 - up0) - mul_i(vp0 - w0);
        const c64 h0 = -(hr * sm0) - 0.5 * t0, h1 = -(hr * sm1) - 0.5 * t1;
        const double C = D.C ? D.C[s] : 0.0, mu = musign * D.mu;
        c64 o0 = c64{D.diag + C, mu} * s0 + h0, o1 = c64{D.diag - C, -mu} * s1 + h1;
        if (g5out) o1 = -o1;
        dst[o + 2 * s] = o0;
        dst[o + 2 * s + 1] = o1;
        if (EPI == EPI_DOTS) {
            const c64 a0 = aux[o + 2 * s], a1 = aux[o + 2 * s + 1];
            const c64 d0 = cdot(o0, a0), d1 = cdot(o1, a1);
            acc[0] += d0.re + d1.re;
            acc[1] += d0.im + d1.im;
            acc[2] += norm2(o0) + norm2(o1);
        }
    }
    if (EPI == EPI_DOTS) finish_seg<3>(acc, fin, d, G, g);
}

// CG passes (src/solvers.cpp:16-72 per lattice): init r = p = b, x = 0 with |r|^2;
// r -= a_g Ap with |r|^2; x += a_g p, p = r + b_g p; |v|^2 per lattice.  The work-items
// of lattice g stride over its sites.
__global__ __launch_bounds__(kBlock) void k_ml_cg_init(const LatDesc* __restrict__ d, int G,
                                                       const c64* __restrict__ b, c64* __restrict__ x,
                                                       c64* __restrict__ r, c64* __restrict__ p, LatFin fin) {
    const int g = my_lattice(d, G);
    const LatDesc D = d[g];
    const int V = D.L * D.L;
    double acc[1] = {0.0};
    for (int s = ((int)blockIdx.x - D.wg0) * kBlock + threadIdx.x; s < V; s += D.nwg * kBlock) {
        const size_t i = 2 * ((size_t)D.site0 + s);
        const c64 b0 = b[i], b1 = b[i + 1];
        r[i] = b0;
        r[i + 1] = b1;
        p[i] = b0;
        p[i + 1] = b1;
        x[i] = c64{0.0, 0.0};
        x[i + 1] = c64{0.0, 0.0};
        acc[0] += norm2(b0) + norm2(b1);
    }
    finish_seg<1>(acc, fin, d, G, g);
}
__global__ __launch_bounds__(kBlock) void k_ml_cg_r(const LatDesc* __restrict__ d, int G, c64* __restrict__ r,
                                                    const c64* __restrict__ Ap, const double* __restrict__ alpha,
                                                    LatFin fin, int* __restrict__ late) {
    if (fin.skip && *fin.skip) {
        if (late && blockIdx.x == 0 && threadIdx.x == 0) *late = 1;
        return;
    }
    const int g = my_lattice(d, G);
    const LatDesc D = d[g];
    const int V = D.L * D.L;
    const c64 a{alpha[2 * g], alpha[2 * g + 1]};
    double acc[1] = {0.0};
    for (int s = ((int)blockIdx.x - D.wg0) * kBlock + threadIdx.x; s < V; s += D.nwg * kBlock) {
        const size_t i = 2 * ((size_t)D.site0 + s);
        const c64 r0 = r[i] - a * Ap[i], r1 = r[i + 1] - a * Ap[i + 1];
        r[i] = r0;
        r[i + 1] = r1;
        acc[0] += norm2(r0) + norm2(r1);
    }
    finish_seg<1>(acc, fin, d, G, g);
}
__global__ __launch_bounds__(kBlock) void k_ml_cg_xp(const LatDesc* __restrict__ d, int G, c64* __restrict__ x,
                                                     c64* __restrict__ p, const c64* __restrict__ r,
                                                     const double* __restrict__ alpha,
                                                     const double* __restrict__ beta, const int* __restrict__ late) {
    if (late && *late) return;
    const int g = my_lattice(d, G);
    const LatDesc D = d[g];
    const int V = D.L * D.L;
    const c64 a{alpha[2 * g], alpha[2 * g + 1]}, bt{beta[2 * g], beta[2 * g + 1]};
    for (int s = ((int)blockIdx.x - D.wg0) * kBlock + threadIdx.x; s < V; s += D.nwg * kBlock) {
        const size_t i = 2 * ((size_t)D.site0 + s);
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const c64 pv = p[i + q];
            x[i + q] = x[i + q] + a * pv;
            p[i + q] = r[i + q] + bt * pv;
        }
    }
}
__global__ __launch_bounds__(kBlock) void k_ml_norm2(const LatDesc* __restrict__ d, int G, const c64* __restrict__ v,
                                                     LatFin fin) {
    const int g = my_lattice(d, G);
    const LatDesc D = d[g];
    const int V = D.L * D.L;
    double acc[1] = {0.0};
    for (int s = ((int)blockIdx.x - D.wg0) * kBlock + threadIdx.x; s < V; s += D.nwg * kBlock) {
        const size_t i = 2 * ((size_t)D.site0 + s);
        acc[0] += norm2(v[i]) + norm2(v[i + 1]);
    }
    finish_seg<1>(acc, fin, d, G, g);
}
// per-lattice CG state: aux_g[5] = tol^2, flag_g = 0; *done = *late = 0
__global__ void k_ml_cg_state(double* aux, int G, double tol2, int* done, int* late) {
    for (int g = threadIdx.x; g < G; g += blockDim.x) {
        aux[(size_t)g * 16 + 5] = tol2;
        *reinterpret_cast<int*>(aux + (size_t)g * 16 + 8) = 0;
    }
    if (threadIdx.x == 0) {
        *done = 0;
        *late = 0;
    }
}

void launch_ml_dslash(const LatDesc* d, int G, int nwg, c64* dst, const c64* src, double musign, int g5in, int g5out,
                      Epi epi, const c64* aux, const LatFin& fin, hipStream_t st) {
    if (epi == EPI_DOTS)
        k_ml_dslash<EPI_DOTS><<<nwg, kBlock, 0, st>>>(d, G, dst, src, musign, g5in, g5out, aux, fin);
    else
        k_ml_dslash<EPI_NONE><<<nwg, kBlock, 0, st>>>(d, G, dst, src, musign, g5in, g5out, aux, fin);
    SMG_LAUNCH_CHECK();
}
void launch_ml_cg_state(double* aux, int G, double tol2, int* done, int* late, hipStream_t st) {
    k_ml_cg_state<<<1, 64, 0, st>>>(aux, G, tol2, done, late);
    SMG_LAUNCH_CHECK();
}
void launch_ml_cg_init(const LatDesc* d, int G, int nwg, const c64* b, c64* x, c64* r, c64* p, const LatFin& fin,
                       hipStream_t st) {
    k_ml_cg_init<<<nwg, kBlock, 0, st>>>(d, G, b, x, r, p, fin);
    SMG_LAUNCH_CHECK();
}
void launch_ml_cg_r(const LatDesc* d, int G, int nwg, c64* r, const c64* Ap, const double* alpha, const LatFin& fin,
                    int* late, hipStream_t st) {
    k_ml_cg_r<<<nwg, kBlock, 0, st>>>(d, G, r, Ap, alpha, fin, late);
    SMG_LAUNCH_CHECK();
}
void launch_ml_cg_xp(const LatDesc* d, int G, int nwg, c64* x, c64* p, const c64* r, const double* alpha,
                     const double* beta, const int* late, hipStream_t st) {
    k_ml_cg_xp<<<nwg, kBlock, 0, st>>>(d, G, x, p, r, alpha, beta, late);
    SMG_LAUNCH_CHECK();
}
void launch_ml_norm2(const LatDesc* d, int G, int nwg, const c64* v, const LatFin& fin, hipStream_t st) {
    k_ml_norm2<<<nwg, kBlock, 0, st>>>(d, G, v, fin);
    SMG_LAUNCH_CHECK();
}

}  // namespace smg
