/*
 * oracle.c -- CPU FP64 restatement of the Schwinger_MG hot path (C99 + OpenMP).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library and the timed CPU
 * baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg load it.  Each function cites the reference file:line it restates; loop
 * order and arithmetic order follow the reference so that, single-threaded, most
 * results agree with the compiled reference to the last bit.
 *
 * Deviation (documented, deliberate): the reference builds its Galerkin
 * operators by probing one coarse column at a time (src/coarse_op.h:26-43,
 * 241-279); orc_coarse_build(colored=1) probes one colour class of aggregates
 * at a time instead (9*k probes instead of dim), which yields the same 9-point
 * stencil whenever every coarse direction is at least 3 aggregates long.
 */
#include "oracle.h"
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#define OMP_MIN 4096
#define OMP_FOR _Pragma("omp parallel for schedule(static) if(n > OMP_MIN)")

static void* xcalloc(size_t n, size_t sz) {
    void* p = calloc(n ? n : 1, sz);
    if (!p) abort();
    return p;
}

/* ================================================================ geometry */
/* src/lattice.h:19-26 */
static inline int site_idx(int L, int x, int y) { return ((x % L + L) % L) + L * ((y % L + L) % L); }
static inline int xp(int L, int s) { return site_idx(L, s % L + 1, s / L); }
static inline int xm(int L, int s) { return site_idx(L, s % L - 1, s / L); }
static inline int yp(int L, int s) { return site_idx(L, s % L, s / L + 1); }
static inline int ym(int L, int s) { return site_idx(L, s % L, s / L - 1); }

/* src/lattice.h:28-41 */
void orc_eo_maps(int L, int* even_sites, int* odd_sites, int* eo_index) {
    int V = L * L, ne = 0, no = 0;
    for (int s = 0; s < V; s++) {
        int par = ((s % L) + (s / L)) % 2;
        if (par == 0) { eo_index[s] = ne; even_sites[ne++] = s; }
        else          { eo_index[s] = no; odd_sites[no++] = s; }
    }
}

/* src/gauge.h:26-29 */
static inline ocx plaq(int L, const ocx* U, int s) {
    int V = L * L;
    int sp0 = xp(L, s), sp1 = yp(L, s);
    return U[s] * U[V + sp0] * conj(U[sp1]) * conj(U[V + s]);
}

/* src/gauge.h:31-36 */
double orc_avg_plaq(int L, const ocx* U) {
    int V = L * L;
    double acc = 0;
    for (int i = 0; i < V; i++) acc += creal(plaq(L, U, i));
    return acc / V;
}

/* src/hmc_forces.cpp:80-86 */
double orc_gauge_action(int L, const ocx* U, double beta) {
    int V = L * L;
    double acc = 0;
    for (int i = 0; i < V; i++) acc += creal(plaq(L, U, i));
    return -beta * acc;
}

/* src/hmc_forces.cpp:88-111 */
void orc_gauge_force(int L, const ocx* U, double beta, double* F) {
    int V = L * L, n = V;
    const ocx* U0 = U; const ocx* U1 = U + V;
    OMP_FOR
    for (int s = 0; s < V; s++) {
        {
            int sx = xp(L, s), sy = yp(L, s), smy = ym(L, s);
            int sxmy = ym(L, sx);
            ocx staple = U1[sx] * conj(U0[sy]) * conj(U1[s]) + conj(U1[sxmy]) * conj(U0[smy]) * U1[smy];
            F[s] = -beta * cimag(U0[s] * staple);
        }
        {
            int sy = yp(L, s), sx = xp(L, s), smx = xm(L, s);
            int symx = xm(L, sy);
            ocx staple = U0[sy] * conj(U1[sx]) * conj(U0[s]) + conj(U0[symx]) * conj(U1[smx]) * U0[smx];
            F[V + s] = -beta * cimag(U1[s] * staple);
        }
    }
}

/* src/dirac.h:26-52 */
void orc_clover(int L, const ocx* U, double* F01) {
    int V = L * L, n = V;
    const ocx* U0 = U; const ocx* U1 = U + V;
    OMP_FOR
    for (int s = 0; s < V; s++) {
        int sxp = xp(L, s), sxm = xm(L, s), syp = yp(L, s), sym = ym(L, s);
        ocx P_ur = U0[s] * U1[sxp] * conj(U0[syp]) * conj(U1[s]);
        int sxmyp = yp(L, sxm);
        ocx P_ul = U1[s] * conj(U0[sxmyp]) * conj(U1[sxm]) * U0[sxm];
        int sxmym = ym(L, sxm);
        ocx P_ll = conj(U0[sxm]) * conj(U1[sxmym]) * U0[sxmym] * U1[sym];
        int sxpym = ym(L, sxp);
        ocx P_lr = conj(U1[sym]) * U0[sym] * U1[sxpym] * conj(U0[s]);
        ocx Q = P_ur + P_ul + P_ll + P_lr;
        F01[s] = cimag(Q) / 4.0;
    }
}

/* ================================================================ Dirac */
struct orc_dirac {
    int L, V;
    ocx* U;
    double* F01;
    double mass, r, csw, mu_t;
    int *even, *odd, *eoi;
};

orc_dirac* orc_dirac_new(int L, const ocx* U, double mass, double r, double csw, double mu_t) {
    orc_dirac* D = xcalloc(1, sizeof(orc_dirac));
    D->L = L; D->V = L * L;
    D->U = xcalloc(2 * (size_t)D->V, sizeof(ocx));
    memcpy(D->U, U, 2 * (size_t)D->V * sizeof(ocx));
    D->mass = mass; D->r = r; D->csw = csw; D->mu_t = mu_t;
    D->F01 = xcalloc(D->V, sizeof(double));
    if (csw != 0.0) orc_clover(L, D->U, D->F01);   /* src/dirac.h:17-22 */
    D->even = xcalloc(D->V / 2 + 1, sizeof(int));
    D->odd = xcalloc(D->V / 2 + 1, sizeof(int));
    D->eoi = xcalloc(D->V, sizeof(int));
    orc_eo_maps(L, D->even, D->odd, D->eoi);
    return D;
}

void orc_dirac_free(orc_dirac* D) {
    if (!D) return;
    free(D->U); free(D->F01); free(D->even); free(D->odd); free(D->eoi); free(D);
}

const double* orc_dirac_clover(const orc_dirac* D) { return D->F01; }

/* src/dirac.h:58-89 -- the four Wilson hops at s, with optional phase scales */
static inline void accumulate_hops(ocx* d0, ocx* d1, int s, const ocx* src, const orc_dirac* D,
                                   ocx fwd_x, ocx bwd_x, ocx fwd_y, ocx bwd_y) {
    const int L = D->L, V = D->V;
    const double r = D->r;
    const ocx* U0 = D->U; const ocx* U1 = D->U + V;
    const ocx ii = I;
    { int sf = xp(L, s);
      ocx u = fwd_x * U0[s];
      ocx up0 = u * src[2 * sf], up1 = u * src[2 * sf + 1];
      *d0 -= 0.5 * (r * up0 - up1);
      *d1 -= 0.5 * (-up0 + r * up1); }
    { int sb = xm(L, s);
      ocx ud = bwd_x * conj(U0[sb]);
      ocx udq0 = ud * src[2 * sb], udq1 = ud * src[2 * sb + 1];
      *d0 -= 0.5 * (r * udq0 + udq1);
      *d1 -= 0.5 * (udq0 + r * udq1); }
    { int sf = yp(L, s);
      ocx u = fwd_y * U1[s];
      ocx up0 = u * src[2 * sf], up1 = u * src[2 * sf + 1];
      *d0 -= 0.5 * (r * up0 + ii * up1);
      *d1 -= 0.5 * (-ii * up0 + r * up1); }
    { int sb = ym(L, s);
      ocx ud = bwd_y * conj(U1[sb]);
      ocx udq0 = ud * src[2 * sb], udq1 = ud * src[2 * sb + 1];
      *d0 -= 0.5 * (r * udq0 - ii * udq1);
      *d1 -= 0.5 * (ii * udq0 + r * udq1); }
}

/* src/dirac.h:93-113 */
static void apply_impl(const orc_dirac* D, const ocx* src, ocx* dst, double mu_eff) {
    int V = D->V, n = V;
    int has_clover = (D->csw != 0.0), has_twist = (mu_eff != 0.0);
    OMP_FOR
    for (int s = 0; s < V; s++) {
        ocx s0 = src[2 * s], s1 = src[2 * s + 1];
        double diag = 2.0 * D->r + D->mass;
        ocx a0, a1;
        if (has_clover || has_twist) {
            double C = has_clover ? 0.5 * D->csw * D->F01[s] : 0.0;
            ocx e0 = CMPLX(diag + C, mu_eff);
            ocx e1 = CMPLX(diag - C, -mu_eff);
            a0 = e0 * s0; a1 = e1 * s1;
        } else {
            a0 = diag * s0; a1 = diag * s1;
        }
        accumulate_hops(&a0, &a1, s, src, D, 1.0, 1.0, 1.0, 1.0);
        dst[2 * s] = a0; dst[2 * s + 1] = a1;
    }
}

/* src/dirac.h:120-134: D^dag = g5 D(-mu) g5 */
static void apply_dag(const orc_dirac* D, const ocx* src, ocx* dst) {
    int V = D->V, n = V;
    ocx* g5src = xcalloc(2 * (size_t)V, sizeof(ocx));
    ocx* g5dst = xcalloc(2 * (size_t)V, sizeof(ocx));
    OMP_FOR
    for (int s = 0; s < V; s++) { g5src[2 * s] = src[2 * s]; g5src[2 * s + 1] = -src[2 * s + 1]; }
    apply_impl(D, g5src, g5dst, -D->mu_t);
    OMP_FOR
    for (int s = 0; s < V; s++) { dst[2 * s] = g5dst[2 * s]; dst[2 * s + 1] = -g5dst[2 * s + 1]; }
    free(g5src); free(g5dst);
}

void orc_dirac_apply(const orc_dirac* D, int kind, const ocx* src, ocx* dst) {
    int V = D->V, n = V;
    switch (kind) {
    case ORC_D: apply_impl(D, src, dst, D->mu_t); break;      /* src/dirac.h:115-117 */
    case ORC_DDAG: apply_dag(D, src, dst); break;
    case ORC_DDAGD: {                                          /* src/dirac.h:136-140 */
        ocx* tmp = xcalloc(2 * (size_t)V, sizeof(ocx));
        apply_impl(D, src, tmp, D->mu_t);
        apply_dag(D, tmp, dst);
        free(tmp);
        break; }
    case ORC_G5D:                                              /* src/dirac.h:145-152 */
        apply_impl(D, src, dst, D->mu_t);
        OMP_FOR
        for (int s = 0; s < V; s++) dst[2 * s + 1] = -dst[2 * s + 1];
        break;
    default: abort();
    }
}

/* src/dirac.h:159-174 */
void orc_dirac_delta(const orc_dirac* D, const ocx* src, ocx* dst, const double* pi, double dt) {
    int V = D->V, n = V, L = D->L;
    OMP_FOR
    for (int s = 0; s < V; s++) {
        ocx a0 = 0.0, a1 = 0.0;
        ocx fwd_x = cexp(CMPLX(0, dt * pi[s])) - 1.0;
        ocx bwd_x = cexp(CMPLX(0, -dt * pi[xm(L, s)])) - 1.0;
        ocx fwd_y = cexp(CMPLX(0, dt * pi[V + s])) - 1.0;
        ocx bwd_y = cexp(CMPLX(0, -dt * pi[V + ym(L, s)])) - 1.0;
        accumulate_hops(&a0, &a1, s, src, D, fwd_x, bwd_x, fwd_y, bwd_y);
        dst[2 * s] = a0; dst[2 * s + 1] = a1;
    }
}

/* ================================================================ even-odd */
/* src/dirac.h:198-250: dst (dst_sites parity) = hopping of src (other parity) */
static void eo_hop(const orc_dirac* D, const ocx* src, ocx* dst, const int* dst_sites) {
    int nh = D->V / 2, n = nh * 8, L = D->L, V = D->V;
    const double r = D->r;
    const ocx* U0 = D->U; const ocx* U1 = D->U + V;
    const ocx ii = I;
    OMP_FOR
    for (int id = 0; id < nh; id++) {
        int s = dst_sites[id];
        ocx a0 = 0.0, a1 = 0.0;
        { int is = D->eoi[xp(L, s)];
          ocx u = U0[s];
          ocx up0 = u * src[2 * is], up1 = u * src[2 * is + 1];
          a0 -= 0.5 * (r * up0 - up1);
          a1 -= 0.5 * (-up0 + r * up1); }
        { int sb = xm(L, s); int is = D->eoi[sb];
          ocx ud = conj(U0[sb]);
          ocx udq0 = ud * src[2 * is], udq1 = ud * src[2 * is + 1];
          a0 -= 0.5 * (r * udq0 + udq1);
          a1 -= 0.5 * (udq0 + r * udq1); }
        { int is = D->eoi[yp(L, s)];
          ocx u = U1[s];
          ocx up0 = u * src[2 * is], up1 = u * src[2 * is + 1];
          a0 -= 0.5 * (r * up0 + ii * up1);
          a1 -= 0.5 * (-ii * up0 + r * up1); }
        { int sb = ym(L, s); int is = D->eoi[sb];
          ocx ud = conj(U1[sb]);
          ocx udq0 = ud * src[2 * is], udq1 = ud * src[2 * is + 1];
          a0 -= 0.5 * (r * udq0 - ii * udq1);
          a1 -= 0.5 * (ii * udq0 + r * udq1); }
        dst[2 * id] = a0; dst[2 * id + 1] = a1;
    }
}

/* src/dirac.h:277-300, 370-380 */
static void diag_inv(const orc_dirac* D, const ocx* src, ocx* dst, const int* sites, double mu_eff) {
    int nh = D->V / 2, n = nh * 8;
    double diag = 2.0 * D->r + D->mass;
    int hc = (D->csw != 0.0), ht = (mu_eff != 0.0);
    if (!hc && !ht) {
        double inv = 1.0 / diag;
        OMP_FOR
        for (int i = 0; i < nh; i++) { dst[2 * i] = inv * src[2 * i]; dst[2 * i + 1] = inv * src[2 * i + 1]; }
        return;
    }
    OMP_FOR
    for (int i = 0; i < nh; i++) {
        int s = sites[i];
        double C = hc ? 0.5 * D->csw * D->F01[s] : 0.0;
        ocx d0 = CMPLX(diag + C, mu_eff), d1 = CMPLX(diag - C, -mu_eff);
        dst[2 * i] = src[2 * i] / d0;
        dst[2 * i + 1] = src[2 * i + 1] / d1;
    }
}

/* src/dirac.h:302-323, 382-392 */
static void diag_mul(const orc_dirac* D, const ocx* src, ocx* dst, const int* sites, double mu_eff) {
    int nh = D->V / 2, n = nh * 8;
    double diag = 2.0 * D->r + D->mass;
    int hc = (D->csw != 0.0), ht = (mu_eff != 0.0);
    OMP_FOR
    for (int i = 0; i < nh; i++) {
        if (!hc && !ht) { dst[2 * i] = diag * src[2 * i]; dst[2 * i + 1] = diag * src[2 * i + 1]; continue; }
        int s = sites[i];
        double C = hc ? 0.5 * D->csw * D->F01[s] : 0.0;
        ocx d0 = CMPLX(diag + C, mu_eff), d1 = CMPLX(diag - C, -mu_eff);
        dst[2 * i] = d0 * src[2 * i];
        dst[2 * i + 1] = d1 * src[2 * i + 1];
    }
}

/* src/dirac.h:399-415 */
static void schur_impl(const orc_dirac* D, const ocx* src_o, ocx* dst_o, double mu_eff) {
    int nh = D->V / 2, n = nh * 8;
    ocx* t1 = xcalloc(2 * (size_t)nh, sizeof(ocx));
    ocx* t2 = xcalloc(2 * (size_t)nh, sizeof(ocx));
    eo_hop(D, src_o, t1, D->even);            /* apply_oe */
    diag_inv(D, t1, t2, D->even, mu_eff);
    eo_hop(D, t2, dst_o, D->odd);             /* apply_eo */
    double diag = 2.0 * D->r + D->mass;
    int hc = (D->csw != 0.0), ht = (mu_eff != 0.0);
    OMP_FOR
    for (int i = 0; i < nh; i++) {            /* schur_diag_loop, src/dirac.h:325-346 */
        if (!hc && !ht) {
            dst_o[2 * i] = diag * src_o[2 * i] - dst_o[2 * i];
            dst_o[2 * i + 1] = diag * src_o[2 * i + 1] - dst_o[2 * i + 1];
            continue;
        }
        int s = D->odd[i];
        double C = hc ? 0.5 * D->csw * D->F01[s] : 0.0;
        ocx d0 = CMPLX(diag + C, mu_eff), d1 = CMPLX(diag - C, -mu_eff);
        dst_o[2 * i] = d0 * src_o[2 * i] - dst_o[2 * i];
        dst_o[2 * i + 1] = d1 * src_o[2 * i + 1] - dst_o[2 * i + 1];
    }
    free(t1); free(t2);
}

/* src/dirac.h:422-437 */
static void schur_dag(const orc_dirac* D, const ocx* src_o, ocx* dst_o) {
    int nh = D->V / 2, n = nh * 8;
    ocx* g = xcalloc(2 * (size_t)nh, sizeof(ocx));
    ocx* h = xcalloc(2 * (size_t)nh, sizeof(ocx));
    OMP_FOR
    for (int i = 0; i < nh; i++) { g[2 * i] = src_o[2 * i]; g[2 * i + 1] = -src_o[2 * i + 1]; }
    schur_impl(D, g, h, -D->mu_t);
    OMP_FOR
    for (int i = 0; i < nh; i++) { dst_o[2 * i] = h[2 * i]; dst_o[2 * i + 1] = -h[2 * i + 1]; }
    free(g); free(h);
}

void orc_eo_apply(const orc_dirac* D, int kind, const ocx* src, ocx* dst) {
    int nh = D->V / 2;
    switch (kind) {
    case ORC_EO_HOP_EO: eo_hop(D, src, dst, D->odd); break;            /* src/dirac.h:253-255 */
    case ORC_EO_HOP_OE: eo_hop(D, src, dst, D->even); break;           /* src/dirac.h:258-260 */
    case ORC_EO_EE_INV: diag_inv(D, src, dst, D->even, D->mu_t); break; /* src/dirac.h:263-265 */
    case ORC_EO_OO_INV: diag_inv(D, src, dst, D->odd, D->mu_t); break;
    case ORC_EO_OO: diag_mul(D, src, dst, D->odd, D->mu_t); break;     /* src/dirac.h:394-396 */
    case ORC_EO_SCHUR: schur_impl(D, src, dst, D->mu_t); break;        /* src/dirac.h:417-419 */
    case ORC_EO_SCHUR_DAG: schur_dag(D, src, dst); break;
    case ORC_EO_SCHUR_DAG_SCHUR: {                                      /* src/dirac.h:440-445 */
        ocx* t = xcalloc(2 * (size_t)nh, sizeof(ocx));
        schur_impl(D, src, t, D->mu_t);
        schur_dag(D, t, dst);
        free(t);
        break; }
    default: abort();
    }
}

/* src/dirac.h:448-463 */
static void eo_scatter(const orc_dirac* D, const ocx* he, const ocx* ho, ocx* full) {
    int nh = D->V / 2;
    for (int i = 0; i < nh; i++) {
        int se = D->even[i];
        full[2 * se] = he[2 * i]; full[2 * se + 1] = he[2 * i + 1];
    }
    for (int i = 0; i < nh; i++) {
        int so = D->odd[i];
        full[2 * so] = ho[2 * i]; full[2 * so + 1] = ho[2 * i + 1];
    }
}

/* src/dirac.h:541-552 */
void orc_eo_reconstruct(const orc_dirac* D, const ocx* x_o, const ocx* phi, ocx* out) {
    int nh = D->V / 2;
    ocx* hop = xcalloc(2 * (size_t)nh, sizeof(ocx));
    ocx* xe = xcalloc(2 * (size_t)nh, sizeof(ocx));
    eo_hop(D, x_o, hop, D->even);
    for (int i = 0; i < nh; i++) {
        int s = D->even[i];
        hop[2 * i] = phi[2 * s] - hop[2 * i];
        hop[2 * i + 1] = phi[2 * s + 1] - hop[2 * i + 1];
    }
    diag_inv(D, hop, xe, D->even, D->mu_t);
    eo_scatter(D, xe, x_o, out);
    free(hop); free(xe);
}

/* src/dirac.h:515-523 */
void orc_eo_reconstruct_full(const orc_dirac* D, const ocx* x_o, ocx* out) {
    int nh = D->V / 2;
    ocx* hop = xcalloc(2 * (size_t)nh, sizeof(ocx));
    ocx* xe = xcalloc(2 * (size_t)nh, sizeof(ocx));
    eo_hop(D, x_o, hop, D->even);
    for (int i = 0; i < 2 * nh; i++) hop[i] = -hop[i];
    diag_inv(D, hop, xe, D->even, D->mu_t);
    eo_scatter(D, xe, x_o, out);
    free(hop); free(xe);
}

/* src/dirac.h:348-366, 528-537 */
double orc_eo_logdet(const orc_dirac* D) {
    double diag = 2.0 * D->r + D->mass;
    int nh = D->V / 2;
    int hc = (D->csw != 0.0), ht = (D->mu_t != 0.0);
    if (!hc && !ht) return nh * 2.0 * log(diag);
    double mu2 = D->mu_t * D->mu_t, acc = 0;
    for (int i = 0; i < nh; i++) {
        int s = D->even[i];
        double C = hc ? 0.5 * D->csw * D->F01[s] : 0.0;
        double d0 = diag + C, d1 = diag - C;
        acc += 0.5 * log(d0 * d0 + mu2) + 0.5 * log(d1 * d1 + mu2);
    }
    return acc;
}

/* ================================================================ BLAS-1 */
/* src/linalg.cpp:8-18 */
ocx orc_dot(int n, const ocx* a, const ocx* b) {
    double re = 0, im = 0;
    #pragma omp parallel for reduction(+:re,im) schedule(static) if(n > OMP_MIN)
    for (int i = 0; i < n; i++) {
        ocx p = conj(a[i]) * b[i];
        re += creal(p); im += cimag(p);
    }
    return CMPLX(re, im);
}
void orc_dot2(int n, const ocx* a, const ocx* b, double* out) {
    ocx d = orc_dot(n, a, b);
    out[0] = creal(d); out[1] = cimag(d);
}
/* src/linalg.cpp:20 */
double orc_norm(int n, const ocx* a) { return sqrt(creal(orc_dot(n, a, a))); }
/* src/linalg.cpp:22-26 */
static void axpy(int n, ocx a, const ocx* x, ocx* y) {
    OMP_FOR
    for (int i = 0; i < n; i++) y[i] += a * x[i];
}
static void vscale(int n, ocx* v, ocx a) {
    OMP_FOR
    for (int i = 0; i < n; i++) v[i] *= a;
}

/* ================================================================ operators */
struct orc_op {
    int n;
    void (*fn)(const orc_op*, const ocx*, ocx*);
    const void* ctx;
    int kind;
};
static void op_dirac_fn(const orc_op* A, const ocx* s, ocx* d) {
    const orc_dirac* D = A->ctx;
    if (A->kind == 100) orc_eo_apply(D, ORC_EO_SCHUR_DAG_SCHUR, s, d);
    else orc_dirac_apply(D, A->kind, s, d);
}
orc_op* orc_op_dirac(const orc_dirac* D, int kind) {
    orc_op* A = xcalloc(1, sizeof(orc_op));
    A->n = (kind == 100) ? D->V : 2 * D->V;
    A->fn = op_dirac_fn; A->ctx = D; A->kind = kind;
    return A;
}
void orc_op_apply(const orc_op* A, const ocx* src, ocx* dst) { A->fn(A, src, dst); }
int orc_op_dim(const orc_op* A) { return A->n; }
void orc_op_free(orc_op* A) { free(A); }

/* ================================================================ prolongator */
/* src/prolongator.h:14-142 (geometric, nd = 2 fine dof per site) and
 * src/prolongator.h:149-228 (linear blocking of block_k consecutive coarse dofs).
 * The linear blocking of 4*k consecutive dofs on a coarse lattice whose x extent
 * is a multiple of 4 is the geometric 4x1 aggregation with nd = k, so one
 * geometric implementation with nd dof per site restates both. */
struct orc_prolong {
    int nx, ny, nd, bx, by, k, nbx, nby, nblocks, bsites, bsize;
    int chiral;       /* 1: k/2 null vectors, each split into its two chiralities */
    ocx* bv;          /* [nblocks][k][bsize], bsize = bx*by*nd */
    int* s2b; int* s2l;
    int* blist;       /* [nblocks][bsites] ascending site index */
};

orc_prolong* orc_prolong_new(int nx, int ny, int nd, int bx, int by, int k, int chiral) {
    orc_prolong* P = xcalloc(1, sizeof(orc_prolong));
    P->nx = nx; P->ny = ny; P->nd = nd; P->bx = bx; P->by = by; P->k = k; P->chiral = chiral;
    P->nbx = nx / bx; P->nby = ny / by; P->nblocks = P->nbx * P->nby;
    P->bsites = bx * by; P->bsize = bx * by * nd;
    P->bv = xcalloc((size_t)P->nblocks * k * P->bsize, sizeof(ocx));
    int V = nx * ny;
    P->s2b = xcalloc(V, sizeof(int)); P->s2l = xcalloc(V, sizeof(int));
    P->blist = xcalloc((size_t)V, sizeof(int));
    int* fill = xcalloc(P->nblocks, sizeof(int));
    for (int s = 0; s < V; s++) {                 /* src/prolongator.h:50-57 */
        int x = s % nx, y = s / nx;
        int b = x / bx + P->nbx * (y / by);
        P->s2b[s] = b; P->s2l[s] = x % bx + bx * (y % by);
        P->blist[b * P->bsites + fill[b]++] = s;
    }
    free(fill);
    return P;
}
void orc_prolong_free(orc_prolong* P) {
    if (!P) return;
    free(P->bv); free(P->s2b); free(P->s2l); free(P->blist); free(P);
}
int orc_prolong_coarse_dim(const orc_prolong* P) { return P->nblocks * P->k; }
void orc_prolong_get(const orc_prolong* P, ocx* bv) {
    memcpy(bv, P->bv, (size_t)P->nblocks * P->k * P->bsize * sizeof(ocx));
}
void orc_prolong_set(orc_prolong* P, const ocx* bv) {
    memcpy(P->bv, bv, (size_t)P->nblocks * P->k * P->bsize * sizeof(ocx));
}

/* src/prolongator.h:111-130 */
static void orthonormalise_block(orc_prolong* P, int b) {
    int bd = P->bsize;
    ocx* base = P->bv + (size_t)b * P->k * bd;
    for (int k = 0; k < P->k; k++) {
        ocx* v = base + (size_t)k * bd;
        for (int j = 0; j < k; j++) {
            const ocx* w = base + (size_t)j * bd;
            ocx proj = 0;
            for (int i = 0; i < bd; i++) proj += conj(w[i]) * v[i];
            for (int i = 0; i < bd; i++) v[i] -= proj * w[i];
        }
        double nn = 0;
        for (int i = 0; i < bd; i++) nn += creal(v[i]) * creal(v[i]) + cimag(v[i]) * cimag(v[i]);
        nn = sqrt(nn);
        if (nn > 1e-14)
            for (int i = 0; i < bd; i++) v[i] /= nn;
    }
}

/* src/prolongator.h:63-74, 133-141.  Chiral aggregation (no reference
 * counterpart): coarse dof j < k/2 holds the gamma5 = +1 part of null vector j
 * (the first nd/2 dof of every site), coarse dof k/2 + j its gamma5 = -1 part;
 * the per-block Gram-Schmidt then runs over all k as usual (the two halves are
 * orthogonal by support). */
void orc_prolong_build(orc_prolong* P, const ocx* null_vecs) {
    int V = P->nx * P->ny, nd = P->nd, nv = P->chiral ? P->k / 2 : P->k;
    size_t fd = (size_t)V * nd;
    for (int b = 0; b < P->nblocks; b++) {
        for (int k = 0; k < P->k; k++) {
            ocx* out = P->bv + ((size_t)b * P->k + k) * P->bsize;
            int src = k % nv, chi = P->chiral ? k / nv : -1;
            for (int t = 0; t < P->bsites; t++) {
                int s = P->blist[b * P->bsites + t];
                int loc = P->s2l[s];
                for (int c = 0; c < nd; c++) {
                    int keep = chi < 0 || (c >= nd / 2) == chi;
                    out[loc * nd + c] = keep ? null_vecs[src * fd + (size_t)s * nd + c] : 0.0;
                }
            }
        }
        orthonormalise_block(P, b);
    }
}

/* src/prolongator.h:77-90 */
void orc_prolong_apply(const orc_prolong* P, const ocx* coarse, ocx* fine) {
    int V = P->nx * P->ny, n = V, nd = P->nd;
    OMP_FOR
    for (int s = 0; s < V; s++) {
        int b = P->s2b[s], loc = P->s2l[s];
        for (int c = 0; c < nd; c++) fine[(size_t)s * nd + c] = 0.0;
        for (int k = 0; k < P->k; k++) {
            ocx cc = coarse[b * P->k + k];
            const ocx* v = P->bv + ((size_t)b * P->k + k) * P->bsize + loc * nd;
            for (int c = 0; c < nd; c++) fine[(size_t)s * nd + c] += cc * v[c];
        }
    }
}

/* src/prolongator.h:94-108 */
void orc_restrict_apply(const orc_prolong* P, const ocx* fine, ocx* coarse) {
    int nb = P->nblocks, n = nb * 16, nd = P->nd;
    OMP_FOR
    for (int b = 0; b < nb; b++) {
        for (int k = 0; k < P->k; k++) coarse[b * P->k + k] = 0.0;
        for (int t = 0; t < P->bsites; t++) {
            int s = P->blist[b * P->bsites + t];
            int loc = P->s2l[s];
            for (int k = 0; k < P->k; k++) {
                const ocx* v = P->bv + ((size_t)b * P->k + k) * P->bsize + loc * nd;
                ocx acc = conj(v[0]) * fine[(size_t)s * nd];
                for (int c = 1; c < nd; c++) acc += conj(v[c]) * fine[(size_t)s * nd + c];
                coarse[b * P->k + k] += acc;
            }
        }
    }
}

/* ================================================================ coarse op */
struct orc_coarse {
    int nx, ny, k, nb;
    ocx* link;   /* [nb][9][k*k] row-major k x k blocks */
};

/* src/coarse_op.h:205-222 */
static inline int cnb(int nbx, int nby, int b, int dir) {
    static const int DX[9] = {0, 1, -1, 0, 0, 1, 1, -1, -1};
    static const int DY[9] = {0, 0, 0, 1, -1, 1, -1, 1, -1};
    int bxi = b % nbx, byi = b / nbx;
    int nx = (bxi + DX[dir] + nbx) % nbx, ny = (byi + DY[dir] + nby) % nby;
    return nx + ny * nbx;
}
/* src/coarse_op.h:225-229 */
static const int REVDIR[9] = {0, 2, 1, 4, 3, 8, 7, 6, 5};

static int color_period(int nb) {
    if (nb < 3) return nb;
    for (int c = 3; c <= nb; c++) if (nb % c == 0) return c;
    return nb;
}

/* src/coarse_op.h:241-279 (colored = 0); colored = 1 probes every aggregate of
 * one colour class at once. */
orc_coarse* orc_coarse_build(const orc_op* A, const orc_prolong* P, int colored) {
    orc_coarse* C = xcalloc(1, sizeof(orc_coarse));
    int k = P->k, nbx = P->nbx, nby = P->nby, nb = P->nblocks, dim = nb * k;
    C->nx = nbx; C->ny = nby; C->k = k; C->nb = nb;
    C->link = xcalloc((size_t)nb * 9 * k * k, sizeof(ocx));
    int fdim = A->n;
    ocx* ej = xcalloc(dim, sizeof(ocx));
    ocx* fine = xcalloc(fdim, sizeof(ocx));
    ocx* Afine = xcalloc(fdim, sizeof(ocx));
    ocx* col = xcalloc(dim, sizeof(ocx));
    int cx_ = colored ? color_period(nbx) : nbx, cy_ = colored ? color_period(nby) : nby;
    int nclass = colored ? cx_ * cy_ : nb;
    for (int cls = 0; cls < nclass; cls++) {
        for (int kj = 0; kj < k; kj++) {
            memset(ej, 0, dim * sizeof(ocx));
            for (int J = 0; J < nb; J++) {
                int in_class = colored ? ((J % nbx) % cx_ == cls % cx_ && (J / nbx) % cy_ == cls / cx_) : (J == cls);
                if (in_class) ej[J * k + kj] = 1.0;
            }
            orc_prolong_apply(P, ej, fine);
            A->fn(A, fine, Afine);
            orc_restrict_apply(P, Afine, col);
            for (int J = 0; J < nb; J++) {
                int in_class = colored ? ((J % nbx) % cx_ == cls % cx_ && (J / nbx) % cy_ == cls / cx_) : (J == cls);
                if (!in_class) continue;
                for (int d = 0; d < 9; d++) {
                    int Ib = cnb(nbx, nby, J, d);
                    int rd = REVDIR[d];
                    /* On a coarse lattice with fewer than 3 aggregates along a
                     * direction, several stencil directions reach the same
                     * aggregate; the coupling is stored once, in the first of
                     * them, so that the stencil equals the dense Galerkin
                     * operator of src/coarse_op.h:26-43.  (The reference's
                     * SparseCoarseOp::build stores it in every such direction.) */
                    int dup = 0;
                    for (int d2 = 0; d2 < d; d2++) if (cnb(nbx, nby, J, d2) == Ib) dup = 1;
                    if (dup) continue;
                    for (int ki = 0; ki < k; ki++)
                        C->link[((size_t)Ib * 9 + rd) * k * k + ki * k + kj] = col[Ib * k + ki];
                }
            }
        }
    }
    free(ej); free(fine); free(Afine); free(col);
    return C;
}
void orc_coarse_free(orc_coarse* C) { if (C) { free(C->link); free(C); } }
void orc_coarse_get(const orc_coarse* C, ocx* link) { memcpy(link, C->link, (size_t)C->nb * 9 * C->k * C->k * sizeof(ocx)); }
void orc_coarse_set(orc_coarse* C, const ocx* link) { memcpy(C->link, link, (size_t)C->nb * 9 * C->k * C->k * sizeof(ocx)); }

/* src/coarse_op.h:282-297 */
static void coarse_apply(const orc_coarse* C, const ocx* src, ocx* dst) {
    int nb = C->nb, k = C->k, n = nb * 16;
    OMP_FOR
    for (int b = 0; b < nb; b++) {
        for (int ki = 0; ki < k; ki++) dst[b * k + ki] = 0.0;
        for (int d = 0; d < 9; d++) {
            int nbb = cnb(C->nx, C->ny, b, d);
            const ocx* Lk = C->link + ((size_t)b * 9 + d) * k * k;
            for (int ki = 0; ki < k; ki++) {
                ocx sum = 0.0;
                for (int kj = 0; kj < k; kj++) sum += Lk[ki * k + kj] * src[nbb * k + kj];
                dst[b * k + ki] += sum;
            }
        }
    }
}
static void op_coarse_fn(const orc_op* A, const ocx* s, ocx* d) { coarse_apply(A->ctx, s, d); }
orc_op* orc_op_coarse(const orc_coarse* C) {
    orc_op* A = xcalloc(1, sizeof(orc_op));
    A->n = C->nb * C->k; A->fn = op_coarse_fn; A->ctx = C; A->kind = 200;
    return A;
}

/* ---------------- dense coarse op, src/coarse_op.h:18-137 ---------------- */
struct orc_dense { int dim; ocx* mat; /* column-major: mat[j*dim + i] */ };

/* src/coarse_op.h:26-37.  P == NULL: the dense matrix of A itself. */
orc_dense* orc_dense_build(const orc_op* A, const orc_prolong* P) {
    orc_dense* M = xcalloc(1, sizeof(orc_dense));
    int dim = P ? orc_prolong_coarse_dim(P) : A->n, fdim = A->n;
    M->dim = dim;
    M->mat = xcalloc((size_t)dim * dim, sizeof(ocx));
    ocx* ej = xcalloc(dim, sizeof(ocx));
    ocx* fine = xcalloc(fdim, sizeof(ocx));
    ocx* Af = xcalloc(fdim, sizeof(ocx));
    for (int j = 0; j < dim; j++) {
        memset(ej, 0, dim * sizeof(ocx));
        ej[j] = 1.0;
        if (P) {
            orc_prolong_apply(P, ej, fine);
            A->fn(A, fine, Af);
            orc_restrict_apply(P, Af, M->mat + (size_t)j * dim);
        } else {
            A->fn(A, ej, M->mat + (size_t)j * dim);
        }
    }
    free(ej); free(fine); free(Af);
    return M;
}
void orc_dense_free(orc_dense* M) { if (M) { free(M->mat); free(M); } }
void orc_dense_get(const orc_dense* M, ocx* cols) { memcpy(cols, M->mat, (size_t)M->dim * M->dim * sizeof(ocx)); }
/* src/coarse_op.h:46-54 */
void orc_dense_apply(const orc_dense* M, const ocx* x, ocx* y) {
    int dim = M->dim;
    for (int i = 0; i < dim; i++) y[i] = 0.0;
    for (int j = 0; j < dim; j++) {
        ocx xj = x[j];
        for (int i = 0; i < dim; i++) y[i] += M->mat[(size_t)j * dim + i] * xj;
    }
}
/* src/coarse_op.h:101-137 */
void orc_dense_solve(const orc_dense* M, const ocx* b, ocx* x) {
    int n = M->dim;
    ocx** A = xcalloc(n, sizeof(ocx*));
    for (int i = 0; i < n; i++) {
        A[i] = xcalloc(n + 1, sizeof(ocx));
        for (int j = 0; j < n; j++) A[i][j] = M->mat[(size_t)j * n + i];
        A[i][n] = b[i];
    }
    for (int k = 0; k < n; k++) {
        double best = 0; int br = k;
        for (int i = k; i < n; i++) { double v = cabs(A[i][k]); if (v > best) { best = v; br = i; } }
        ocx* t = A[k]; A[k] = A[br]; A[br] = t;
        if (best < 1e-30) continue;
        for (int i = k + 1; i < n; i++) {
            ocx f = A[i][k] / A[k][k];
            for (int j = k; j <= n; j++) A[i][j] -= f * A[k][j];
        }
    }
    for (int i = n - 1; i >= 0; i--) {
        ocx s = A[i][n];
        for (int j = i + 1; j < n; j++) s -= A[i][j] * x[j];
        x[i] = (cabs(A[i][i]) > 1e-30) ? s / A[i][i] : 0.0;
    }
    for (int i = 0; i < n; i++) free(A[i]);
    free(A);
}

/* ================================================================ smoothers */
/* src/smoother.cpp:26-40 (and :9-23, the same loop with A = D^dag D) */
void orc_mr_smooth(const orc_op* A, ocx* x, const ocx* b, int niter) {
    int n = A->n;
    ocx* r = xcalloc(n, sizeof(ocx));
    ocx* Ar = xcalloc(n, sizeof(ocx));
    for (int it = 0; it < niter; it++) {
        A->fn(A, x, r);
        OMP_FOR
        for (int i = 0; i < n; i++) r[i] = b[i] - r[i];
        A->fn(A, r, Ar);
        ocx rAr = orc_dot(n, Ar, r);
        ocx ArAr = orc_dot(n, Ar, Ar);
        if (cabs(ArAr) < 1e-30) break;
        ocx omega = rAr / ArAr;
        axpy(n, omega, r, x);
    }
    free(r); free(Ar);
}
/* src/smoother.cpp:63-73 */
void orc_richardson(const orc_op* A, ocx* x, const ocx* b, int niter, double omega) {
    int n = A->n;
    ocx* r = xcalloc(n, sizeof(ocx));
    for (int it = 0; it < niter; it++) {
        A->fn(A, x, r);
        OMP_FOR
        for (int i = 0; i < n; i++) x[i] += omega * (b[i] - r[i]);
    }
    free(r);
}
/* src/smoother.cpp:43-59 with the start vector supplied explicitly */
double orc_lambda_max(const orc_op* A, const ocx* v0, int niter) {
    int n = A->n;
    ocx* v = xcalloc(n, sizeof(ocx));
    ocx* Av = xcalloc(n, sizeof(ocx));
    memcpy(v, v0, n * sizeof(ocx));
    double vn = orc_norm(n, v);
    vscale(n, v, 1.0 / vn);
    double lambda = 0;
    for (int i = 0; i < niter; i++) {
        A->fn(A, v, Av);
        lambda = orc_norm(n, Av);
        if (lambda < 1e-30) break;
        vscale(n, Av, 1.0 / lambda);
        memcpy(v, Av, n * sizeof(ocx));
    }
    free(v); free(Av);
    return lambda;
}

/* src/multigrid.cpp:545-587 (and :435-484, identical with A = D^dag D), all k
 * start vectors supplied explicitly (the reference's warm_start argument). */
void orc_null_space(const orc_op* A, int k, int outer_iters, const ocx* starts, ocx* out) {
    int n = A->n;
    ocx* xnew = xcalloc(n, sizeof(ocx));
    ocx* rhs = xcalloc(n, sizeof(ocx));
    for (int i = 0; i < k; i++) {
        ocx* v = out + (size_t)i * n;
        memcpy(v, starts + (size_t)i * n, n * sizeof(ocx));
        double nv = orc_norm(n, v);
        vscale(n, v, 1.0 / nv);
        for (int it = 0; it < outer_iters; it++) {
            memcpy(rhs, v, n * sizeof(ocx));
            memset(xnew, 0, n * sizeof(ocx));
            orc_mr_smooth(A, xnew, rhs, 5);
            for (int j = 0; j < i; j++) {
                ocx proj = orc_dot(n, out + (size_t)j * n, xnew);
                axpy(n, -proj, out + (size_t)j * n, xnew);
            }
            nv = orc_norm(n, xnew);
            if (nv > 1e-14) { vscale(n, xnew, 1.0 / nv); memcpy(v, xnew, n * sizeof(ocx)); }
            else break;
        }
    }
    for (int i = 0; i < k; i++) {
        ocx* v = out + (size_t)i * n;
        for (int j = 0; j < i; j++) {
            ocx proj = orc_dot(n, out + (size_t)j * n, v);
            axpy(n, -proj, out + (size_t)j * n, v);
        }
        double nv = orc_norm(n, v);
        if (nv > 1e-14) vscale(n, v, 1.0 / nv);
    }
    free(xnew); free(rhs);
}

/* ================================================================ multigrid */
#define MAXLEV 5
struct orc_mg {
    orc_mg_params prm;
    int nlev;
    const orc_op* op[MAXLEV];
    orc_op* own_op[MAXLEV];
    int dim[MAXLEV];
    orc_prolong* P[MAXLEV];
    orc_coarse* C[MAXLEV];
    ocx* nullv[MAXLEV];
    double omega[MAXLEV];
    orc_dense* dense;
};

/* src/multigrid.cpp:595-687, with geometric aggregation at every level and the
 * 9-point operator (src/coarse_op.h:154-297) as the coarse operator. */
orc_mg* orc_mg_new(const orc_op* fine, int nx, int ny, int nd, const orc_mg_params* prm, const ocx* starts) {
    orc_mg* mg = xcalloc(1, sizeof(orc_mg));
    mg->prm = *prm;
    mg->nlev = prm->n_levels;
    mg->op[0] = fine;
    mg->dim[0] = fine->n;
    int cnx = nx, cny = ny, cnd = nd;
    size_t soff = 0;
    for (int l = 0; l + 1 < mg->nlev; l++) {
        int k = prm->k[l], n = mg->dim[l], nv = prm->chiral ? k / 2 : k;
        mg->nullv[l] = xcalloc((size_t)nv * n, sizeof(ocx));
        orc_null_space(mg->op[l], nv, prm->null_iters, starts + soff, mg->nullv[l]);
        soff += (size_t)nv * n;
        mg->P[l] = orc_prolong_new(cnx, cny, cnd, prm->bx[l], prm->by[l], k, prm->chiral);
        orc_prolong_build(mg->P[l], mg->nullv[l]);
        mg->C[l] = orc_coarse_build(mg->op[l], mg->P[l], prm->colored_build);
        mg->own_op[l + 1] = orc_op_coarse(mg->C[l]);
        mg->op[l + 1] = mg->own_op[l + 1];
        mg->dim[l + 1] = mg->P[l]->nblocks * k;
        cnx = mg->P[l]->nbx; cny = mg->P[l]->nby; cnd = k;
    }
    if (prm->coarse_solver == 1 && mg->nlev >= 2)
        mg->dense = orc_dense_build(mg->op[mg->nlev - 1], NULL);
    return mg;
}
/* The hierarchy of orc_mg_new with its parts given instead of computed: per
 * transfer l the block vectors (orc_prolong_get layout) and the coarse links
 * (orc_coarse_get layout), each concatenated over levels.  The cycle is
 * src/multigrid.cpp:43-102 as in orc_mg_new; used to time the solve on a
 * hierarchy set up elsewhere. */
orc_mg* orc_mg_assemble(const orc_op* fine, int nx, int ny, int nd, const orc_mg_params* prm,
                        const ocx* block_vecs, const ocx* links) {
    orc_mg* mg = xcalloc(1, sizeof(orc_mg));
    mg->prm = *prm;
    mg->nlev = prm->n_levels;
    mg->op[0] = fine;
    mg->dim[0] = fine->n;
    int cnx = nx, cny = ny, cnd = nd;
    size_t boff = 0, loff = 0;
    for (int l = 0; l + 1 < mg->nlev; l++) {
        int k = prm->k[l];
        mg->P[l] = orc_prolong_new(cnx, cny, cnd, prm->bx[l], prm->by[l], k, prm->chiral);
        orc_prolong_set(mg->P[l], block_vecs + boff);
        boff += (size_t)mg->P[l]->nblocks * k * mg->P[l]->bsize;
        orc_coarse* C = xcalloc(1, sizeof(orc_coarse));
        C->nx = mg->P[l]->nbx; C->ny = mg->P[l]->nby; C->k = k; C->nb = mg->P[l]->nblocks;
        C->link = xcalloc((size_t)C->nb * 9 * k * k, sizeof(ocx));
        orc_coarse_set(C, links + loff);
        loff += (size_t)C->nb * 9 * k * k;
        mg->C[l] = C;
        mg->own_op[l + 1] = orc_op_coarse(C);
        mg->op[l + 1] = mg->own_op[l + 1];
        mg->dim[l + 1] = C->nb * k;
        cnx = C->nx; cny = C->ny; cnd = k;
    }
    if (prm->coarse_solver == 1 && mg->nlev >= 2)
        mg->dense = orc_dense_build(mg->op[mg->nlev - 1], NULL);
    return mg;
}

void orc_mg_free(orc_mg* mg) {
    if (!mg) return;
    for (int l = 0; l < MAXLEV; l++) {
        orc_prolong_free(mg->P[l]); orc_coarse_free(mg->C[l]);
        if (mg->own_op[l]) orc_op_free(mg->own_op[l]);
        free(mg->nullv[l]);
    }
    orc_dense_free(mg->dense);
    free(mg);
}
const orc_prolong* orc_mg_prolong(const orc_mg* mg, int l) { return mg->P[l]; }
const orc_coarse* orc_mg_coarse(const orc_mg* mg, int l) { return mg->C[l]; }
double orc_mg_omega(const orc_mg* mg, int l) { return mg->omega[l]; }
void orc_mg_null_vecs(const orc_mg* mg, int l, ocx* out) {
    int nv = mg->prm.chiral ? mg->prm.k[l] / 2 : mg->prm.k[l];
    if (mg->nullv[l]) memcpy(out, mg->nullv[l], (size_t)nv * mg->dim[l] * sizeof(ocx));
    else memset(out, 0, (size_t)nv * mg->dim[l] * sizeof(ocx));   /* assembled from parts */
}

/* src/multigrid.cpp:146-152 with explicit power-iteration starts per level */
void orc_mg_set_symmetric(orc_mg* mg, double damping, const ocx* starts) {
    size_t off = 0;
    for (int l = 0; l + 1 < mg->nlev; l++) {
        double lmax = orc_lambda_max(mg->op[l], starts + off, 30);
        mg->omega[l] = damping / lmax;
        off += mg->dim[l];
    }
}

/* src/coarse_op.h:492-547 without deflation vectors (x0 = 0) */
int orc_coarse_cg(const orc_op* A, const ocx* b, int maxit, double tol, ocx* x) {
    int n = A->n;
    memset(x, 0, n * sizeof(ocx));
    double bnorm = orc_norm(n, b);
    if (bnorm < 1e-30) return 0;
    ocx* Ax = xcalloc(n, sizeof(ocx));
    ocx* r = xcalloc(n, sizeof(ocx));
    ocx* p = xcalloc(n, sizeof(ocx));
    ocx* Ap = xcalloc(n, sizeof(ocx));
    A->fn(A, x, Ax);
    for (int i = 0; i < n; i++) r[i] = b[i] - Ax[i];
    memcpy(p, r, n * sizeof(ocx));
    ocx rr = orc_dot(n, r, r);
    int iter = 0;
    while (iter < maxit) {
        A->fn(A, p, Ap);
        ocx pAp = orc_dot(n, p, Ap);
        if (cabs(pAp) < 1e-30) break;
        ocx alpha = rr / pAp;
        for (int i = 0; i < n; i++) { x[i] += alpha * p[i]; r[i] -= alpha * Ap[i]; }
        iter++;
        double rn = orc_norm(n, r);
        if (rn / bnorm < tol) break;
        if (isnan(rn) || isinf(rn)) break;
        ocx rn2 = orc_dot(n, r, r);
        ocx beta = rn2 / rr;
        rr = rn2;
        for (int i = 0; i < n; i++) p[i] = r[i] + beta * p[i];
    }
    free(Ax); free(r); free(p); free(Ap);
    return iter;
}

/* src/multigrid.cpp:43-102 */
static void mg_cycle(orc_mg* mg, int lev, const ocx* b, ocx* x) {
    int n = mg->dim[lev], last = mg->nlev - 1;
    const orc_op* A = mg->op[lev];
    if (lev == last) {
        if (mg->dense) orc_dense_solve(mg->dense, b, x);
        else orc_coarse_cg(A, b, mg->prm.coarse_maxit, mg->prm.coarse_tol, x);
        return;
    }
    memset(x, 0, n * sizeof(ocx));
    if (mg->omega[lev] > 0) orc_richardson(A, x, b, mg->prm.pre_smooth, mg->omega[lev]);
    else orc_mr_smooth(A, x, b, mg->prm.pre_smooth);
    ocx* r = xcalloc(n, sizeof(ocx));
    A->fn(A, x, r);
    OMP_FOR
    for (int i = 0; i < n; i++) r[i] = b[i] - r[i];
    int nc = mg->dim[lev + 1];
    ocx* rc = xcalloc(nc, sizeof(ocx));
    ocx* ec = xcalloc(nc, sizeof(ocx));
    orc_restrict_apply(mg->P[lev], r, rc);
    mg_cycle(mg, lev + 1, rc, ec);
    if (mg->prm.w_cycle && lev + 1 < last) {
        ocx* t = xcalloc(nc, sizeof(ocx));
        ocx* ec2 = xcalloc(nc, sizeof(ocx));
        mg->op[lev + 1]->fn(mg->op[lev + 1], ec, t);
        for (int i = 0; i < nc; i++) t[i] = rc[i] - t[i];
        mg_cycle(mg, lev + 1, t, ec2);
        for (int i = 0; i < nc; i++) ec[i] += ec2[i];
        free(t); free(ec2);
    }
    ocx* e = r;   /* reuse */
    orc_prolong_apply(mg->P[lev], ec, e);
    OMP_FOR
    for (int i = 0; i < n; i++) x[i] += e[i];
    if (mg->omega[lev] > 0) orc_richardson(A, x, b, mg->prm.post_smooth, mg->omega[lev]);
    else orc_mr_smooth(A, x, b, mg->prm.post_smooth);
    free(r); free(rc); free(ec);
}
void orc_mg_precondition(orc_mg* mg, const ocx* b, ocx* x) { mg_cycle(mg, 0, b, x); }

static void precond(orc_mg* mg, int n, const ocx* r, ocx* z) {
    if (mg) orc_mg_precondition(mg, r, z);
    else memcpy(z, r, n * sizeof(ocx));
}

/* ================================================================ Krylov */
/* src/solvers.cpp:16-72 */
int orc_cg(const orc_op* A, const ocx* rhs, const ocx* x0, orc_mg* mg, int maxit, double tol,
           ocx* x, double* final_res) {
    int n = A->n;
    double rhs_norm = orc_norm(n, rhs);
    if (rhs_norm < 1e-30) { memset(x, 0, n * sizeof(ocx)); *final_res = 0; return 0; }
    ocx* r = xcalloc(n, sizeof(ocx));
    ocx* z = xcalloc(n, sizeof(ocx));
    ocx* p = xcalloc(n, sizeof(ocx));
    ocx* Ap = xcalloc(n, sizeof(ocx));
    if (x0) {
        memcpy(x, x0, n * sizeof(ocx));
        A->fn(A, x, Ap);
        OMP_FOR
        for (int i = 0; i < n; i++) r[i] = rhs[i] - Ap[i];
    } else {
        memset(x, 0, n * sizeof(ocx));
        memcpy(r, rhs, n * sizeof(ocx));
    }
    precond(mg, n, r, z);
    memcpy(p, z, n * sizeof(ocx));
    ocx rz = orc_dot(n, r, z);
    int iter = 0;
    while (iter < maxit) {
        A->fn(A, p, Ap);
        ocx pAp = orc_dot(n, p, Ap);
        ocx alpha = rz / pAp;
        OMP_FOR
        for (int i = 0; i < n; i++) { x[i] += alpha * p[i]; r[i] -= alpha * Ap[i]; }
        iter++;
        double rn = orc_norm(n, r);
        if (rn / rhs_norm < tol) break;
        precond(mg, n, r, z);
        ocx rz_new = orc_dot(n, r, z);
        ocx beta = rz_new / rz;
        rz = rz_new;
        OMP_FOR
        for (int i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
    }
    *final_res = orc_norm(n, r) / rhs_norm;
    free(r); free(z); free(p); free(Ap);
    return iter;
}

/* src/solvers.cpp:223-300 (truncated GCR, sliding window of trunc directions) */
int orc_fcg(const orc_op* A, const ocx* rhs, orc_mg* mg, int maxit, double tol, int trunc,
            int ref_projection, ocx* x, double* final_res) {
    int n = A->n;
    double rhs_norm = orc_norm(n, rhs);
    memset(x, 0, n * sizeof(ocx));
    if (rhs_norm < 1e-30) { *final_res = 0; return 0; }
    ocx* r = xcalloc(n, sizeof(ocx));
    memcpy(r, rhs, n * sizeof(ocx));
    int cap = trunc + 1;
    ocx** Pd = xcalloc(cap, sizeof(ocx*));
    ocx** APd = xcalloc(cap, sizeof(ocx*));
    double* ApAp = xcalloc(cap, sizeof(double));
    for (int i = 0; i < cap; i++) { Pd[i] = xcalloc(n, sizeof(ocx)); APd[i] = xcalloc(n, sizeof(ocx)); }
    ocx* z = xcalloc(n, sizeof(ocx));
    ocx* Az = xcalloc(n, sizeof(ocx));
    int nd = 0, iter = 0;
    while (iter < maxit) {
        double rn = orc_norm(n, r);
        if (rn / rhs_norm < tol) break;
        precond(mg, n, r, z);
        A->fn(A, z, Az);
        ocx* p = Pd[nd];
        ocx* Ap = APd[nd];
        memcpy(p, z, n * sizeof(ocx));
        memcpy(Ap, Az, n * sizeof(ocx));
        for (int i = 0; i < nd; i++) {
            /* src/solvers.cpp:261 computes dot(Az, AP_i) = <Az, AP_i>, the complex
             * conjugate of the GCR projection coefficient <AP_i, Az>; ref_projection
             * selects that form, otherwise the textbook coefficient is used. */
            ocx beta = (ref_projection ? orc_dot(n, Az, APd[i]) : orc_dot(n, APd[i], Az)) / ApAp[i];
            OMP_FOR
            for (int j = 0; j < n; j++) { p[j] -= beta * Pd[i][j]; Ap[j] -= beta * APd[i][j]; }
        }
        double apap = creal(orc_dot(n, Ap, Ap));
        if (apap < 1e-30) { nd = 0; continue; }
        ocx alpha = orc_dot(n, Ap, r) / apap;
        OMP_FOR
        for (int j = 0; j < n; j++) { x[j] += alpha * p[j]; r[j] -= alpha * Ap[j]; }
        iter++;
        ApAp[nd] = apap;
        nd++;
        if (nd > trunc) {   /* erase the oldest direction (the reference erases before push) */
            ocx* t0 = Pd[0]; ocx* t1 = APd[0];
            for (int i = 0; i < nd - 1; i++) { Pd[i] = Pd[i + 1]; APd[i] = APd[i + 1]; ApAp[i] = ApAp[i + 1]; }
            Pd[nd - 1] = t0; APd[nd - 1] = t1;
            nd--;
        }
    }
    *final_res = orc_norm(n, r) / rhs_norm;
    for (int i = 0; i < cap; i++) { free(Pd[i]); free(APd[i]); }
    free(Pd); free(APd); free(ApAp); free(z); free(Az); free(r);
    return iter;
}

/* fgmres_solve_generic, src/solvers.cpp:555-772: right-preconditioned FGMRES(restart)
 * with two Gram-Schmidt passes and complex Givens rotations; with n_ritz > 0, after each
 * restart cycle of at least n_ritz steps, the n_ritz lowest eigenpairs of G = H~^dag H~
 * (the unreduced Hessenberg) give Ritz pairs sum_j evecs[j][p] Z_j (normalised), which
 * deflate the later preconditioned vectors (Z_k += sum (v^dag V_k / l) v); beyond
 * 2 n_ritz pairs the n_ritz smallest are kept.  The reference sorts with std::sort; an
 * insertion sort gives the same order for distinct values.  Returns the iterations;
 * *n_ritz_out pairs in ritz_vals / ritz_vecs [.][n] (capacity 2 n_ritz + restart). */
int orc_fgmres_ritz(const orc_op* A, const ocx* rhs, orc_mg* mg, int restart, int maxit, double tol,
                    int n_ritz, ocx* x, double* final_res, int* n_ritz_out, double* ritz_vals, ocx* ritz_vecs) {
    int n = A->n;
    double rhs_norm = orc_norm(n, rhs);
    memset(x, 0, n * sizeof(ocx));
    if (rhs_norm < 1e-30) { *final_res = 0; return 0; }
    int total = 0;
    ocx* Ax = xcalloc(n, sizeof(ocx));
    ocx* w = xcalloc(n, sizeof(ocx));
    int m0 = restart;
    ocx** V = xcalloc(m0 + 1, sizeof(ocx*));
    ocx** Z = xcalloc(m0, sizeof(ocx*));
    for (int i = 0; i <= m0; i++) V[i] = xcalloc(n, sizeof(ocx));
    for (int i = 0; i < m0; i++) Z[i] = xcalloc(n, sizeof(ocx));
    ocx* H = xcalloc((size_t)(m0 + 1) * m0, sizeof(ocx));   /* H[i*m0 + j] */
    ocx* g = xcalloc(m0 + 1, sizeof(ocx));
    ocx* cs = xcalloc(m0, sizeof(ocx));
    ocx* sn = xcalloc(m0, sizeof(ocx));
    ocx* y = xcalloc(m0, sizeof(ocx));
    ocx* Hr = xcalloc((size_t)(m0 + 1) * m0, sizeof(ocx));   /* unreduced Hessenberg */
    int cap = 2 * (n_ritz > 0 ? n_ritz : 1) + m0, nr = 0;
    double* rv = xcalloc(cap, sizeof(double));               /* harvested Ritz pairs */
    ocx** rvec = xcalloc(cap, sizeof(ocx*));
    if (n_ritz_out) *n_ritz_out = 0;
    while (total < maxit) {
        A->fn(A, x, Ax);
        for (int i = 0; i < n; i++) V[0][i] = rhs[i] - Ax[i];
        double beta = orc_norm(n, V[0]);
        if (beta / rhs_norm < tol) break;
        int m = restart < maxit - total ? restart : maxit - total;
        vscale(n, V[0], 1.0 / beta);
        memset(H, 0, (size_t)(m0 + 1) * m0 * sizeof(ocx));
        memset(Hr, 0, (size_t)(m0 + 1) * m0 * sizeof(ocx));
        memset(g, 0, (m0 + 1) * sizeof(ocx));
        g[0] = beta;
        int k;
        for (k = 0; k < m; k++) {
            precond(mg, n, V[k], Z[k]);
            for (int q = 0; q < nr; q++)   /* intra-solve deflation by earlier Ritz pairs */
                if (rv[q] > 1e-14) axpy(n, orc_dot(n, rvec[q], V[k]) / rv[q], rvec[q], Z[k]);
            A->fn(A, Z[k], w);
            for (int j = 0; j <= k; j++) {
                H[j * m0 + k] = orc_dot(n, V[j], w);
                axpy(n, -H[j * m0 + k], V[j], w);
            }
            for (int j = 0; j <= k; j++) {
                ocx c = orc_dot(n, V[j], w);
                H[j * m0 + k] += c;
                axpy(n, -c, V[j], w);
            }
            H[(k + 1) * m0 + k] = orc_norm(n, w);
            for (int j = 0; j <= k + 1; j++) Hr[j * m0 + k] = H[j * m0 + k];
            if (cabs(H[(k + 1) * m0 + k]) < 1e-14) { k++; break; }
            memcpy(V[k + 1], w, n * sizeof(ocx));
            vscale(n, V[k + 1], 1.0 / H[(k + 1) * m0 + k]);
            for (int j = 0; j < k; j++) {
                ocx hj = H[j * m0 + k], hj1 = H[(j + 1) * m0 + k];
                H[j * m0 + k] = conj(cs[j]) * hj + conj(sn[j]) * hj1;
                H[(j + 1) * m0 + k] = -sn[j] * hj + cs[j] * hj1;
            }
            double a = cabs(H[k * m0 + k]), bb = cabs(H[(k + 1) * m0 + k]);
            double den = sqrt(a * a + bb * bb);
            cs[k] = H[k * m0 + k] / den;
            sn[k] = H[(k + 1) * m0 + k] / den;
            g[k + 1] = -sn[k] * g[k];
            g[k] = conj(cs[k]) * g[k];
            H[k * m0 + k] = den;
            H[(k + 1) * m0 + k] = 0.0;
            total++;
            if (cabs(g[k + 1]) / rhs_norm < tol) { k++; break; }
        }
        int kk = k;
        for (int i = kk - 1; i >= 0; i--) {
            ocx s = g[i];
            for (int j = i + 1; j < kk; j++) s -= H[i * m0 + j] * y[j];
            y[i] = s / H[i * m0 + i];
        }
        for (int j = 0; j < kk; j++) axpy(n, y[j], Z[j], x);
        if (kk >= n_ritz && n_ritz > 0) {   /* Ritz pairs of G = H~^dag H~ (:707-751) */
            ocx* G = xcalloc((size_t)kk * kk, sizeof(ocx));   /* G[j*kk + i] */
            for (int j = 0; j < kk; j++)
                for (int i = 0; i < kk; i++)
                    for (int p = 0; p <= kk; p++) G[(size_t)j * kk + i] += conj(Hr[p * m0 + i]) * Hr[p * m0 + j];
            for (int i = 0; i < kk; i++)
                for (int j = i + 1; j < kk; j++) {
                    ocx av = 0.5 * (G[(size_t)j * kk + i] + conj(G[(size_t)i * kk + j]));
                    G[(size_t)j * kk + i] = av;
                    G[(size_t)i * kk + j] = conj(av);
                }
            double* ev = xcalloc(kk, sizeof(double));
            ocx* Ev = xcalloc((size_t)kk * kk, sizeof(ocx));   /* Ev[row*kk + col] */
            orc_lanczos_eigen(kk, G, ev, Ev);
            int nh = n_ritz < kk ? n_ritz : kk;
            for (int p = 0; p < nh; p++) {
                ocx* v = xcalloc(n, sizeof(ocx));
                for (int j = 0; j < kk; j++) axpy(n, Ev[(size_t)j * kk + p], Z[j], v);
                double nv = orc_norm(n, v);
                if (nv > 1e-14) vscale(n, v, 1.0 / nv);
                rv[nr] = ev[p];
                rvec[nr] = v;
                nr++;
            }
            if (nr > 2 * n_ritz) {   /* keep the n_ritz smallest (insertion sort by value) */
                for (int a = 1; a < nr; a++)
                    for (int b2 = a; b2 > 0 && rv[b2 - 1] > rv[b2]; b2--) {
                        double tv = rv[b2]; rv[b2] = rv[b2 - 1]; rv[b2 - 1] = tv;
                        ocx* tp = rvec[b2]; rvec[b2] = rvec[b2 - 1]; rvec[b2 - 1] = tp;
                    }
                for (int q = n_ritz; q < nr; q++) free(rvec[q]);
                nr = n_ritz;
            }
            free(G); free(ev); free(Ev);
        }
        if (cabs(g[kk]) / rhs_norm < tol) break;
    }
    A->fn(A, x, Ax);
    double res = 0;
    for (int i = 0; i < n; i++) { ocx d = rhs[i] - Ax[i]; res += creal(d) * creal(d) + cimag(d) * cimag(d); }
    *final_res = sqrt(res) / rhs_norm;
    if (n_ritz_out) *n_ritz_out = nr;
    for (int q = 0; q < nr; q++) {
        if (ritz_vals) ritz_vals[q] = rv[q];
        if (ritz_vecs) memcpy(ritz_vecs + (size_t)q * n, rvec[q], n * sizeof(ocx));
        free(rvec[q]);
    }
    free(rv); free(rvec); free(Hr);
    for (int i = 0; i <= m0; i++) free(V[i]);
    for (int i = 0; i < m0; i++) free(Z[i]);
    free(V); free(Z); free(H); free(g); free(cs); free(sn); free(y); free(Ax); free(w);
    return total;
}

/* src/solvers.cpp:555-772 with n_ritz_harvest = 0 (no Ritz harvest). */
int orc_fgmres(const orc_op* A, const ocx* rhs, orc_mg* mg, int restart, int maxit, double tol,
               ocx* x, double* final_res) {
    return orc_fgmres_ritz(A, rhs, mg, restart, maxit, tol, 0, x, final_res, NULL, NULL, NULL);
}

/* ================================================================ forces */
/* src/hmc_forces.cpp:540-569 */
void orc_hopping_force(const orc_dirac* D, const ocx* chi, const ocx* X, double* F) {
    int L = D->L, V = D->V, n = V;
    const double r = D->r;
    const ocx* U0 = D->U; const ocx* U1 = D->U + V;
    const ocx ii = I;
    OMP_FOR
    for (int s = 0; s < V; s++) {
        ocx c0 = chi[2 * s], c1 = chi[2 * s + 1], x0 = X[2 * s], x1 = X[2 * s + 1];
        { int sf = xp(L, s);
          ocx u = U0[s], ud = conj(u);
          ocx ux0 = u * X[2 * sf], ux1 = u * X[2 * sf + 1];
          ocx f0 = r * ux0 - ux1, f1 = -ux0 + r * ux1;
          ocx c_fwd = conj(c0) * f0 + conj(c1) * f1;
          ocx udx0 = ud * x0, udx1 = ud * x1;
          ocx b0 = r * udx0 + udx1, b1 = udx0 + r * udx1;
          ocx c_bwd = conj(chi[2 * sf]) * b0 + conj(chi[2 * sf + 1]) * b1;
          F[s] = creal(ii * c_bwd - ii * c_fwd); }
        { int sf = yp(L, s);
          ocx u = U1[s], ud = conj(u);
          ocx ux0 = u * X[2 * sf], ux1 = u * X[2 * sf + 1];
          ocx f0 = r * ux0 + ii * ux1, f1 = -ii * ux0 + r * ux1;
          ocx c_fwd = conj(c0) * f0 + conj(c1) * f1;
          ocx udx0 = ud * x0, udx1 = ud * x1;
          ocx b0 = r * udx0 - ii * udx1, b1 = ii * udx0 + r * udx1;
          ocx c_bwd = conj(chi[2 * sf]) * b0 + conj(chi[2 * sf + 1]) * b1;
          F[V + s] = creal(ii * c_bwd - ii * c_fwd); }
    }
}

/* src/hmc_forces.cpp:501-534 */
void orc_clover_deriv_insert(const orc_dirac* D, const double* w, double factor, double* F) {
    int L = D->L, V = D->V, n = V;
    const ocx* U0 = D->U; const ocx* U1 = D->U + V;
    OMP_FOR
    for (int s = 0; s < V; s++) {
        int sxp = xp(L, s), sxm = xm(L, s), syp = yp(L, s), sym = ym(L, s);
        int sxpym = ym(L, sxp), sxmyp = yp(L, sxm), sxpyp = yp(L, sxp);
        ocx P_fwd = U0[s] * U1[sxp] * conj(U0[syp]) * conj(U1[s]);
        double re_fwd = creal(P_fwd);
        double w_fwd = w[s] + w[sxp] + w[sxpyp] + w[syp];
        ocx P_bwd = U0[sym] * U1[sxpym] * conj(U0[s]) * conj(U1[sym]);
        double re_bwd = creal(P_bwd);
        double w_bwd = w[sym] + w[sxpym] + w[sxp] + w[s];
        F[s] += factor * (re_fwd * w_fwd - re_bwd * w_bwd);
        ocx P_fwd1 = U0[sxm] * U1[s] * conj(U0[sxmyp]) * conj(U1[sxm]);
        double re_fwd1 = creal(P_fwd1);
        double w_fwd1 = w[sxm] + w[s] + w[syp] + w[sxmyp];
        double re_bwd1 = re_fwd;
        double w_bwd1 = w[s] + w[sxp] + w[sxpyp] + w[syp];
        F[V + s] += factor * (re_fwd1 * w_fwd1 - re_bwd1 * w_bwd1);
    }
}

/* src/hmc_forces.cpp:141-194 + 198-495 (the clover part is the insertion with
 * w = Re(chi0* X0 - chi1* X1) and factor c_sw/4) */
void orc_fermion_force_bilinear(const orc_dirac* D, const ocx* chi, const ocx* X, double* F) {
    orc_hopping_force(D, chi, X, F);
    if (D->csw != 0.0) {
        int V = D->V;
        double* w = xcalloc(V, sizeof(double));
        for (int s = 0; s < V; s++)
            w[s] = creal(conj(chi[2 * s]) * X[2 * s] - conj(chi[2 * s + 1]) * X[2 * s + 1]);
        orc_clover_deriv_insert(D, w, D->csw / 4.0, F);
        free(w);
    }
}

/* src/hmc_forces.cpp:134-139 */
void orc_fermion_force(const orc_dirac* D, const ocx* X, double* F) {
    ocx* chi = xcalloc(2 * (size_t)D->V, sizeof(ocx));
    orc_dirac_apply(D, ORC_D, X, chi);
    orc_fermion_force_bilinear(D, chi, X, F);
    free(chi);
}

/* src/hmc_forces.cpp:578-637 */
void orc_logdet_ee_force(const orc_dirac* D, double* F) {
    int V = D->V;
    memset(F, 0, 2 * (size_t)V * sizeof(double));
    if (D->csw == 0.0) return;
    double csw = D->csw, diag = 2.0 * D->r + D->mass, mu2 = D->mu_t * D->mu_t;
    double* w = xcalloc(V, sizeof(double));
    for (int ie = 0; ie < V / 2; ie++) {
        int e = D->even[ie];
        double C = 0.5 * csw * D->F01[e];
        double d0 = diag + C, d1 = diag - C;
        w[e] = 0.5 * csw * (d0 / (d0 * d0 + mu2) - d1 / (d1 * d1 + mu2));
    }
    orc_clover_deriv_insert(D, w, 0.5, F);
    free(w);
}

/* src/hmc_forces.cpp:763-828 */
static void schur_deriv_plus(const orc_dirac* D, const ocx* chi_o, const ocx* psi_o, double* F) {
    int V = D->V, nh = V / 2;
    ocx* t1 = xcalloc(2 * (size_t)nh, sizeof(ocx));
    ocx* tmp2 = xcalloc(2 * (size_t)nh, sizeof(ocx));
    eo_hop(D, psi_o, t1, D->even);
    diag_inv(D, t1, tmp2, D->even, D->mu_t);
    ocx* g5c = xcalloc(2 * (size_t)nh, sizeof(ocx));
    for (int i = 0; i < nh; i++) { g5c[2 * i] = chi_o[2 * i]; g5c[2 * i + 1] = -chi_o[2 * i + 1]; }
    ocx* hop = xcalloc(2 * (size_t)nh, sizeof(ocx));
    eo_hop(D, g5c, hop, D->even);
    for (int i = 0; i < nh; i++) hop[2 * i + 1] = -hop[2 * i + 1];
    ocx* tmp3 = xcalloc(2 * (size_t)nh, sizeof(ocx));
    diag_inv(D, hop, tmp3, D->even, -D->mu_t);
    if (D->csw != 0.0) {
        double* w = xcalloc(V, sizeof(double));
        for (int io = 0; io < nh; io++)
            w[D->odd[io]] = creal(conj(chi_o[2 * io]) * psi_o[2 * io] - conj(chi_o[2 * io + 1]) * psi_o[2 * io + 1]);
        for (int ie = 0; ie < nh; ie++)
            w[D->even[ie]] = creal(conj(tmp3[2 * ie]) * tmp2[2 * ie] - conj(tmp3[2 * ie + 1]) * tmp2[2 * ie + 1]);
        orc_clover_deriv_insert(D, w, D->csw / 8.0, F);
        free(w);
    }
    ocx* chif = xcalloc(2 * (size_t)V, sizeof(ocx));
    ocx* psif = xcalloc(2 * (size_t)V, sizeof(ocx));
    eo_scatter(D, tmp3, chi_o, chif);
    eo_scatter(D, tmp2, psi_o, psif);
    double* ff = xcalloc(2 * (size_t)V, sizeof(double));
    orc_hopping_force(D, chif, psif, ff);
    for (int i = 0; i < 2 * V; i++) F[i] -= 0.5 * ff[i];
    free(t1); free(tmp2); free(g5c); free(hop); free(tmp3); free(chif); free(psif); free(ff);
}

/* src/hmc_forces.cpp:830-867 */
void orc_eo_fermion_force(const orc_dirac* D, const ocx* x_o, const ocx* y_o, double* F) {
    int V = D->V;
    memset(F, 0, 2 * (size_t)V * sizeof(double));
    schur_deriv_plus(D, y_o, x_o, F);
    for (int i = 0; i < 2 * V; i++) F[i] *= 2.0;
    if (D->csw != 0.0) {
        double* fl = xcalloc(2 * (size_t)V, sizeof(double));
        orc_logdet_ee_force(D, fl);
        for (int i = 0; i < 2 * V; i++) F[i] += fl[i];
        free(fl);
    }
}

/* ================================================================ U(1) heatbath
 * No reference counterpart (the reference thermalises by HMC).  Checkerboard heatbath of
 * the Wilson action S = -beta sum_x Re P(x): links U_mu(x) of one direction and one
 * parity are independent given the rest; each is drawn from exp(beta Re(U A)), A the sum
 * of its two staples, as U = e^{i phi} conj(A)/|A| with phi from the von Mises density
 * exp(kappa cos phi), kappa = beta |A| (Best & Fisher 1979, Appl. Statist. 28:152-157).
 * Random numbers: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC11), key = seed, counter
 * = (link index mu*V + x, 4*sweep + 2*mu + parity, attempt, 0). */
static inline uint32_t mulhilo32(uint32_t a, uint32_t b, uint32_t* hi) {
    uint64_t p = (uint64_t)a * b;
    *hi = (uint32_t)(p >> 32);
    return (uint32_t)p;
}
void orc_philox4x32(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; r++) {
        uint32_t hi0, hi1;
        uint32_t lo0 = mulhilo32(0xD2511F53u, c0, &hi0);
        uint32_t lo1 = mulhilo32(0xCD9E8D57u, c2, &hi1);
        uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}
#define ORC_PI 3.14159265358979323846
static inline double u32_to_unit(uint32_t x) { return ((double)x + 0.5) * (1.0 / 4294967296.0); }
/* one von Mises(kappa) draw from the Philox stream (ctr[2] counts attempts); a stream
 * that rejects 1000 times returns nan (the caller keeps the old link) */
double orc_von_mises(double kappa, uint32_t ctr0, uint32_t ctr1, const uint32_t key[2]) {
    uint32_t ctr[4] = {ctr0, ctr1, 0u, 0u}, o[4];
    if (kappa < 1e-8) {
        orc_philox4x32(ctr, key, o);
        return ORC_PI * (2.0 * u32_to_unit(o[0]) - 1.0);
    }
    double tau = 1.0 + sqrt(1.0 + 4.0 * kappa * kappa);
    double rho = (tau - sqrt(2.0 * tau)) / (2.0 * kappa);
    double rr = (1.0 + rho * rho) / (2.0 * rho);
    for (uint32_t a = 0; a < 1000; a++) {
        ctr[2] = a;
        orc_philox4x32(ctr, key, o);
        double u1 = u32_to_unit(o[0]), u2 = u32_to_unit(o[1]), u3 = u32_to_unit(o[2]);
        double z = cos(ORC_PI * u1);
        double f = (1.0 + rr * z) / (rr + z);
        double c = kappa * (rr - f);
        if (c * (2.0 - c) - u2 > 0.0 || log(c / u2) + 1.0 - c >= 0.0) {
            double ph = acos(fmax(-1.0, fmin(1.0, f)));
            return u3 > 0.5 ? ph : -ph;
        }
    }
    return NAN;
}
/* the staple sum A of U_mu(x): Re(U_mu(x) A) is the x-dependent part of sum Re P */
static ocx staple(int L, const ocx* U, int mu, int s) {
    int V = L * L, x = s % L, y = s / L;
    int xp = (x + 1) % L + L * y, xm = (x - 1 + L) % L + L * y;
    int yp = x + L * ((y + 1) % L), ym = x + L * ((y - 1 + L) % L);
    const ocx* U0 = U;
    const ocx* U1 = U + V;
    if (mu == 0) {
        int xp_ym = (x + 1) % L + L * ((y - 1 + L) % L);
        return U1[xp] * conj(U0[yp]) * conj(U1[s]) + conj(U0[ym]) * conj(U1[xp_ym]) * U1[ym];
    }
    int xm_yp = (x - 1 + L) % L + L * ((y + 1) % L);
    return conj(U0[s]) * conj(U1[xp]) * U0[yp] + U0[xm] * conj(U0[xm_yp]) * conj(U1[xm]);
}
/* n_sweeps sweeps (each: mu = 0, 1; parity 0, 1), the first numbered first_sweep */
void orc_heatbath(int L, ocx* U, double beta, int n_sweeps, uint64_t seed, uint32_t first_sweep) {
    int V = L * L;
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (int sw = 0; sw < n_sweeps; sw++)
        for (int mu = 0; mu < 2; mu++)
            for (int par = 0; par < 2; par++)
                for (int s = 0; s < V; s++) {
                    int x = s % L, y = s / L;
                    if (((x + y) & 1) != par) continue;
                    ocx A = staple(L, U, mu, s);
                    double a = cabs(A);
                    uint32_t sub = 4u * (first_sweep + (uint32_t)sw) + 2u * (uint32_t)mu + (uint32_t)par;
                    double ph = orc_von_mises(beta * a, (uint32_t)(mu * V + s), sub, key);
                    if (isnan(ph)) continue;
                    ocx dir = a > 0.0 ? conj(A) / a : 1.0;
                    U[(size_t)mu * V + s] = cexp(I * ph) * dir;
                }
}

/* ================================================================ HMC molecular dynamics */
/* MomentumField::kinetic_energy, src/hmc_forces.cpp:70-77: 0.5 sum pi^2 over 2V reals */
double orc_kinetic_energy(int n, const double* pi) {
    double ke = 0.0;
    for (int i = 0; i < n; i++) ke += pi[i] * pi[i];
    return 0.5 * ke;
}

/* the MD force of src/hmc_trajectory.cpp:254-343 without tracking: gauge force plus
 * the fermion force of X = (D^dag D)^-1 phi (CG from x0 = 0), both -dS/dtheta */
static int md_force(int L, const ocx* U, double mass, double r, double csw, double mu_t, double beta, int cg_maxit,
                    double cg_tol, const ocx* phi, double* F) {
    int V = L * L;
    orc_gauge_force(L, U, beta, F);
    orc_dirac* D = orc_dirac_new(L, U, mass, r, csw, mu_t);
    orc_op* A = orc_op_dirac(D, ORC_DDAGD);
    ocx* X = xcalloc(2 * (size_t)V, sizeof(ocx));
    double res;
    int it = orc_cg(A, phi, NULL, NULL, cg_maxit, cg_tol, X, &res);
    double* ff = xcalloc(2 * (size_t)V, sizeof(double));
    orc_fermion_force(D, X, ff);
    for (int i = 0; i < 2 * V; i++) F[i] += ff[i];
    free(ff); free(X);
    orc_op_free(A);
    orc_dirac_free(D);
    return it;
}
/* src/hmc_trajectory.cpp:345-357 */
static void md_update_mom(int n, double* pi, const double* F, double scale) {
    for (int i = 0; i < n; i++) pi[i] += scale * F[i];
}
static void md_update_gauge(int n, ocx* U, const double* pi, double scale) {
    for (int i = 0; i < n; i++) U[i] *= cexp(I * (scale * pi[i]));
}

/* plain_leapfrog_evolve, src/hmc_trajectory.cpp:243-409: U (2V complex) and pi (2V
 * reals) evolved in place over tau in n_steps steps, leapfrog P(dt/2) [Q(dt) P(dt)]^(N-1)
 * Q(dt) P(dt/2) or Omelyan (2MN, lambda = 0.1932); returns the CG iterations. */
int orc_hmc_evolve(int L, ocx* U, double mass, double r, double csw, double mu_t, double beta, double tau,
                   int n_steps, int omelyan, int cg_maxit, double cg_tol, double* pi, const ocx* phi) {
    int n = 2 * L * L, cg = 0;
    double dt = tau / n_steps;
    double* F = xcalloc(n, sizeof(double));
    if (omelyan) {
        const double lambda = 0.1932;
        cg += md_force(L, U, mass, r, csw, mu_t, beta, cg_maxit, cg_tol, phi, F);
        md_update_mom(n, pi, F, lambda * dt);
        for (int step = 0; step < n_steps; step++) {
            md_update_gauge(n, U, pi, 0.5 * dt);
            cg += md_force(L, U, mass, r, csw, mu_t, beta, cg_maxit, cg_tol, phi, F);
            md_update_mom(n, pi, F, (1.0 - 2.0 * lambda) * dt);
            md_update_gauge(n, U, pi, 0.5 * dt);
            cg += md_force(L, U, mass, r, csw, mu_t, beta, cg_maxit, cg_tol, phi, F);
            md_update_mom(n, pi, F, (step < n_steps - 1 ? 2.0 : 1.0) * lambda * dt);
        }
    } else {
        cg += md_force(L, U, mass, r, csw, mu_t, beta, cg_maxit, cg_tol, phi, F);
        md_update_mom(n, pi, F, 0.5 * dt);
        for (int step = 0; step < n_steps - 1; step++) {
            md_update_gauge(n, U, pi, dt);
            cg += md_force(L, U, mass, r, csw, mu_t, beta, cg_maxit, cg_tol, phi, F);
            md_update_mom(n, pi, F, dt);
        }
        md_update_gauge(n, U, pi, dt);
        cg += md_force(L, U, mass, r, csw, mu_t, beta, cg_maxit, cg_tol, phi, F);
        md_update_mom(n, pi, F, 0.5 * dt);
    }
    free(F);
    return cg;
}

/* The nested (Sexton-Weingarten) integrator with the gauge force on the fine scale; no
 * reference counterpart (the reference's multi-timescale split, src/hmc_multiscale.cpp:
 * 424-506, puts the gauge force on the outer scale).  The outer integrator -- leapfrog
 * P(h/2) [Q(h) P(h)]^(N-1) Q(h) P(h/2) or Omelyan 2MN, lambda = 0.1932, the step
 * sequences of src/hmc_trajectory.cpp:243-409 -- kicks with the fermion force alone;
 * each of its position updates Q(h) is n_gauge steps of the inner integrator (leapfrog
 * or Omelyan) of step h / n_gauge kicking with the gauge force alone.  Parity is pinned
 * by reductions: beta = 0 makes every inner integrator one drift, which is the plain
 * trajectory of the fermion force; phi = 0 makes the fermion force vanish, which is
 * the plain trajectory of the gauge force with the inner steps; and by reversibility. */
static void nested_inner(int L, ocx* U, double beta, double h, int n_gauge, int omelyan, double* pi, double* F) {
    int n = 2 * L * L;
    double dt = h / n_gauge;
    if (omelyan) {
        const double lambda = 0.1932;
        orc_gauge_force(L, U, beta, F);
        md_update_mom(n, pi, F, lambda * dt);
        for (int s = 0; s < n_gauge; s++) {
            md_update_gauge(n, U, pi, 0.5 * dt);
            orc_gauge_force(L, U, beta, F);
            md_update_mom(n, pi, F, (1.0 - 2.0 * lambda) * dt);
            md_update_gauge(n, U, pi, 0.5 * dt);
            orc_gauge_force(L, U, beta, F);
            md_update_mom(n, pi, F, (s < n_gauge - 1 ? 2.0 : 1.0) * lambda * dt);
        }
    } else {
        orc_gauge_force(L, U, beta, F);
        md_update_mom(n, pi, F, 0.5 * dt);
        for (int s = 0; s < n_gauge - 1; s++) {
            md_update_gauge(n, U, pi, dt);
            orc_gauge_force(L, U, beta, F);
            md_update_mom(n, pi, F, dt);
        }
        md_update_gauge(n, U, pi, dt);
        orc_gauge_force(L, U, beta, F);
        md_update_mom(n, pi, F, 0.5 * dt);
    }
}
int orc_hmc_nested_evolve(int L, ocx* U, double mass, double r, double csw, double mu_t, double beta, double tau,
                          int n_outer, int n_gauge, int omelyan_outer, int omelyan_inner, int cg_maxit,
                          double cg_tol, double* pi, const ocx* phi) {
    int n = 2 * L * L, cg = 0;
    double h = tau / n_outer;
    double* F = xcalloc(n, sizeof(double));
    /* md_force at beta = 0 is the fermion force alone */
#define FFORCE() (cg += md_force(L, U, mass, r, csw, mu_t, 0.0, cg_maxit, cg_tol, phi, F))
    if (omelyan_outer) {
        const double lambda = 0.1932;
        FFORCE();
        md_update_mom(n, pi, F, lambda * h);
        for (int o = 0; o < n_outer; o++) {
            nested_inner(L, U, beta, 0.5 * h, n_gauge, omelyan_inner, pi, F);
            FFORCE();
            md_update_mom(n, pi, F, (1.0 - 2.0 * lambda) * h);
            nested_inner(L, U, beta, 0.5 * h, n_gauge, omelyan_inner, pi, F);
            FFORCE();
            md_update_mom(n, pi, F, (o < n_outer - 1 ? 2.0 : 1.0) * lambda * h);
        }
    } else {
        FFORCE();
        md_update_mom(n, pi, F, 0.5 * h);
        for (int o = 0; o < n_outer; o++) {
            nested_inner(L, U, beta, h, n_gauge, omelyan_inner, pi, F);
            FFORCE();
            md_update_mom(n, pi, F, o < n_outer - 1 ? h : 0.5 * h);
        }
    }
#undef FFORCE
    free(F);
    return cg;
}

/* rr_evolve, src/eigen_evolve.cpp:48-108: Rayleigh-Ritz of the k vectors V [k][n]
 * (in/out) on A: M_ij = <v_i, A v_j> (i <= j, M_ji = conj), diagonalised by
 * orc_lanczos_eigen; v_i <- sum_j evecs[j][i] v_j, vals <- the ascending eigenvalues */
void orc_rr_evolve(const orc_op* A, int k, ocx* V, double* vals) {
    int n = A->n;
    ocx* AV = xcalloc((size_t)k * n, sizeof(ocx));
    for (int i = 0; i < k; i++) A->fn(A, V + (size_t)i * n, AV + (size_t)i * n);
    ocx* M = xcalloc((size_t)k * k, sizeof(ocx));   /* M[j*k + i] = M_ij (columns) */
    for (int j = 0; j < k; j++)
        for (int i = 0; i <= j; i++) {
            ocx val = orc_dot(n, V + (size_t)i * n, AV + (size_t)j * n);
            M[(size_t)j * k + i] = val;
            if (i != j) M[(size_t)i * k + j] = conj(val);
        }
    double* ev = xcalloc(k, sizeof(double));
    ocx* Z = xcalloc((size_t)k * k, sizeof(ocx));   /* Z[row*k + col] */
    orc_lanczos_eigen(k, M, ev, Z);
    ocx* W = xcalloc((size_t)k * n, sizeof(ocx));
    for (int i = 0; i < k; i++) {
        for (int j = 0; j < k; j++) {
            ocx c = Z[(size_t)j * k + i];
            for (int t = 0; t < n; t++) W[(size_t)i * n + t] += c * V[(size_t)j * n + t];
        }
        vals[i] = ev[i];
    }
    memcpy(V, W, (size_t)k * n * sizeof(ocx));
    free(AV); free(M); free(ev); free(Z); free(W);
}

/* CoarseDeflState::lowmode_solution, src/hmc.h:311-323: P (sum_i v_i v_i^dag / l_i) P^dag phi
 * over the l_i > 1e-14 */
static void lowmode_solution(const orc_prolong* P, int k, const ocx* V, const double* vals, const ocx* phi,
                             ocx* out) {
    int nc = orc_prolong_coarse_dim(P);
    ocx* pc = xcalloc(nc, sizeof(ocx));
    ocx* xc = xcalloc(nc, sizeof(ocx));
    orc_restrict_apply(P, phi, pc);
    for (int i = 0; i < k; i++)
        if (vals[i] > 1e-14) {
            ocx c = orc_dot(nc, V + (size_t)i * nc, pc) / vals[i];
            for (int t = 0; t < nc; t++) xc[t] += c * V[(size_t)i * nc + t];
        }
    orc_prolong_apply(P, xc, out);
    free(pc); free(xc);
}

/* coarse_lowmode_force, src/hmc_multiscale.cpp:237-276: smooth = 0 takes the force of
 * the low-mode solution; otherwise x = MR(smooth) on D^dag D x = phi, plus the low-mode
 * solution of its residual, then MR(smooth) again from x; F = fermion_force(D, x) */
void orc_coarse_lowmode_force(const orc_dirac* D, const orc_prolong* P, int k, const ocx* V, const double* vals,
                              const ocx* phi, int smooth, double* F) {
    orc_op* A = orc_op_dirac(D, ORC_DDAGD);
    int n = A->n;
    ocx* x = xcalloc(n, sizeof(ocx));
    if (smooth <= 0) {
        lowmode_solution(P, k, V, vals, phi, x);
    } else {
        ocx* Ax = xcalloc(n, sizeof(ocx));
        ocx* e = xcalloc(n, sizeof(ocx));
        orc_mr_smooth(A, x, phi, smooth);
        A->fn(A, x, Ax);
        for (int i = 0; i < n; i++) Ax[i] = phi[i] - Ax[i];
        lowmode_solution(P, k, V, vals, Ax, e);
        for (int i = 0; i < n; i++) x[i] += e[i];
        orc_mr_smooth(A, x, phi, smooth);
        free(Ax); free(e);
    }
    orc_fermion_force(D, x, F);
    free(x);
    orc_op_free(A);
}

/* The molecular dynamics of hmc_trajectory_mg_multiscale, src/hmc_multiscale.cpp:340-888,
 * without eigenspace tracking, forecasting or prolongator perturbation: the outer
 * (leapfrog or Omelyan, lambda = 0.1932) integrator kicks with F_g + F_f - F_low (F_f from
 * the CG solve of D^dag D X = phi; the oracle's CG is unpreconditioned), the inner leapfrog
 * (n_inner steps per call) with F_low = coarse_lowmode_force; every defl_refresh-th inner
 * step the deflation space V [k][nc] (in/out) is Rayleigh-Ritz evolved on the Galerkin
 * operator of P on the current gauge field (evolve_coarse_deflation, :278-338).  U and pi
 * evolve in place; returns the outer CG iterations, *lowmode_evals the F_low count. */
int orc_hmc_ms_evolve(int L, ocx* U, double mass, double r, double csw, double mu_t, double beta, double tau,
                      int n_outer, int n_inner, int omelyan, int inner_smooth, int defl_refresh, int cg_maxit,
                      double cg_tol, const orc_prolong* P, int k, ocx* V, double* vals, double* pi, const ocx* phi,
                      int* lowmode_evals) {
    int n = 2 * L * L, cg = 0, nlow = 0, counter = 0;
    double h = tau / n_outer;
    double* F = xcalloc(n, sizeof(double));
    double* fl = xcalloc(n, sizeof(double));
#define LOW_FORCE()                                                                        \
    do {                                                                                   \
        orc_dirac* Dc = orc_dirac_new(L, U, mass, r, csw, mu_t);                           \
        orc_coarse_lowmode_force(Dc, P, k, V, vals, phi, inner_smooth, fl);                \
        orc_dirac_free(Dc);                                                                \
    } while (0)
#define INNER_FORCE()  /* compute_inner_force, :593-599 (the one that counts) */            \
    do {                                                                                   \
        LOW_FORCE();                                                                       \
        nlow++;                                                                            \
    } while (0)
#define OUTER_FORCE()  /* compute_outer_force, :446-512 */                                  \
    do {                                                                                   \
        cg += md_force(L, U, mass, r, csw, mu_t, beta, cg_maxit, cg_tol, phi, F);          \
        LOW_FORCE();                                                                       \
        for (int i = 0; i < n; i++) F[i] -= fl[i];                                         \
    } while (0)
    /* inner_integrator(dt_total), :667-685 */
#define INNER(DT)                                                                          \
    do {                                                                                   \
        double dti = (DT) / n_inner;                                                       \
        INNER_FORCE();                                                                     \
        md_update_mom(n, pi, fl, 0.5 * dti);                                               \
        for (int i = 0; i < n_inner; i++) {                                                \
            md_update_gauge(n, U, pi, dti);                                                \
            counter++;                                                                     \
            if (defl_refresh > 0 && counter % defl_refresh == 0) {                         \
                orc_dirac* Dr = orc_dirac_new(L, U, mass, r, csw, mu_t);                   \
                orc_op* Ar = orc_op_dirac(Dr, ORC_DDAGD);                                  \
                orc_coarse* Cr = orc_coarse_build(Ar, P, 1);                               \
                orc_op* Ac = orc_op_coarse(Cr);                                            \
                orc_rr_evolve(Ac, k, V, vals);                                             \
                orc_op_free(Ac); orc_coarse_free(Cr); orc_op_free(Ar); orc_dirac_free(Dr); \
            }                                                                              \
            INNER_FORCE();                                                                 \
            md_update_mom(n, pi, fl, (i < n_inner - 1) ? dti : 0.5 * dti);                 \
        }                                                                                  \
    } while (0)
    if (omelyan) {
        const double lam = 0.1932;
        OUTER_FORCE();
        md_update_mom(n, pi, F, lam * h);
        for (int o = 0; o < n_outer; o++) {
            INNER(0.5 * h);
            OUTER_FORCE();
            md_update_mom(n, pi, F, (1.0 - 2.0 * lam) * h);
            INNER(0.5 * h);
            OUTER_FORCE();
            md_update_mom(n, pi, F, (o < n_outer - 1) ? 2.0 * lam * h : lam * h);
        }
    } else {
        OUTER_FORCE();
        md_update_mom(n, pi, F, 0.5 * h);
        for (int o = 0; o < n_outer; o++) {
            INNER(h);
            OUTER_FORCE();
            md_update_mom(n, pi, F, (o < n_outer - 1) ? h : 0.5 * h);
        }
    }
#undef INNER
#undef OUTER_FORCE
#undef INNER_FORCE
#undef LOW_FORCE
    free(F); free(fl);
    if (lowmode_evals) *lowmode_evals = nlow;
    return cg;
}

/* H = KE + S_g + S_f of src/hmc_trajectory.cpp:87-103 (S_f = Re phi^dag (D^dag D)^-1 phi,
 * CG from x0 = 0); out = {KE, S_g, S_f}; returns the CG iterations */
int orc_hmc_hamiltonian(int L, const ocx* U, double mass, double r, double csw, double mu_t, double beta,
                        int cg_maxit, double cg_tol, const double* pi, const ocx* phi, double* out) {
    int V = L * L;
    out[0] = orc_kinetic_energy(2 * V, pi);
    out[1] = orc_gauge_action(L, U, beta);
    orc_dirac* D = orc_dirac_new(L, U, mass, r, csw, mu_t);
    orc_op* A = orc_op_dirac(D, ORC_DDAGD);
    ocx* X = xcalloc(2 * (size_t)V, sizeof(ocx));
    double res;
    int it = orc_cg(A, phi, NULL, NULL, cg_maxit, cg_tol, X, &res);
    out[2] = creal(orc_dot(2 * V, phi, X));
    free(X);
    orc_op_free(A);
    orc_dirac_free(D);
    return it;
}

/* ---------------- even-odd HMC, src/hmc_trajectory.cpp:53-135 ----------------
 * S_eo = phi_o^dag (M^dag M)^-1 phi_o - 2 log det D_ee on the odd-site pseudofermion
 * phi_o (V complex = 2 V/2); CG from x0 = 0 on M^dag M; always leapfrog. */
static int md_force_eo(int L, const ocx* U, double mass, double r, double csw, double mu_t, double beta,
                       int cg_maxit, double cg_tol, const ocx* phi_o, double* F) {
    int V = L * L;
    orc_gauge_force(L, U, beta, F);
    orc_dirac* D = orc_dirac_new(L, U, mass, r, csw, mu_t);
    orc_op* A = orc_op_dirac(D, 100);
    ocx* x = xcalloc((size_t)V, sizeof(ocx));
    ocx* y = xcalloc((size_t)V, sizeof(ocx));
    double res;
    int it = orc_cg(A, phi_o, NULL, NULL, cg_maxit, cg_tol, x, &res);
    orc_eo_apply(D, ORC_EO_SCHUR, x, y);
    double* ff = xcalloc(2 * (size_t)V, sizeof(double));
    orc_eo_fermion_force(D, x, y, ff);
    for (int i = 0; i < 2 * V; i++) F[i] += ff[i];
    free(ff); free(x); free(y);
    orc_op_free(A);
    orc_dirac_free(D);
    return it;
}
/* the leapfrog of the even-odd trajectory, src/hmc_trajectory.cpp:135-170 */
int orc_hmc_evolve_eo(int L, ocx* U, double mass, double r, double csw, double mu_t, double beta, double tau,
                      int n_steps, int cg_maxit, double cg_tol, double* pi, const ocx* phi_o) {
    int n = 2 * L * L, cg = 0;
    double dt = tau / n_steps;
    double* F = xcalloc(n, sizeof(double));
    cg += md_force_eo(L, U, mass, r, csw, mu_t, beta, cg_maxit, cg_tol, phi_o, F);
    md_update_mom(n, pi, F, 0.5 * dt);
    for (int step = 0; step < n_steps - 1; step++) {
        md_update_gauge(n, U, pi, dt);
        cg += md_force_eo(L, U, mass, r, csw, mu_t, beta, cg_maxit, cg_tol, phi_o, F);
        md_update_mom(n, pi, F, dt);
    }
    md_update_gauge(n, U, pi, dt);
    cg += md_force_eo(L, U, mass, r, csw, mu_t, beta, cg_maxit, cg_tol, phi_o, F);
    md_update_mom(n, pi, F, 0.5 * dt);
    free(F);
    return cg;
}
/* out = {KE, S_g, S_f = Re phi_o^dag (M^dag M)^-1 phi_o, LD = -2 log det D_ee (0 if c_sw = 0)},
 * src/hmc_trajectory.cpp:115-127 */
int orc_hmc_hamiltonian_eo(int L, const ocx* U, double mass, double r, double csw, double mu_t, double beta,
                           int cg_maxit, double cg_tol, const double* pi, const ocx* phi_o, double* out) {
    int V = L * L;
    out[0] = orc_kinetic_energy(2 * V, pi);
    out[1] = orc_gauge_action(L, U, beta);
    orc_dirac* D = orc_dirac_new(L, U, mass, r, csw, mu_t);
    orc_op* A = orc_op_dirac(D, 100);
    ocx* x = xcalloc((size_t)V, sizeof(ocx));
    double res;
    int it = orc_cg(A, phi_o, NULL, NULL, cg_maxit, cg_tol, x, &res);
    out[2] = creal(orc_dot(V, phi_o, x));
    out[3] = csw != 0.0 ? -2.0 * orc_eo_logdet(D) : 0.0;
    free(x);
    orc_op_free(A);
    orc_dirac_free(D);
    return it;
}

/* ================================================================ eigensolvers */
/* lanczos_eigen, src/eigensolver_core.cpp:103-265: A (n x n Hermitian, column-major
 * A[j*n + i]) -> ascending eigenvalues and evecs[row*n + col] (column col = eigenvector
 * col).  Lanczos tridiagonalisation from e_0 with two MGS passes, then implicit QL
 * with the Wilkinson shift on the real tridiagonal. */
void orc_lanczos_eigen(int n, const ocx* A, double* evals, ocx* evecs) {
    if (n <= 0) return;
    if (n == 1) { evals[0] = creal(A[0]); evecs[0] = 1.0; return; }
    ocx* Q = xcalloc((size_t)n * n, sizeof(ocx));      /* Q[j*n + i]: vector j */
    double* al = xcalloc(n, sizeof(double));
    double* be = xcalloc(n + 1, sizeof(double));
    ocx* w = xcalloc(n, sizeof(ocx));
#define MATVEC(x, y) do { for (int i_ = 0; i_ < n; i_++) (y)[i_] = 0; \
        for (int j_ = 0; j_ < n; j_++) { ocx xj_ = (x)[j_]; for (int i_ = 0; i_ < n; i_++) (y)[i_] += A[(size_t)j_ * n + i_] * xj_; } } while (0)
    Q[0] = 1.0;
    MATVEC(Q, w);
    al[0] = creal(orc_dot(n, Q, w));
    for (int i = 0; i < n; i++) w[i] -= al[0] * Q[i];
    for (int j = 1; j < n; j++) {
        ocx* qj = Q + (size_t)j * n;
        be[j] = orc_norm(n, w);
        if (be[j] < 1e-14) {
            int found = 0;
            for (int t = 0; t < n && !found; t++) {
                for (int i = 0; i < n; i++) qj[i] = 0;
                qj[t] = 1.0;
                for (int pass = 0; pass < 2; pass++)
                    for (int i = 0; i < j; i++) {
                        ocx pr = orc_dot(n, Q + (size_t)i * n, qj);
                        for (int k = 0; k < n; k++) qj[k] -= pr * Q[(size_t)i * n + k];
                    }
                double nq = orc_norm(n, qj);
                if (nq > 1e-10) { for (int k = 0; k < n; k++) qj[k] /= nq; found = 1; }
            }
            be[j] = 0.0;
        } else {
            for (int i = 0; i < n; i++) qj[i] = w[i] / be[j];
        }
        for (int pass = 0; pass < 2; pass++)
            for (int i = 0; i < j; i++) {
                ocx pr = orc_dot(n, Q + (size_t)i * n, qj);
                for (int k = 0; k < n; k++) qj[k] -= pr * Q[(size_t)i * n + k];
            }
        double nq = orc_norm(n, qj);
        if (nq > 1e-14) for (int k = 0; k < n; k++) qj[k] /= nq;
        MATVEC(qj, w);
        al[j] = creal(orc_dot(n, qj, w));
        for (int i = 0; i < n; i++) w[i] -= al[j] * qj[i] + be[j] * Q[(size_t)(j - 1) * n + i];
        for (int pass = 0; pass < 2; pass++)
            for (int i = 0; i <= j; i++) {
                ocx pr = orc_dot(n, Q + (size_t)i * n, w);
                for (int k = 0; k < n; k++) w[k] -= pr * Q[(size_t)i * n + k];
            }
    }
#undef MATVEC
    double* d = xcalloc(n, sizeof(double));
    double* e = xcalloc(n, sizeof(double));
    double* Z = xcalloc((size_t)n * n, sizeof(double));   /* Z[k*n + i] */
    for (int i = 0; i < n; i++) { d[i] = al[i]; Z[(size_t)i * n + i] = 1.0; }
    for (int i = 0; i < n - 1; i++) e[i] = be[i + 1];
    for (int l = 0; l < n; l++) {
        int iter = 0;
        for (;;) {
            int m;
            for (m = l; m < n - 1; m++) {
                double dd = fabs(d[m]) + fabs(d[m + 1]);
                if (fabs(e[m]) <= 1e-15 * dd) break;
            }
            if (m == l) break;
            if (++iter > 200) break;
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = hypot(g, 1.0);
            g = d[m] - d[l] + e[l] / (g + copysign(r, g));
            double s = 1.0, c = 1.0, p = 0.0;
            int early = 0;
            for (int i = m - 1; i >= l; i--) {
                double f = s * e[i], b = c * e[i];
                r = hypot(f, g);
                e[i + 1] = r;
                if (r < 1e-30) { d[i + 1] -= p; e[m] = 0.0; early = 1; break; }
                s = f / r;
                c = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * s + 2.0 * c * b;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - b;
                for (int k = 0; k < n; k++) {
                    double t = Z[(size_t)k * n + i + 1];
                    Z[(size_t)k * n + i + 1] = s * Z[(size_t)k * n + i] + c * t;
                    Z[(size_t)k * n + i] = c * Z[(size_t)k * n + i] - s * t;
                }
            }
            if (!early) { d[l] -= p; e[l] = g; e[m] = 0.0; }
        }
    }
    /* ascending order (stable, as std::sort on distinct values) then back-transform */
    int* idx = xcalloc(n, sizeof(int));
    for (int i = 0; i < n; i++) idx[i] = i;
    for (int i = 1; i < n; i++) {
        int v = idx[i], j = i;
        while (j > 0 && d[idx[j - 1]] > d[v]) { idx[j] = idx[j - 1]; j--; }
        idx[j] = v;
    }
    for (int col = 0; col < n; col++) {
        evals[col] = d[idx[col]];
        for (int row = 0; row < n; row++) {
            ocx v = 0;
            for (int j = 0; j < n; j++) v += Q[(size_t)j * n + row] * Z[(size_t)j * n + idx[col]];
            evecs[(size_t)row * n + col] = v;
        }
    }
    free(idx); free(d); free(e); free(Z); free(Q); free(al); free(be); free(w);
}

/* trlm_eigensolver without the Chebyshev filter (poly_deg = 0), src/eigensolver_core.cpp:
 * 691-1057: the n_ev smallest eigenpairs of the Hermitian operator A (dimension n), Krylov
 * space n_kr, start vector start (n complex; the reference draws random_vec(n,
 * mt19937(42))).  evecs: n_ev vectors of n; returns 1 if converged. */
int orc_trlm(const orc_op* A, int n_ev, int n_kr, int max_restarts, double tol, const ocx* start,
             double* evals, ocx* evecs, int* restarts_out, int* iters_out) {
    int n = A->n;
    if (n_kr <= 0) {   /* the default Krylov space without the Chebyshev filter */
        n_kr = 2 * n_ev + 6 > n_ev + 32 ? 2 * n_ev + 6 : n_ev + 32;
        if (n_kr > n) n_kr = n;
    }
    if (n_kr > n) n_kr = n;
    if (n_kr < n_ev + 6) n_kr = n_ev + 6 < n ? n_ev + 6 : n;
    ocx* K = xcalloc((size_t)(n_kr + 1) * n, sizeof(ocx));
    double* alpha = xcalloc(n_kr, sizeof(double));
    double* beta = xcalloc(n_kr, sizeof(double));
    double* resid = xcalloc(n_kr, sizeof(double));
    ocx* r = xcalloc(n, sizeof(ocx));
    memcpy(K, start, n * sizeof(ocx));
    double nv = orc_norm(n, K);
    for (int i = 0; i < n; i++) K[i] /= nv;
    int num_locked = 0, num_converged = 0, num_keep = 0, converged = 0, restart, iters = 0;
#define KV(j) (K + (size_t)(j) * n)
    for (restart = 0; restart < max_restarts && !converged; restart++) {
        iters += n_kr - num_keep;   /* operator applications */
        for (int j = num_keep; j < n_kr; j++) {   /* lanczosStep */
            A->fn(A, KV(j), r);
            alpha[j] = creal(orc_dot(n, KV(j), r));
            for (int i = 0; i < n; i++) r[i] -= alpha[j] * KV(j)[i];
            int start_i = j > num_keep ? j - 1 : 0;
            for (int i = start_i; i < j; i++)
                for (int k = 0; k < n; k++) r[k] -= beta[i] * KV(i)[k];
            for (int pass = 0; pass < 2; pass++)
                for (int k = 0; k <= j; k++) {
                    ocx pr = orc_dot(n, KV(k), r);
                    for (int i = 0; i < n; i++) r[i] -= pr * KV(k)[i];
                }
            beta[j] = orc_norm(n, r);
            if (beta[j] > 1e-14)
                for (int i = 0; i < n; i++) KV(j + 1)[i] = r[i] / beta[j];
        }
        int dim = n_kr - num_locked, arrow = num_keep - num_locked;
        ocx* M = xcalloc((size_t)dim * dim, sizeof(ocx));
        for (int i = 0; i < dim; i++) M[(size_t)i * dim + i] = alpha[i + num_locked];
        for (int i = 0; i < arrow; i++) {
            M[(size_t)arrow * dim + i] = beta[i + num_locked];
            M[(size_t)i * dim + arrow] = beta[i + num_locked];
        }
        for (int i = arrow; i < dim - 1; i++) {
            M[(size_t)(i + 1) * dim + i] = beta[i + num_locked];
            M[(size_t)i * dim + i + 1] = beta[i + num_locked];
        }
        double* ev = xcalloc(dim, sizeof(double));
        ocx* Z = xcalloc((size_t)dim * dim, sizeof(ocx));   /* Z[row*dim + col] */
        orc_lanczos_eigen(dim, M, ev, Z);
        for (int i = 0; i < dim; i++) {
            resid[i + num_locked] = fabs(beta[n_kr - 1] * creal(Z[(size_t)(dim - 1) * dim + i]));
            alpha[i + num_locked] = ev[i];
        }
        int it_locked = 0, it_conv;
        for (int i = 1; i < dim; i++) {
            if (resid[i + num_locked] < 1e-15 * fabs(alpha[i + num_locked])) it_locked = i;
            else break;
        }
        it_conv = it_locked;
        for (int i = it_locked + 1; i < dim; i++) {
            if (resid[i + num_locked] < tol * fabs(alpha[i + num_locked])) it_conv = i;
            else break;
        }
        int keep = it_conv + (n_kr - num_converged) / 2;
        if (keep > n_kr - num_locked - 12) keep = n_kr - num_locked - 12;
        if (keep < 1) keep = 1;
        if (keep > dim - 1) keep = dim - 1;
        ocx* rot = xcalloc((size_t)keep * n, sizeof(ocx));
        for (int i = 0; i < keep; i++)
            for (int j = 0; j < dim; j++) {
                ocx c = Z[(size_t)j * dim + i];
                for (int k = 0; k < n; k++) rot[(size_t)i * n + k] += c * KV(num_locked + j)[k];
            }
        memcpy(KV(num_locked), rot, (size_t)keep * n * sizeof(ocx));
        memcpy(KV(num_locked + keep), KV(n_kr), n * sizeof(ocx));
        for (int i = 0; i < keep; i++) beta[i + num_locked] = beta[n_kr - 1] * creal(Z[(size_t)(dim - 1) * dim + i]);
        free(rot); free(M); free(ev); free(Z);
        num_converged = num_locked + it_conv;
        num_keep = num_locked + keep;
        num_locked += it_locked;
        if (num_converged >= n_ev) {
            converged = 1;
            for (int i = 1; i < n_kr; i++) {   /* ascending, insertion sort */
                int j = i;
                while (j > 0 && alpha[j - 1] > alpha[j]) {
                    double t = alpha[j]; alpha[j] = alpha[j - 1]; alpha[j - 1] = t;
                    for (int k = 0; k < n; k++) { ocx s = KV(j)[k]; KV(j)[k] = KV(j - 1)[k]; KV(j - 1)[k] = s; }
                    j--;
                }
            }
        }
    }
    for (int i = 0; i < n_ev; i++) {
        evals[i] = alpha[i];
        memcpy(evecs + (size_t)i * n, KV(i), n * sizeof(ocx));
    }
#undef KV
    if (restarts_out) *restarts_out = restart;
    if (iters_out) *iters_out = iters;
    free(K); free(alpha); free(beta); free(resid); free(r);
    return converged;
}

/* lobpcg_update, src/eigensolver_core.cpp:279-533: block LOBPCG on [X, W, P] with
 * W = M^-1 R (M = the MG cycle or the identity), the Rayleigh-Ritz problem reduced by
 * the Cholesky factor of the Gram matrix and solved by orc_lanczos_eigen; a failed
 * Cholesky falls back to one preconditioned steepest-descent step per vector.
 * The closing re-orthonormalisation updates AX_i by -<X_j,X_i> X_j, as the reference
 * does, so AX is A X only up to that term.  X0, X: [k][n]; returns the iterations run. */
static int lob_exact_ax = 0;   /* 1: AX_i -= pr AX_j, which keeps AX = A X */
static void lob_mgs(int n, int k, ocx* X, ocx* AX) {
    for (int i = 0; i < k; i++) {
        for (int j = 0; j < i; j++) {
            ocx pr = orc_dot(n, X + (size_t)j * n, X + (size_t)i * n);
            for (int t = 0; t < n; t++) X[(size_t)i * n + t] -= pr * X[(size_t)j * n + t];
            if (AX) {   /* the reference subtracts pr X_j (not pr AX_j) from AX_i (:514-517) */
                const ocx* src = lob_exact_ax ? AX + (size_t)j * n : X + (size_t)j * n;
                for (int t = 0; t < n; t++) AX[(size_t)i * n + t] -= pr * src[t];
            }
        }
        double nv = orc_norm(n, X + (size_t)i * n);
        if (nv > 1e-14) {
            for (int t = 0; t < n; t++) X[(size_t)i * n + t] /= nv;
            if (AX)
                for (int t = 0; t < n; t++) AX[(size_t)i * n + t] /= nv;
        }
    }
}
/* orc_lobpcg with exact_ax = 1: the re-orthonormalisation keeps AX = A X (AX_i -= pr AX_j)
 * in place of the reference's AX_i -= pr X_j; no reference counterpart */
int orc_lobpcg_ex(const orc_op* A, orc_mg* mg, int k, const ocx* X0, int max_iter, double tol, int exact_ax,
                  ocx* Xout, double* evals) {
    lob_exact_ax = exact_ax != 0;
    int it = orc_lobpcg(A, mg, k, X0, max_iter, tol, Xout, evals);
    lob_exact_ax = 0;
    return it;
}
int orc_lobpcg(const orc_op* A, orc_mg* mg, int k, const ocx* X0, int max_iter, double tol, ocx* Xout,
               double* evals) {
    int n = A->n;
    size_t kn = (size_t)k * n;
    ocx* X = xcalloc(kn, sizeof(ocx));
    ocx* AX = xcalloc(kn, sizeof(ocx));
    ocx* W = xcalloc(kn, sizeof(ocx));
    ocx* AW = xcalloc(kn, sizeof(ocx));
    ocx* P = xcalloc(kn, sizeof(ocx));
    ocx* AP = xcalloc(kn, sizeof(ocx));
    ocx* R = xcalloc(kn, sizeof(ocx));
    ocx* Xn = xcalloc(kn, sizeof(ocx));
    ocx* AXn = xcalloc(kn, sizeof(ocx));
    double* theta = xcalloc(k, sizeof(double));
    int sd = 3 * k;
    ocx* H = xcalloc((size_t)sd * sd, sizeof(ocx));   /* H[j*sd + i] = <S_i, A S_j> (columns) */
    ocx* M = xcalloc((size_t)sd * sd, sizeof(ocx));
    ocx* Lc = xcalloc((size_t)sd * sd, sizeof(ocx));  /* Lc[j*sd + i] = L[j][i], the row-j factor */
    ocx* Y = xcalloc((size_t)sd * sd, sizeof(ocx));
    ocx* Ht = xcalloc((size_t)sd * sd, sizeof(ocx));
    ocx* z = xcalloc(sd, sizeof(ocx));
    double* ev = xcalloc(sd, sizeof(double));
    ocx* evec = xcalloc((size_t)sd * sd, sizeof(ocx));
    ocx* Yb = xcalloc((size_t)sd * sd, sizeof(ocx));   /* Yb[col*sd + j] */
    memcpy(X, X0, kn * sizeof(ocx));
    lob_mgs(n, k, X, NULL);
    for (int i = 0; i < k; i++) A->fn(A, X + (size_t)i * n, AX + (size_t)i * n);
    int have_P = 0, iter;
    for (iter = 0; iter < max_iter; iter++) {
        double max_r = 0;
        for (int i = 0; i < k; i++) {
            theta[i] = creal(orc_dot(n, X + (size_t)i * n, AX + (size_t)i * n));
            for (int t = 0; t < n; t++) R[(size_t)i * n + t] = AX[(size_t)i * n + t] - theta[i] * X[(size_t)i * n + t];
            double ta = fabs(theta[i]) > 1e-30 ? fabs(theta[i]) : 1e-30;
            double rn = orc_norm(n, R + (size_t)i * n) / ta;
            if (rn > max_r) max_r = rn;
        }
        if (max_r < tol) break;
        for (int i = 0; i < k; i++) {
            precond(mg, n, R + (size_t)i * n, W + (size_t)i * n);
            A->fn(A, W + (size_t)i * n, AW + (size_t)i * n);
        }
        int sdim = (have_P ? 3 : 2) * k;
        const ocx* S[3] = {X, W, P};
        const ocx* AS[3] = {AX, AW, AP};
#define SV(j) (S[(j) / k] + (size_t)((j) % k) * n)
#define ASV(j) (AS[(j) / k] + (size_t)((j) % k) * n)
        for (int j = 0; j < sdim; j++)
            for (int i = 0; i < sdim; i++) {
                H[(size_t)j * sd + i] = orc_dot(n, SV(i), ASV(j));
                M[(size_t)j * sd + i] = orc_dot(n, SV(i), SV(j));
            }
        for (int i = 0; i < sdim; i++)
            for (int j = i + 1; j < sdim; j++) {
                ocx ah = 0.5 * (H[(size_t)j * sd + i] + conj(H[(size_t)i * sd + j]));
                H[(size_t)j * sd + i] = ah; H[(size_t)i * sd + j] = conj(ah);
                ocx am = 0.5 * (M[(size_t)j * sd + i] + conj(M[(size_t)i * sd + j]));
                M[(size_t)j * sd + i] = am; M[(size_t)i * sd + j] = conj(am);
            }
        /* Cholesky M = L L^dag, row by row: L[j][i] for i < j, then L[j][j] */
        memset(Lc, 0, (size_t)sd * sd * sizeof(ocx));
        int chol_ok = 1;
        for (int j = 0; j < sdim && chol_ok; j++) {
            for (int i = 0; i < j; i++) {
                ocx sm = M[(size_t)j * sd + i];
                for (int p = 0; p < i; p++) sm -= Lc[(size_t)i * sd + p] * conj(Lc[(size_t)j * sd + p]);
                Lc[(size_t)j * sd + i] = sm / Lc[(size_t)i * sd + i];
            }
            ocx sm = M[(size_t)j * sd + j];
            for (int p = 0; p < j; p++) sm -= Lc[(size_t)j * sd + p] * conj(Lc[(size_t)j * sd + p]);
            double d = creal(sm);
            if (d < 1e-14) chol_ok = 0;
            else Lc[(size_t)j * sd + j] = sqrt(d);
        }
        if (!chol_ok) {
            for (int i = 0; i < k; i++) {
                ocx* xi = X + (size_t)i * n;
                ocx xAw = orc_dot(n, xi, AW + (size_t)i * n);
                ocx wAw = orc_dot(n, W + (size_t)i * n, AW + (size_t)i * n);
                if (cabs(wAw) > 1e-30) {
                    ocx al = -xAw / wAw;
                    for (int t = 0; t < n; t++) xi[t] += al * W[(size_t)i * n + t];
                }
                double nv = orc_norm(n, xi);
                if (nv > 1e-14) for (int t = 0; t < n; t++) xi[t] /= nv;
            }
            lob_mgs(n, k, X, NULL);
            for (int i = 0; i < k; i++) A->fn(A, X + (size_t)i * n, AX + (size_t)i * n);
            have_P = 0;
            continue;
        }
        /* Ht = L^-1 H L^-dag: Y_col = L^-1 H_col, then L^dag z = Y_col */
        for (int col = 0; col < sdim; col++) {
            for (int i = 0; i < sdim; i++) {
                ocx sm = H[(size_t)col * sd + i];
                for (int p = 0; p < i; p++) sm -= Lc[(size_t)i * sd + p] * Y[(size_t)col * sd + p];
                Y[(size_t)col * sd + i] = sm / Lc[(size_t)i * sd + i];
            }
            for (int i = sdim - 1; i >= 0; i--) {
                ocx sm = Y[(size_t)col * sd + i];
                for (int p = i + 1; p < sdim; p++) sm -= conj(Lc[(size_t)p * sd + i]) * z[p];
                z[i] = sm / Lc[(size_t)i * sd + i];
            }
            for (int i = 0; i < sdim; i++) Ht[(size_t)col * sdim + i] = z[i];   /* packed sdim x sdim */
        }
        for (int i = 0; i < sdim; i++)
            for (int j = i + 1; j < sdim; j++) {
                ocx a = 0.5 * (Ht[(size_t)j * sdim + i] + conj(Ht[(size_t)i * sdim + j]));
                Ht[(size_t)j * sdim + i] = a; Ht[(size_t)i * sdim + j] = conj(a);
            }
        orc_lanczos_eigen(sdim, Ht, ev, evec);   /* evec[row*sdim + col] */
        for (int col = 0; col < k; col++)
            for (int i = sdim - 1; i >= 0; i--) {
                ocx sm = evec[(size_t)i * sdim + col];
                for (int p = i + 1; p < sdim; p++) sm -= conj(Lc[(size_t)p * sd + i]) * Yb[(size_t)col * sd + p];
                Yb[(size_t)col * sd + i] = sm / Lc[(size_t)i * sd + i];
            }
        for (int i = 0; i < k; i++) {
            ocx* xn = Xn + (size_t)i * n;
            ocx* axn = AXn + (size_t)i * n;
            memset(xn, 0, n * sizeof(ocx));
            memset(axn, 0, n * sizeof(ocx));
            for (int j = 0; j < sdim; j++) {
                ocx c = Yb[(size_t)i * sd + j];
                const ocx* sv = SV(j);
                const ocx* asv = ASV(j);
                for (int t = 0; t < n; t++) { xn[t] += c * sv[t]; axn[t] += c * asv[t]; }
            }
        }
#undef SV
#undef ASV
        for (int i = 0; i < k; i++) {
            ocx* p = P + (size_t)i * n;
            ocx* ap = AP + (size_t)i * n;
            for (int t = 0; t < n; t++) {
                p[t] = Xn[(size_t)i * n + t] - X[(size_t)i * n + t];
                ap[t] = AXn[(size_t)i * n + t] - AX[(size_t)i * n + t];
            }
            double np = orc_norm(n, p);
            if (np > 1e-14)
                for (int t = 0; t < n; t++) { p[t] /= np; ap[t] /= np; }
        }
        memcpy(X, Xn, kn * sizeof(ocx));
        memcpy(AX, AXn, kn * sizeof(ocx));
        have_P = 1;
        lob_mgs(n, k, X, AX);
    }
    for (int i = 0; i < k; i++) evals[i] = creal(orc_dot(n, X + (size_t)i * n, AX + (size_t)i * n));
    memcpy(Xout, X, kn * sizeof(ocx));
    free(X); free(AX); free(W); free(AW); free(P); free(AP); free(R); free(Xn); free(AXn); free(theta);
    free(H); free(M); free(Lc); free(Y); free(Ht); free(z); free(ev); free(evec); free(Yb);
    return iter;
}

/* cg_solve_deflated, src/solvers.cpp:111-213: E = U^dag A U (E_ij = <u_i, A u_j>),
 * x0 = U E^-1 U^dag b, then preconditioned CG whose preconditioned residual is
 * projected, z = M^-1 r - U E^-1 (AU)^dag M^-1 r; E solves by CoarseOp::solve
 * (partial-pivot elimination, src/coarse_op.h:101-137, as orc_dense_solve) */
int orc_cg_deflated(const orc_op* A, const ocx* rhs, orc_mg* mg, int ndefl, const ocx* U, int maxit, double tol,
                    ocx* x, double* final_res) {
    int n = A->n;
    memset(x, 0, n * sizeof(ocx));
    double rhs_norm = orc_norm(n, rhs);
    if (rhs_norm < 1e-30) { *final_res = 0; return 0; }
    ocx* AU = xcalloc((size_t)(ndefl > 0 ? ndefl : 1) * n, sizeof(ocx));
    for (int i = 0; i < ndefl; i++) A->fn(A, U + (size_t)i * n, AU + (size_t)i * n);
    orc_dense E = {ndefl, xcalloc((size_t)(ndefl > 0 ? ndefl * ndefl : 1), sizeof(ocx))};
    for (int i = 0; i < ndefl; i++)
        for (int j = 0; j < ndefl; j++) E.mat[(size_t)j * ndefl + i] = orc_dot(n, U + (size_t)i * n, AU + (size_t)j * n);
    ocx* t = xcalloc(ndefl + 1, sizeof(ocx));
    ocx* c = xcalloc(ndefl + 1, sizeof(ocx));
    for (int i = 0; i < ndefl; i++) t[i] = orc_dot(n, U + (size_t)i * n, rhs);
    if (ndefl) orc_dense_solve(&E, t, c);
    for (int i = 0; i < ndefl; i++)
        for (int k = 0; k < n; k++) x[k] += c[i] * U[(size_t)i * n + k];
    ocx* r = xcalloc(n, sizeof(ocx));
    ocx* z = xcalloc(n, sizeof(ocx));
    ocx* p = xcalloc(n, sizeof(ocx));
    ocx* Ap = xcalloc(n, sizeof(ocx));
    A->fn(A, x, Ap);
    for (int i = 0; i < n; i++) r[i] = rhs[i] - Ap[i];
#define PROJECT()                                                                          \
    do {                                                                                   \
        precond(mg, n, r, z);                                                              \
        for (int i = 0; i < ndefl; i++) t[i] = orc_dot(n, AU + (size_t)i * n, z);          \
        if (ndefl) orc_dense_solve(&E, t, c);                                              \
        for (int i = 0; i < ndefl; i++)                                                    \
            for (int k = 0; k < n; k++) z[k] -= c[i] * U[(size_t)i * n + k];               \
    } while (0)
    PROJECT();
    memcpy(p, z, n * sizeof(ocx));
    ocx rz = orc_dot(n, r, z);
    int iter = 0;
    while (iter < maxit) {
        A->fn(A, p, Ap);
        ocx alpha = rz / orc_dot(n, p, Ap);
        for (int i = 0; i < n; i++) { x[i] += alpha * p[i]; r[i] -= alpha * Ap[i]; }
        iter++;
        if (orc_norm(n, r) / rhs_norm < tol) break;
        PROJECT();
        ocx rz_new = orc_dot(n, r, z);
        ocx beta = rz_new / rz;
        rz = rz_new;
        for (int i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
    }
#undef PROJECT
    *final_res = orc_norm(n, r) / rhs_norm;
    free(AU); free(E.mat); free(t); free(c); free(r); free(z); free(p); free(Ap);
    return iter;
}

/* SparseCoarseOp::solve with deflation vectors, src/coarse_op.h:492-547: x0 = sum_i
 * (v_i^dag b / lambda_i) v_i (lambda_i > 1e-14), then CG from x0 with the stopping rule
 * |r|/|b| < tol checked after each update; returns the iterations */
int orc_deflated_coarse_cg(const orc_op* A, int ndefl, const ocx* vecs, const double* vals, const ocx* b,
                           int maxit, double tol, ocx* x) {
    int n = A->n;
    memset(x, 0, n * sizeof(ocx));
    double bnorm = orc_norm(n, b);
    if (bnorm < 1e-30) return 0;
    for (int i = 0; i < ndefl; i++)
        if (vals[i] > 1e-14) {
            ocx c = orc_dot(n, vecs + (size_t)i * n, b) / vals[i];
            for (int k = 0; k < n; k++) x[k] += c * vecs[(size_t)i * n + k];
        }
    ocx* r = xcalloc(n, sizeof(ocx));
    ocx* p = xcalloc(n, sizeof(ocx));
    ocx* Ap = xcalloc(n, sizeof(ocx));
    A->fn(A, x, Ap);
    for (int i = 0; i < n; i++) r[i] = b[i] - Ap[i];
    memcpy(p, r, n * sizeof(ocx));
    ocx rr = orc_dot(n, r, r);
    int it = 0;
    while (it < maxit) {
        A->fn(A, p, Ap);
        ocx pAp = orc_dot(n, p, Ap);
        if (cabs(pAp) < 1e-30) break;
        ocx a = rr / pAp;
        for (int i = 0; i < n; i++) { x[i] += a * p[i]; r[i] -= a * Ap[i]; }
        it++;
        double rn = orc_norm(n, r);
        if (rn / bnorm < tol || isnan(rn) || isinf(rn)) break;
        ocx rn2 = orc_dot(n, r, r);
        ocx be = rn2 / rr;
        rr = rn2;
        for (int i = 0; i < n; i++) p[i] = r[i] + be * p[i];
    }
    free(r); free(p); free(Ap);
    return it;
}

/* the OpenMP thread count of the oracle's parallel loops (the process's OpenMP runtime
 * may have started before OMP_NUM_THREADS was set, so it is set here explicitly) */
void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int orc_get_threads(void) { return omp_get_max_threads(); }
