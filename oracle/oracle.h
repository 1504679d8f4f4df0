/*
 * oracle.h -- CPU FP64 restatement of the Schwinger_MG hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it may be loaded by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and
 * nowhere else.  The product path (schwinger_mg_amd, libsmg.so) never links,
 * loads or calls it.
 *
 * Every function cites the reference file:line whose algorithm it restates.
 * Complex numbers are C99 `double complex`; at the ABI every complex array is
 * interleaved (re, im) doubles, i.e. the same bytes as std::vector<cx>.
 *
 * Layouts (identical to the reference):
 *   spinor psi[2*s + spin], s = x + L*y           (src/lattice.h:19, src/dirac.h:99)
 *   gauge  U[mu*V + s], mu in {0 = x, 1 = y}       (src/gauge.h:12)
 *   coarse vector c[b*k + j]                       (src/prolongator.h:13)
 *   coarse links link[(b*9 + d)*k*k + ki*k + kj]    (src/coarse_op.h:159-166)
 */
#ifndef SMG_ORACLE_H
#define SMG_ORACLE_H
#include <complex.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef double complex ocx;

/* ---------------- geometry / gauge ---------------- */
void orc_eo_maps(int L, int* even_sites, int* odd_sites, int* eo_index);
void orc_clover(int L, const ocx* U, double* F01);
double orc_avg_plaq(int L, const ocx* U);
double orc_gauge_action(int L, const ocx* U, double beta);
void orc_gauge_force(int L, const ocx* U, double beta, double* F);

/* ---------------- Dirac operator handle ---------------- */
typedef struct orc_dirac orc_dirac;
orc_dirac* orc_dirac_new(int L, const ocx* U, double mass, double r, double csw, double mu_t);
void orc_dirac_free(orc_dirac* D);
void orc_set_threads(int n);
int orc_get_threads(void);
const double* orc_dirac_clover(const orc_dirac* D);
enum { ORC_D = 0, ORC_DDAG = 1, ORC_DDAGD = 2, ORC_G5D = 3 };
void orc_dirac_apply(const orc_dirac* D, int kind, const ocx* src, ocx* dst);
void orc_dirac_delta(const orc_dirac* D, const ocx* src, ocx* dst, const double* pi, double dt);

/* even-odd: kinds */
enum { ORC_EO_HOP_EO = 0, ORC_EO_HOP_OE = 1, ORC_EO_EE_INV = 2, ORC_EO_OO_INV = 3,
       ORC_EO_OO = 4, ORC_EO_SCHUR = 5, ORC_EO_SCHUR_DAG = 6, ORC_EO_SCHUR_DAG_SCHUR = 7 };
void orc_eo_apply(const orc_dirac* D, int kind, const ocx* src, ocx* dst);
void orc_eo_reconstruct(const orc_dirac* D, const ocx* x_o, const ocx* phi_full, ocx* out_full);
void orc_eo_reconstruct_full(const orc_dirac* D, const ocx* x_o, ocx* out_full);
double orc_eo_logdet(const orc_dirac* D);

/* ---------------- BLAS-1 ---------------- */
ocx orc_dot(int n, const ocx* a, const ocx* b);
void orc_dot2(int n, const ocx* a, const ocx* b, double* out);
double orc_norm(int n, const ocx* a);

/* ---------------- operators ---------------- */
typedef struct orc_op orc_op;
orc_op* orc_op_dirac(const orc_dirac* D, int kind); /* ORC_DDAGD or 100 = Schur M^dag M */
void orc_op_apply(const orc_op* A, const ocx* src, ocx* dst);
int orc_op_dim(const orc_op* A);
void orc_op_free(orc_op* A);

/* ---------------- prolongator (geometric aggregation) ---------------- */
typedef struct orc_prolong orc_prolong;
/* k coarse dof per aggregate; chiral = 1 splits each of k/2 null vectors into its
 * gamma5 = +1 (first nd/2 dof of a site) and gamma5 = -1 parts */
orc_prolong* orc_prolong_new(int nx, int ny, int nd, int bx, int by, int k, int chiral);
void orc_prolong_free(orc_prolong* P);
void orc_prolong_build(orc_prolong* P, const ocx* null_vecs);
void orc_prolong_get(const orc_prolong* P, ocx* bv);
void orc_prolong_set(orc_prolong* P, const ocx* bv);
void orc_prolong_apply(const orc_prolong* P, const ocx* coarse, ocx* fine);
void orc_restrict_apply(const orc_prolong* P, const ocx* fine, ocx* coarse);
int orc_prolong_coarse_dim(const orc_prolong* P);

/* ---------------- 9-point coarse operator ---------------- */
typedef struct orc_coarse orc_coarse;
orc_coarse* orc_coarse_build(const orc_op* A, const orc_prolong* P, int colored);
void orc_coarse_free(orc_coarse* C);
void orc_coarse_get(const orc_coarse* C, ocx* link);
void orc_coarse_set(orc_coarse* C, const ocx* link);
orc_op* orc_op_coarse(const orc_coarse* C);

/* dense Galerkin operator + LU (src/coarse_op.h:18-137) */
typedef struct orc_dense orc_dense;
orc_dense* orc_dense_build(const orc_op* A, const orc_prolong* P);
void orc_dense_free(orc_dense* M);
void orc_dense_apply(const orc_dense* M, const ocx* x, ocx* y);
void orc_dense_solve(const orc_dense* M, const ocx* b, ocx* x);
void orc_dense_get(const orc_dense* M, ocx* cols);

/* ---------------- smoothers ---------------- */
void orc_mr_smooth(const orc_op* A, ocx* x, const ocx* b, int niter);
void orc_richardson(const orc_op* A, ocx* x, const ocx* b, int niter, double omega);
double orc_lambda_max(const orc_op* A, const ocx* v0, int niter);
void orc_null_space(const orc_op* A, int k, int outer_iters, const ocx* starts, ocx* out);

/* ---------------- multigrid ---------------- */
typedef struct orc_mg orc_mg;
typedef struct {
    int n_levels;
    int bx[4], by[4];      /* aggregate shape per transfer (level l -> l+1) */
    int k[4];              /* near-null vectors per transfer */
    int null_iters;
    int pre_smooth, post_smooth;
    int w_cycle;
    int coarse_solver;     /* 0 = CG on the 9-point op (tol, maxit); 1 = dense LU */
    int coarse_maxit;
    double coarse_tol;
    int colored_build;     /* 1 = colored probing, 0 = one probe per coarse column */
    int chiral;            /* 1 = chiral aggregation at every level */
} orc_mg_params;
/* fine op acts on nx*ny sites with nd dof each.  starts: the k[l] start vectors of
 * every level, concatenated (level 0 first). */
orc_mg* orc_mg_new(const orc_op* fine, int nx, int ny, int nd, const orc_mg_params* prm,
                   const ocx* starts);
/* the same hierarchy from given parts: block vectors and coarse links of every
 * transfer, concatenated (level 0 first), in orc_prolong_get / orc_coarse_get layout */
orc_mg* orc_mg_assemble(const orc_op* fine, int nx, int ny, int nd, const orc_mg_params* prm,
                        const ocx* block_vecs, const ocx* links);
void orc_mg_free(orc_mg* mg);
void orc_mg_set_symmetric(orc_mg* mg, double damping, const ocx* starts);
void orc_mg_precondition(orc_mg* mg, const ocx* b, ocx* x);
const orc_prolong* orc_mg_prolong(const orc_mg* mg, int level);
const orc_coarse* orc_mg_coarse(const orc_mg* mg, int level);
double orc_mg_omega(const orc_mg* mg, int level);
void orc_mg_null_vecs(const orc_mg* mg, int level, ocx* out);

/* ---------------- Krylov ---------------- */
int orc_cg(const orc_op* A, const ocx* rhs, const ocx* x0, orc_mg* mg, int maxit, double tol,
           ocx* x, double* final_res);
int orc_fcg(const orc_op* A, const ocx* rhs, orc_mg* mg, int maxit, double tol, int trunc,
            int ref_projection, ocx* x, double* final_res);
int orc_fgmres(const orc_op* A, const ocx* rhs, orc_mg* mg, int restart, int maxit, double tol,
               ocx* x, double* final_res);
int orc_coarse_cg(const orc_op* A, const ocx* b, int maxit, double tol, ocx* x);
int orc_fgmres_ritz(const orc_op* A, const ocx* rhs, orc_mg* mg, int restart, int maxit, double tol,
                    int n_ritz, ocx* x, double* final_res, int* n_ritz_out, double* ritz_vals, ocx* ritz_vecs);

/* ---------------- forces ---------------- */
void orc_fermion_force(const orc_dirac* D, const ocx* X, double* F);
void orc_fermion_force_bilinear(const orc_dirac* D, const ocx* chi, const ocx* X, double* F);
void orc_hopping_force(const orc_dirac* D, const ocx* chi, const ocx* X, double* F);
void orc_clover_deriv_insert(const orc_dirac* D, const double* w, double factor, double* F);
void orc_logdet_ee_force(const orc_dirac* D, double* F);
void orc_eo_fermion_force(const orc_dirac* D, const ocx* x_o, const ocx* y_o, double* F);

/* ---------------- eigensolvers and deflation ---------------- */
void orc_lanczos_eigen(int n, const ocx* A, double* evals, ocx* evecs);
int orc_trlm(const orc_op* A, int n_ev, int n_kr, int max_restarts, double tol, const ocx* start,
             double* evals, ocx* evecs, int* restarts_out, int* iters_out);
int orc_lobpcg(const orc_op* A, orc_mg* mg, int k, const ocx* X0, int max_iter, double tol, ocx* Xout,
               double* evals);
int orc_lobpcg_ex(const orc_op* A, orc_mg* mg, int k, const ocx* X0, int max_iter, double tol, int exact_ax,
                  ocx* Xout, double* evals);
int orc_cg_deflated(const orc_op* A, const ocx* rhs, orc_mg* mg, int ndefl, const ocx* U, int maxit, double tol,
                    ocx* x, double* final_res);
int orc_deflated_coarse_cg(const orc_op* A, int ndefl, const ocx* vecs, const double* vals, const ocx* b,
                           int maxit, double tol, ocx* x);

/* ---------------- U(1) heatbath (no reference counterpart) ---------------- */
void orc_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double orc_von_mises(double kappa, uint32_t ctr0, uint32_t ctr1, const uint32_t key[2]);
void orc_heatbath(int L, ocx* U, double beta, int n_sweeps, uint64_t seed, uint32_t first_sweep);

/* ---------------- HMC molecular dynamics ---------------- */
double orc_kinetic_energy(int n, const double* pi);
int orc_hmc_evolve(int L, ocx* U, double mass, double r, double csw, double mu_t, double beta, double tau,
                   int n_steps, int omelyan, int cg_maxit, double cg_tol, double* pi, const ocx* phi);
void orc_rr_evolve(const orc_op* A, int k, ocx* V, double* vals);
void orc_coarse_lowmode_force(const orc_dirac* D, const orc_prolong* P, int k, const ocx* V, const double* vals,
                              const ocx* phi, int smooth, double* F);
int orc_hmc_ms_evolve(int L, ocx* U, double mass, double r, double csw, double mu_t, double beta, double tau,
                      int n_outer, int n_inner, int omelyan, int inner_smooth, int defl_refresh, int cg_maxit,
                      double cg_tol, const orc_prolong* P, int k, ocx* V, double* vals, double* pi, const ocx* phi,
                      int* lowmode_evals);
int orc_hmc_nested_evolve(int L, ocx* U, double mass, double r, double csw, double mu_t, double beta, double tau,
                          int n_outer, int n_gauge, int omelyan_outer, int omelyan_inner, int cg_maxit,
                          double cg_tol, double* pi, const ocx* phi);
int orc_hmc_hamiltonian(int L, const ocx* U, double mass, double r, double csw, double mu_t, double beta,
                        int cg_maxit, double cg_tol, const double* pi, const ocx* phi, double* out);
int orc_hmc_evolve_eo(int L, ocx* U, double mass, double r, double csw, double mu_t, double beta, double tau,
                      int n_steps, int cg_maxit, double cg_tol, double* pi, const ocx* phi_o);
int orc_hmc_hamiltonian_eo(int L, const ocx* U, double mass, double r, double csw, double mu_t, double beta,
                           int cg_maxit, double cg_tol, const double* pi, const ocx* phi_o, double* out);

#ifdef __cplusplus
}
#endif
#endif
