/*
 * smg.h -- C-ABI of the MI355X (gfx950) Schwinger-model Wilson-clover / multigrid /
 * fermion-force library (libsmg.so).
 *
 * Conventions
 *  - Every entry point returns 0 on success and a negative code on failure
 *    (-1 invalid argument, -2 HIP runtime error, -3 internal error); the message
 *    of the last failure on the calling thread is returned by smg_last_error().
 *    No C++ exception crosses this boundary.
 *  - Complex arrays are interleaved (re, im) doubles: the bytes of the
 *    reference's std::vector<std::complex<double>> (src/types.h:7-8).  Layouts are
 *    the reference's: spinor psi[2*s + spin] with s = x + L*y (src/lattice.h:19,
 *    src/dirac.h:99); gauge U[mu*V + s] (src/gauge.h:12); forces F[mu*V + s];
 *    coarse vectors c[b*k + j] (src/prolongator.h:13); coarse links
 *    link[(b*9 + d)*k*k + ki*k + kj] (src/coarse_op.h:159-166); half-lattice
 *    vectors are indexed by eo_index (src/lattice.h:28-41).
 *  - Unsuffixed calls take HOST arrays (the reference's Vec semantics: inputs
 *    const, outputs pre-sized by the caller).  *_dev calls take DEVICE pointers
 *    of precision `prec` and an optional hipStream_t (NULL: the library orders
 *    itself after all prior work of the calling thread's null stream).
 *  - Handles are opaque, owned by the caller, and used by one host thread at a
 *    time (the reference's objects carry the same restriction).
 */
#ifndef SMG_H
#define SMG_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SMG_F64 0
#define SMG_F32 1
#define SMG_F16 2 /* coarse-link storage only (FP32 arithmetic) */
#define SMG_BF16V 3 /* FP16 coarse links with BF16 vector storage (FP32 arithmetic) */

/* Dirac operator kinds: D, D^dag, D^dag D, gamma5 D (src/dirac.h:115-152) */
#define SMG_D 0
#define SMG_DDAG 1
#define SMG_DDAGD 2
#define SMG_G5D 3

/* even-odd kinds (src/dirac.h:253-445) */
#define SMG_EO_HOP_EO 0          /* apply_eo: even -> odd hopping      */
#define SMG_EO_HOP_OE 1          /* apply_oe: odd -> even hopping      */
#define SMG_EO_EE_INV 2          /* apply_ee_inv                       */
#define SMG_EO_OO_INV 3          /* apply_oo_inv                       */
#define SMG_EO_OO 4              /* apply_oo                           */
#define SMG_EO_SCHUR 5           /* apply_schur: M                     */
#define SMG_EO_SCHUR_DAG 6       /* apply_schur_dag: M^dag             */
#define SMG_EO_SCHUR_DAG_SCHUR 7 /* apply_schur_dag_schur: M^dag M     */

/* solver operators */
#define SMG_OP_DDAGD 0 /* D^dag D on the full lattice (DiracOp::as_DdagD_op, src/dirac.h:177-179) */
#define SMG_OP_SCHUR 1 /* M^dag M on odd sites (EvenOddDiracOp::apply_schur_dag_schur) */

/* smoother kinds (src/smoother.cpp) */
#define SMG_SMOOTH_MR 0
#define SMG_SMOOTH_RICHARDSON 1

typedef struct smg_dirac_s smg_dirac;
typedef struct smg_prolongator_s smg_prolongator;
typedef struct smg_coarse_s smg_coarse;
typedef struct smg_mg_s smg_mg;
typedef struct smg_lattices_s smg_lattices;
typedef struct smg_lattices_mg_s smg_lattices_mg;
typedef struct smg_eo_mg_s smg_eo_mg;

/* ------------------------------------------------------------------ runtime */
/* Select the HIP device and create the library's stream.  Optional: the first
 * call of any other entry point initialises device 0. */
int smg_init(int device);
const char* smg_last_error(void);
int smg_version(void);
/* Block until all library work has completed. */
int smg_synchronize(void);
/* The hipStream_t the library launches on (for event timing by the caller). */
int smg_get_stream(void** stream);

/* ------------------------------------------------------------------ geometry */
/* Lattice::init_eo, src/lattice.h:28-41 (host arrays of V/2, V/2 and V ints) */
int smg_eo_maps(int L, int* even_sites, int* odd_sites, int* eo_index);
/* GaugeField::avg_plaq, src/gauge.h:31-36 */
int smg_avg_plaq(int L, const double* U, double* out);
/* gauge_action, src/hmc_forces.cpp:80-86 */
int smg_gauge_action(int L, const double* U, double beta, double* out);
/* gauge_force, src/hmc_forces.cpp:88-111 (F = -dS_g/dtheta) */
int smg_gauge_force(int L, const double* U, double beta, double* F);

/* ------------------------------------------------------------------ Dirac operator */
/* DiracOp(lat, gauge, m, r, csw, mu_t), src/dirac.h:17-22 (U copied to the device,
 * clover field computed on the device, src/dirac.h:26-52) */
int smg_dirac_create(smg_dirac** out, int L, const double* U, double mass, double r, double csw, double mu_t);
int smg_dirac_destroy(smg_dirac* D);
/* new gauge field for the same operator (clover recomputed) */
int smg_dirac_set_gauge(smg_dirac* D, const double* U);
/* DiracOp::clover_field (V doubles) */
int smg_dirac_clover(smg_dirac* D, double* F01);
/* DiracOp::apply / apply_dag / apply_DdagD / apply_g5D, src/dirac.h:115-152 */
int smg_dirac_apply(smg_dirac* D, int kind, const double* src, double* dst);
/* the same on device vectors of precision prec, nrhs vectors of 2*V complex each */
int smg_dirac_apply_dev(smg_dirac* D, int kind, int prec, const void* src, void* dst, int nrhs, void* stream);
/* DiracOp::apply_delta_D, src/dirac.h:159-174 (pi: 2*V doubles) */
int smg_dirac_delta(smg_dirac* D, const double* src, double* dst, const double* pi, double dt);
/* EvenOddDiracOp operators (half-lattice vectors of V complex), src/dirac.h:253-445 */
int smg_eo_apply(smg_dirac* D, int kind, const double* src, double* dst);
int smg_eo_apply_dev(smg_dirac* D, int kind, int prec, const void* src, void* dst, void* stream);
/* EvenOddDiracOp::reconstruct (phi != NULL) / reconstruct_full (phi == NULL), src/dirac.h:515-552 */
int smg_eo_reconstruct(smg_dirac* D, const double* x_o, const double* phi, double* out_full);
/* EvenOddDiracOp::log_det_ee, src/dirac.h:528-537 */
int smg_eo_logdet(smg_dirac* D, double* out);

/* ------------------------------------------------------------------ BLAS-1 */
/* dot (src/linalg.cpp:8-18): out2 = {Re, Im} of sum conj(a) b; device vectors of n complex */
int smg_dot_dev(const void* a, const void* b, size_t n, int prec, double* out2, void* stream);
/* norm (src/linalg.cpp:20) */
int smg_norm_dev(const void* a, size_t n, int prec, double* out, void* stream);
/* No counterpart in the reference: smg_dot_dev for callers that cannot allocate
 * (code under stream capture up to the final read-back, tight serving loops).  It
 * calls no hipMalloc/hipFree: the reduction's partial rows (at most 1024
 * workgroups), its completion counter and its result live in a 16.5 KiB file-scope
 * device array that each GPU holds until the process ends.  Kernels and the
 * read-back run on `stream` itself (NULL: the null stream), which is synchronised
 * before the call returns; the result is bit-identical to smg_dot_dev's.  Every
 * call on one GPU shares that array: calls that overlap on one GPU (other host
 * threads, other streams) race and must not be issued. */
int smg_dot_dev_noalloc(const void* a, const void* b, size_t n, int prec, double* out2, void* stream);

/* ------------------------------------------------------------------ prolongator */
/* Prolongator(lat, bx, by, k), src/prolongator.h:35-58, generalised to nd dof per
 * site so that it also restates CoarseProlongator (src/prolongator.h:149-228):
 * linear blocks of 4*k consecutive coarse dofs on an nx x ny coarse lattice are
 * the 4 x 1 aggregates with nd = k.  k is the coarse dof per aggregate.
 * chiral = 0: the reference's aggregation (k null vectors, both chiralities in
 * each block vector).  chiral = 1 (no reference counterpart): k/2 null vectors,
 * each split into its gamma5 = +1 part (coarse dof j < k/2; the first nd/2 dof of
 * a site) and its gamma5 = -1 part (coarse dof k/2 + j), so the coarse operator
 * keeps a diagonal gamma5. */
int smg_prolongator_create(smg_prolongator** out, int nx, int ny, int nd, int bx, int by, int k, int chiral);
int smg_prolongator_destroy(smg_prolongator* P);
/* Prolongator::build_from_vectors, src/prolongator.h:133-141 (k host vectors of nx*ny*nd
 * complex; k/2 with chiral aggregation) */
int smg_prolongator_build(smg_prolongator* P, const double* null_vecs);
/* block vectors [nblocks][k][bx*by*nd] (Prolongator::block_vecs) */
int smg_prolongator_get(smg_prolongator* P, double* bv);
int smg_prolongator_set(smg_prolongator* P, const double* bv);
/* Prolongator::prolong / restrict_vec, src/prolongator.h:77-108 */
int smg_prolong(smg_prolongator* P, const double* coarse, double* fine);
int smg_restrict(smg_prolongator* P, const double* fine, double* coarse);

/* ------------------------------------------------------------------ coarse operator */
/* Galerkin A_c = P^dag (D^dag D) P as the 9-point operator (SparseCoarseOp::build,
 * src/coarse_op.h:241-279; equal to CoarseOp::build, src/coarse_op.h:40-43, for
 * aggregates of at least 2 x 2 sites) */
int smg_coarse_create_galerkin(smg_coarse** out, smg_dirac* D, smg_prolongator* P);
/* next level: A_c = P^dag A P for a coarse operator A (CoarseOp::build_generic, src/coarse_op.h:26-37) */
int smg_coarse_create_from_coarse(smg_coarse** out, smg_coarse* A, smg_prolongator* P);
/* a coarse operator from explicit links */
int smg_coarse_create_links(smg_coarse** out, int nx, int ny, int k, const double* links);
int smg_coarse_destroy(smg_coarse* C);
int smg_coarse_get_links(smg_coarse* C, double* links);
/* CoarseOp (dense A_c), src/coarse_op.h:18-59: the coarse operator as a dense
 * n x n matrix in the reference's column layout, M[j*n + i] = A_c[i][j]
 * (n = coarse dim <= 16384).  Equal to CoarseOp::build(D, P) for a Galerkin
 * operator built from the same prolongator. */
int smg_coarse_dense(smg_coarse* C, double* M);
/* CoarseOp::solve, src/coarse_op.h:101-137: x = A_c^-1 b (FP64 Gauss-Jordan
 * inverse; A_c must be Hermitian positive definite, as Galerkin projections of
 * D^dag D are; n <= 8192) */
int smg_coarse_dense_solve(smg_coarse* C, const double* b, double* x);
/* SparseCoarseOp::apply_to, src/coarse_op.h:282-297, links read in precision prec */
int smg_coarse_apply(smg_coarse* C, int prec, const double* src, double* dst);
/* SparseCoarseOp::solve, src/coarse_op.h:492-547: CG from x0 = 0, or from the
 * deflated guess once smg_coarse_setup_deflation / smg_coarse_set_deflation has
 * given C a deflation space */
int smg_coarse_solve(smg_coarse* C, const double* b, double* x, int maxit, double tol, int* iters);

/* ------------------------------------------------------------------ smoothers */
/* mr_smooth (src/smoother.cpp:9-23) / richardson_smooth_op (:63-73) on D^dag D; x in/out */
int smg_smooth_dirac(smg_dirac* D, int kind, double* x, const double* b, int niter, double omega);
/* mr_smooth_op / richardson_smooth_op on a coarse operator */
int smg_smooth_coarse(smg_coarse* C, int kind, double* x, const double* b, int niter, double omega);
/* estimate_lambda_max (src/smoother.cpp:43-59) of D^dag D from the start vector v0 */
int smg_lambda_max_dirac(smg_dirac* D, const double* v0, int niter, double* out);

/* ------------------------------------------------------------------ multigrid */
typedef struct {
    int n_levels;          /* 2..5 */
    int bx[4], by[4];      /* aggregate shape of transfer l (level l -> l+1) */
    int k[4];              /* near-null vectors of transfer l */
    int null_iters;        /* outer inverse-iteration steps (5 MR sweeps each) */
    int pre_smooth, post_smooth;
    int w_cycle;
    int coarse_maxit;      /* coarsest CG */
    double coarse_tol;
    int coarse_half;       /* FP32 cycle: 1 = coarse links in FP16; 2 = also BF16 vectors in the
                              batched coarsest CG (N >= 16), half its vector traffic */
    int coarse_solver;     /* coarsest level: 0 CG (coarse_maxit, coarse_tol); 1 dense inverse
                              (formed once at setup; coarsest dimension <= 8192); 2 CG in the
                              FP64 cycles and, in the batched FP32 cycle with coarse_half 2, a
                              Chebyshev iteration of coarse_maxit steps on the coarsest
                              operator's spectral interval (Lanczos at setup): no reductions */
    int use_graph;         /* 1: replay the cycle from a captured hipGraph */
    int chiral;            /* 1: chiral aggregation at every level (k[l] = 2 x null vectors) */
    int null_prec;         /* near-null MR sweeps: 0 FP64 (the reference's); 1 FP32 on the FP32
                              operator, vectors and orthogonalisation kept in FP64 */
} smg_mg_params;
/* build_mg_hierarchy, src/multigrid.cpp:595-687, with every level's null-space start
 * vectors supplied (k[l] vectors of level-l dimension, k[l]/2 with chiral
 * aggregation; level 0 first), geometric aggregation at every level and 9-point
 * Galerkin operators */
int smg_mg_setup(smg_mg** out, smg_dirac* D, const smg_mg_params* prm, const double* starts);
int smg_mg_destroy(smg_mg* mg);
/* MGHierarchy::set_symmetric, src/multigrid.cpp:146-152 (power-iteration start
 * vectors for every non-coarsest level, concatenated) */
int smg_mg_set_symmetric(smg_mg* mg, double damping, const double* starts);
/* MGHierarchy::precondition, src/multigrid.cpp:155-157; the cycle runs in prec */
int smg_mg_precondition(smg_mg* mg, int prec, const double* b, double* x);
int smg_mg_precondition_dev(smg_mg* mg, int prec, const void* b, void* x, void* stream);
int smg_mg_level_dim(smg_mg* mg, int level, size_t* n);
int smg_mg_null_vecs(smg_mg* mg, int level, double* out);
int smg_mg_block_vecs(smg_mg* mg, int level, double* out);
int smg_mg_coarse_links(smg_mg* mg, int level, double* out);
int smg_mg_omega(smg_mg* mg, int level, double* out);
/* change the cycle's smoothing steps, coarsest-solve budget and graph replay after setup */
int smg_mg_set_cycle(smg_mg* mg, int pre_smooth, int post_smooth, int coarse_maxit, double coarse_tol,
                     int use_graph);

/* ------------------------------------------------------------------ Krylov solvers */
typedef struct {
    int iterations;
    double final_residual; /* the solver's own residual, as the reference reports it */
    double true_residual;  /* |b - A x|/|b| recomputed in FP64 on the device */
    int restarts;          /* outer FP64 refinements (mixed precision) */
} smg_solve_info;
/* cg_solve / cg_solve_x0 / cg_solve_precond / cg_solve_x0_precond, src/solvers.cpp:16-96
 * (x0 and mg may be NULL); op = SMG_OP_DDAGD or SMG_OP_SCHUR (mg must be NULL) */
int smg_cg_solve(smg_dirac* D, int op, smg_mg* mg, const double* b, const double* x0, double* x, int maxit,
                 double tol, smg_solve_info* info);
/* fcg_solve / fcg_solve_mg (truncated GCR), src/solvers.cpp:223-300, 803-811.
 * prec = SMG_F64: FP64 GCR(trunc) with the FP64 cycle.  prec = SMG_F32: mixed
 * precision -- FP64 iterative refinement on the FP64 true residual around FP32
 * GCR(trunc) + FP32 (optionally FP16-link) cycle, each inner solve to inner_tol.
 * ref_projection = 1 uses the reference's conjugated projection coefficient
 * (src/solvers.cpp:261); 0 the textbook GCR coefficient. */
int smg_gcr_solve(smg_dirac* D, smg_mg* mg, int prec, const double* b, double* x, int maxit, double tol, int trunc,
                  int ref_projection, double inner_tol, smg_solve_info* info);
int smg_gcr_solve_dev(smg_dirac* D, smg_mg* mg, int prec, const void* b, void* x, int maxit, double tol, int trunc,
                      double inner_tol, smg_solve_info* info, void* stream);
/* fgmres_solve_generic / fgmres_solve_mg with n_ritz_harvest = 0, src/solvers.cpp:555-800 */
int smg_fgmres_solve(smg_dirac* D, smg_mg* mg, const double* b, double* x, int restart, int maxit, double tol,
                     smg_solve_info* info);
/* fgmres_solve_generic / fgmres_solve_mg with n_ritz_harvest = n_ritz >= 0,
 * src/solvers.cpp:555-800 (harvest and intra-solve deflation :601-611, :693-751; the
 * FGMRESResult::ritz_pairs of src/solvers.h:18-23): *n_pairs (<= 2 n_ritz) Ritz pairs,
 * values in ritz_vals[2 n_ritz], vectors one after the other in ritz_vecs (2 n_ritz
 * vectors of ndof complex).  mg may be null (identity preconditioner). */
int smg_fgmres_solve_ritz(smg_dirac* D, smg_mg* mg, const double* b, double* x, int restart, int maxit, double tol,
                          int n_ritz, smg_solve_info* info, int* n_pairs, double* ritz_vals, double* ritz_vecs);

/* ------------------------------------------------------------------ caller-supplied operators
 * The reference's plug-in types: OpApply = std::function<void(const Vec& src, Vec& dst)>
 * (src/types.h:10) and the preconditioner std::function<Vec(const Vec&)>
 * (src/solvers.h:45-84).  An smg_op names exactly one of: a Dirac handle (its D^dag D,
 * on the device), a coarse handle (its 9-point operator, on the device) or a host
 * callback fn(ctx, src, dst) on n complex numbers (interleaved doubles; return 0, or
 * nonzero to abort the call with an error).  An smg_precond names a hierarchy (its
 * FP64 cycle), a host callback, or neither (the identity).  The Krylov vectors,
 * reductions and updates stay on the device; a callback's input is copied to the host
 * before each call and its output back after it.  Callbacks run on the calling thread,
 * one at a time. */
typedef int (*smg_op_fn)(void* ctx, const double* src, double* dst);
typedef struct {
    size_t n;           /* dimension (complex entries): required for fn, checked otherwise (0: any) */
    smg_op_fn fn;
    void* ctx;
    smg_dirac* dirac;
    smg_coarse* coarse;
} smg_op;
typedef struct {
    smg_op_fn fn;
    void* ctx;
    smg_mg* mg;
} smg_precond;
/* cg_solve / cg_solve_x0 / cg_solve_precond / cg_solve_x0_precond, src/solvers.cpp:16-96,
 * on any operator and preconditioner (x0 may be NULL; M may be NULL: none) */
int smg_cg_solve_op(const smg_op* A, const smg_precond* M, const double* b, const double* x0, double* x, int maxit,
                    double tol, smg_solve_info* info);
/* fcg_solve, src/solvers.cpp:223-300 (ref_projection = 1: the reference's coefficient) */
int smg_gcr_solve_op(const smg_op* A, const smg_precond* M, const double* b, double* x, int maxit, double tol,
                     int trunc, int ref_projection, smg_solve_info* info);
/* fgmres_solve_generic, src/solvers.cpp:555-800, with n_ritz >= 0 Ritz pairs harvested
 * (outputs as smg_fgmres_solve_ritz) */
int smg_fgmres_solve_op(const smg_op* A, const smg_precond* M, const double* b, double* x, int restart, int maxit,
                        double tol, int n_ritz, smg_solve_info* info, int* n_pairs, double* ritz_vals,
                        double* ritz_vecs);
/* ------------------------------------------------------------------ even-odd multigrid
 * EvenOddMG (src/eo_multigrid.h:26-287; the reference defines it and never calls it):
 * two levels on the Schur operator M^dag M of the odd half-lattice (vectors of V = L*L
 * complex numbers, odd half-index order, two spins).  Setup (EvenOddMG::build,
 * :229-254): k near-null vectors of M^dag M by the generic inverse iteration
 * (compute_near_null_space_generic, null_iters outer steps of 5 MR sweeps) from the
 * caller's k start vectors (starts: k x V, the reference's rng draws), the
 * aggregation of the odd sites of bx x by blocks (EvenOddProlongator; bx even,
 * bx and by dividing L), the dense coarse operator P^dag M^dag M P (EvenOddCoarseOp,
 * column by column) and Richardson's omega = 0.8 / lambda_max (estimate_lambda_max,
 * 30 power iterations from lambda_start).  precondition (:257-286): Richardson
 * pre-smoothing from zero, the residual restricted, the dense coarse solve, prolonged
 * and added, Richardson post-smoothing; op is the smoothing operator (NULL: the
 * setup's M^dag M; or a callback on V complex numbers, the reference's A_MdagM). */
int smg_eo_mg_setup(smg_eo_mg** out, smg_dirac* D, int bx, int by, int k, int pre_smooth, int post_smooth,
                    int null_iters, const double* starts, const double* lambda_start);
int smg_eo_mg_destroy(smg_eo_mg* M);
/* richardson_omega, the lambda_max estimate and the coarse dimension (any may be NULL) */
int smg_eo_mg_info(smg_eo_mg* M, double* omega, double* lambda_max, int* coarse_dim);
/* the dense coarse operator, row-major dim x dim complex (EvenOddCoarseOp::mat, transposed) */
int smg_eo_mg_coarse_matrix(smg_eo_mg* M, double* out);
int smg_eo_mg_precondition(smg_eo_mg* M, const smg_op* op, const double* b, double* x);
int smg_eo_mg_precondition_dev(smg_eo_mg* M, const void* b, void* x, void* stream);

/* mr_smooth_op / richardson_smooth_op, src/smoother.cpp:26-73 */
int smg_smooth_op(const smg_op* A, int kind, double* x, const double* b, int niter, double omega);
/* estimate_lambda_max, src/smoother.cpp:43-59, from the start vector v0 */
int smg_lambda_max_op(const smg_op* A, const double* v0, int niter, double* out);
/* SparseCoarseOp::build(P, A, fine_dim), src/coarse_op.h:241-279, for any operator A */
int smg_coarse_create_galerkin_op(smg_coarse** out, const smg_op* A, smg_prolongator* P);

/* ------------------------------------------------------------------ eigensolvers and deflation */
/* random_vec (src/linalg.cpp:44-49) from std::mt19937(seed): n complex numbers of
 * unit-variance normal parts, in the order the reference's g++ build draws them.
 * With seed 42 this is the default start vector of trlm_eigensolver
 * (src/eigensolver_core.cpp:720-727).  Host only. */
int smg_random_vec(int n, unsigned seed, double* out);
/* lanczos_eigen (src/eigensolver_core.cpp:103-265; declared src/eigensolver.h:30):
 * every eigenpair of the n x n Hermitian matrix A, complex, in the reference's
 * column layout A[j*n + i] = A_ij (A_cols[j][i]); evals ascending; evecs[row*n + col]
 * holds component row of eigenvector col (the reference's evecs[col][row]).  Host only. */
int smg_dense_eigh(int n, const double* A, double* evals, double* evecs);
/* trlm_eigensolver with poly_deg = 0 (src/eigensolver_core.cpp:691-1057; declared
 * src/eigensolver.h:68-73) on D^dag D (op SMG_OP_DDAGD) or M^dag M (SMG_OP_SCHUR):
 * the n_ev smallest eigenpairs, evecs[n_ev][n].  n_kr <= 0 takes the reference's
 * default Krylov size.  start: the start vector (the reference's default is
 * smg_random_vec(n, 42)).  info (may be NULL) reports a run that stopped at
 * max_restarts unconverged, as the reference's TRLMResult does. */
typedef struct {
    int iterations; /* operator applications (TRLMResult::iterations) */
    int restarts;   /* TRLMResult::num_restarts */
    int converged;  /* TRLMResult::converged */
} smg_trlm_info;
int smg_dirac_trlm(smg_dirac* D, int op, int n_ev, int n_kr, int max_restarts, double tol, const double* start,
                   double* evals, double* evecs, smg_trlm_info* info);
/* trlm_eigensolver as smg_dirac_trlm, on the coarse operator (SparseCoarseOp::as_op,
 * src/coarse_op.h:299-301) over its dimension; nothing is stored in C */
int smg_coarse_trlm(smg_coarse* C, int n_ev, int n_kr, int max_restarts, double tol, const double* start,
                    double* evals, double* evecs, smg_trlm_info* info);
/* SparseCoarseOp::setup_deflation with solver "trlm" (src/coarse_op.cpp:10-54;
 * declared src/coarse_op.h:480-483): TRLM on A_c, the pairs kept in C; every later
 * smg_coarse_solve then starts from the deflated guess x0 = sum_i (v_i^dag b / l_i) v_i
 * (SparseCoarseOp::solve, src/coarse_op.h:492-547).  start: as for smg_dirac_trlm
 * over the coarse dimension. */
int smg_coarse_setup_deflation(smg_coarse* C, int n_ev, int n_kr, int max_restarts, double tol, const double* start,
                               smg_trlm_info* info);
/* SparseCoarseOp::defl_vals / defl_vecs (src/coarse_op.h:169-170): read (evals,
 * evecs[ndefl][n] may be NULL) or replace (ndefl = 0 clears) the deflation space */
int smg_coarse_deflation(smg_coarse* C, int* ndefl, double* evals, double* evecs);
int smg_coarse_set_deflation(smg_coarse* C, int ndefl, const double* evals, const double* evecs);
/* lobpcg_update (src/eigensolver_core.cpp:279-533; declared src/eigensolver.h:39-44)
 * on D^dag D from the k start vectors X0[k][n] (1 <= k <= 32), preconditioned by the
 * FP64 MG cycle (mg) or not (mg = NULL, the reference's identity): evals[k] (the final
 * Rayleigh quotients, as the reference returns them) and X[k][n].  The k vectors run
 * as one block of multi-right-hand-side launches. */
int smg_dirac_lobpcg(smg_dirac* D, smg_mg* mg, int k, const double* X0, int max_iter, double tol, double* evals,
                     double* X, int* iterations);
/* smg_dirac_lobpcg with exact_ax = 1: the re-orthonormalisation of X keeps AX = A X
 * (AX_i -= <X_j, X_i> AX_j) where the reference subtracts <X_j, X_i> X_j
 * (src/eigensolver_core.cpp:514-517), which leaves AX inconsistent and the warm-started
 * Rayleigh quotients non-monotone; exact_ax = 0 is smg_dirac_lobpcg.  No reference
 * counterpart. */
int smg_dirac_lobpcg_ex(smg_dirac* D, smg_mg* mg, int k, const double* X0, int max_iter, double tol, int exact_ax,
                        double* evals, double* X, int* iterations);
/* lobpcg_update on the coarse operator with the identity preconditioner */
int smg_coarse_lobpcg(smg_coarse* C, int k, const double* X0, int max_iter, double tol, double* evals, double* X,
                      int* iterations);
/* MGHierarchy::update_coarsest_deflation (src/multigrid.cpp:111-141): lobpcg_iters
 * LOBPCG iterations (tol 1e-10) on the coarsest operator, warm-started from the
 * stored space when it holds n_defl vectors, else from random_vec(dim) drawn n_defl
 * times from mt19937(42); n_defl <= 32.  smg_mg_coarsest_deflation reads the space
 * (evals[ndefl], evecs[ndefl][dim]; either may be NULL). */
int smg_mg_update_coarsest_deflation(smg_mg* mg, int n_defl, int lobpcg_iters);
/* MGHierarchy::setup_sparse_coarse (src/multigrid.cpp:754-794; declared
 * src/multigrid.h:73-80) with solver "trlm": TRLM deflation (n_defl vectors, n_kr =
 * 2 n_defl + 10, tol 1e-10, from start) of the level-1 Galerkin operator, then the FP64
 * cycle's coarsest solve is its deflated CG (SparseCoarseOp::solve, to cg_tol in at most
 * max_cg_iter steps).  The reference wires the level-1 operator as the coarsest solve of
 * any hierarchy, which only fits two levels: other depths are refused.  The FP32 cycle
 * keeps its own coarsest solver. */
int smg_mg_setup_sparse_coarse(smg_mg* mg, int n_defl, double cg_tol, int max_cg_iter, const double* start);
/* levels[0].Ac.build(D, P0) then MGHierarchy::rebuild_deeper_levels (src/multigrid.cpp:
 * 163-194): every level's Galerkin operator re-projected, with the existing
 * prolongators, for the current gauge field of the hierarchy's operator (after
 * smg_dirac_set_gauge or an HMC update); the dense coarsest inverse follows, the
 * deflation spaces are kept as they are (as in the reference) */
int smg_mg_rebuild(smg_mg* mg);
/* No reference counterpart (an extension; the reference's warm start, build_full_mg
 * with warm_start in src/mg_builder.cpp:3-31, restarts the fine level only with 5 null
 * iterations and redraws the deeper levels): every level's near-null vectors iterated
 * null_iters more times from their current values on the current gauge field, then
 * the aggregations, the Galerkin operators and the dense coarsest inverse rebuilt;
 * the symmetric-cycle weights and deflation spaces are kept */
int smg_mg_refresh(smg_mg* mg, int null_iters);
int smg_mg_coarsest_deflation(smg_mg* mg, int* ndefl, double* evals, double* evecs);
/* cg_solve_deflated (src/solvers.cpp:111-213; declared src/solvers.h:50-56):
 * CG from x0 = U E^-1 U^dag b, E = U^dag A U, with the preconditioned residual
 * projected, z = M^-1 r - U E^-1 (AU)^dag M^-1 r; M = the FP64 MG cycle or none
 * (mg = NULL); op as smg_cg_solve.  defl_vecs[ndefl][n]; the reference's defl_vals
 * are not read by the solver and are not taken. */
int smg_cg_solve_deflated(smg_dirac* D, int op, smg_mg* mg, const double* b, int ndefl, const double* defl_vecs,
                          double* x, int maxit, double tol, smg_solve_info* info);

/* ------------------------------------------------------------------ many right-hand sides
 * N right-hand sides per launch, N a power of two (<= 32 for the solver, <= 64
 * for the cycle, the stencil and the coarse apply).  Host and device arrays
 * of the calls below hold the N vectors one after the other (each a reference
 * Vec); internally they run in the batched layout v[i*N + j] (RHS innermost) so
 * that every kernel of the cycle and solver carries all N.  No reference
 * counterpart: the reference solves one right-hand side at a time. */
/* the stencil on the batched layout (device pointers): src/dst[(2*s + spin)*N + j] */
int smg_dirac_apply_batched_dev(smg_dirac* D, int kind, int prec, const void* src, void* dst, int nrhs,
                                void* stream);
/* SparseCoarseOp::apply_to (src/coarse_op.h:282-297) on N right-hand sides at
 * once: host arrays of N coarse vectors one after the other; links read in
 * precision prec (SMG_F64, SMG_F32, or SMG_F16 = FP16 links with FP32
 * arithmetic).  For k in {8, 16, 32} and N in {16, 32, 64} the product runs on
 * the matrix cores (FP16 links: bf16 MFMA with the FP32 sources split in two
 * bf16 terms).  SMG_BF16V (host entry only; k in {8, 16, 32}, N in {16, 32, 64}):
 * FP16 links with the sources and results rounded to BF16, the storage of the
 * batched coarsest CG of a hierarchy with coarse_half = 2. */
int smg_coarse_apply_batched(smg_coarse* C, int prec, int nrhs, const double* src, double* dst);
/* the same on device arrays in the batched layout src/dst[(b*k + kk)*N + j]
 * (complex128 for SMG_F64, complex64 otherwise) */
int smg_coarse_apply_batched_dev(smg_coarse* C, int prec, int nrhs, const void* src, void* dst, void* stream);
/* one multigrid cycle applied to N right-hand sides at once (host arrays) */
int smg_mg_precondition_batched(smg_mg* mg, int prec, int nrhs, const double* B, double* X);
/* The coarsest-level solve of the batched FP32 cycle on nrhs right-hand sides (host
 * arrays [nrhs][coarsest dimension], complex128; computed in the cycle's precisions):
 * SparseCoarseOp::solve, src/coarse_op.h:492-547, as CG (coarse_solver 0; BF16 vectors
 * with coarse_half 2), or the dense inverse (1), or (2, with coarse_half 2 and nrhs >= 16)
 * the Chebyshev iteration, which has no reference counterpart */
int smg_mg_coarsest_batched(smg_mg* mg, int nrhs, const double* B, double* X);
/* the spectral interval [lo, hi] of a coarse_solver 2 hierarchy's Chebyshev iteration:
 * the extreme Ritz values of 60 Lanczos steps on the FP64 coarsest operator, hi
 * widened by 10% (no reference counterpart) */
int smg_mg_cheb_interval(smg_mg* mg, double* lo, double* hi);
/* replace that interval (0 < lo < hi) until the next smg_mg_rebuild / smg_mg_refresh
 * estimates it again: a lo above the smallest eigenvalue trades accuracy on the lowest
 * modes for a faster fall of the rest (no reference counterpart) */
int smg_mg_set_cheb_interval(smg_mg* mg, double lo, double hi);
/* Chebyshev smoothing in the batched cycles, FP32 and FP64 (ratio > 0; 0 restores MR):
 * every non-coarsest level smooths with pre_smooth / post_smooth Chebyshev steps on
 * [hi / ratio, hi], hi = 1.1 x a 40-step power-iteration estimate of the level's largest
 * eigenvalue (re-estimated by smg_mg_rebuild and smg_mg_refresh); the
 * single-right-hand-side cycles keep MR (src/smoother.cpp:9-40).  fused = 1: each step is
 * one launch of the level's operator with the vector updates in its epilogue where the
 * operator has that form (the fine D^dag D with L a multiple of 16 and nrhs a multiple
 * of 16 in FP32, 8 in FP64; the FP16-link coarse operators with k in {8, 16, 32} in the
 * FP32 cycle), else the apply and separate passes; fused = 0: always the latter.  hi (optional, n_levels - 1 entries) receives
 * each level's upper end.  No reference counterpart. */
int smg_mg_set_cheb_smoother(smg_mg* mg, double ratio, int fused, double* hi);
/* the hierarchy's level-`level` (>= 1) coarse operator on nrhs device vectors in the
 * batched layout (SparseCoarseOp::apply_to, src/coarse_op.h:282-297, as the cycle runs
 * it): prec SMG_F64 / SMG_F32 (links of that precision), SMG_F16 (FP16 links, FP32
 * vectors) or SMG_BF16V (FP16 links, BF16 vectors: one dword per complex entry); with
 * dots = 1 the launch also reduces {<dst, src>, |dst|^2} per right-hand side, as the
 * smoother and the coarse CG do (results discarded); with dots = 2 (SMG_BF16V) it is one
 * step of the Chebyshev coarsest solve (src = the direction, dst = the next direction,
 * the residual and iterate in library scratch; no reference counterpart) */
int smg_mg_coarse_apply_dev(smg_mg* mg, int level, int prec, int nrhs, const void* src, void* dst, int dots,
                            void* stream);
/* smg_gcr_solve for N right-hand sides in lockstep (textbook GCR projection);
 * true_res (optional, N doubles) receives each RHS's FP64 true relative residual */
int smg_gcr_solve_batched(smg_dirac* D, smg_mg* mg, int prec, int nrhs, const double* B, double* X, int maxit,
                          double tol, int trunc, double inner_tol, double* true_res, smg_solve_info* info);
int smg_gcr_solve_batched_dev(smg_dirac* D, smg_mg* mg, int prec, int nrhs, const void* B, void* X, int maxit,
                              double tol, int trunc, double inner_tol, double* true_res, smg_solve_info* info,
                              void* stream);

/* ------------------------------------------------------------------ forces (F = -dS/dtheta) */
/* forward half of fermion_force, src/hmc_forces.cpp:134-139: chi = D X */
int smg_force_forward(smg_dirac* D, const double* X, double* chi);
/* backward half: fermion_force_bilinear (hopping + clover), src/hmc_forces.cpp:141-495 */
int smg_force_backward(smg_dirac* D, const double* chi, const double* X, double* F);
/* fermion_force, src/hmc_forces.cpp:134-139: forward + backward with chi kept on the device */
int smg_fermion_force(smg_dirac* D, const double* X, double* F);
int smg_fermion_force_dev(smg_dirac* D, const void* X, void* F, void* stream);
/* hopping_force_bilinear, src/hmc_forces.cpp:540-569 */
int smg_hopping_force(smg_dirac* D, const double* chi, const double* X, double* F);
/* logdet_ee_force, src/hmc_forces.cpp:578-637 */
int smg_logdet_force(smg_dirac* D, double* F);
/* eo_fermion_force, src/hmc_forces.cpp:830-867 (x_o, y_o half-lattice) */
int smg_eo_fermion_force(smg_dirac* D, const double* x_o, const double* y_o, double* F);
/* fermion_action, src/hmc_forces.cpp:113-119: S = Re phi^dag (D^dag D)^-1 phi via CG;
 * X (optional) receives (D^dag D)^-1 phi, iters (optional) the CG iterations */
int smg_fermion_action(smg_dirac* D, const double* phi, int maxit, double tol, double* S, double* X, int* iters);

/* ------------------------------------------------------------------ HMC molecular dynamics
 * The gauge field evolves on the device inside the operator handle (its clover
 * term and FP32 links follow); momenta are 2V reals pi[mu*V + s]. */
typedef struct {
    double beta;       /* gauge coupling of the Wilson action */
    double tau;        /* trajectory length */
    int n_steps;       /* MD steps */
    int omelyan;       /* 0: leapfrog; 1: Omelyan 2MN with lambda = 0.1932 */
    int cg_maxiter;    /* CG on D^dag D for every force and action, from x0 = 0 */
    double cg_tol;
    int force_accept;  /* always accept (thermalisation) */
    int use_eo;        /* 1: even-odd trajectory (Schur operator, log det D_ee; leapfrog), even L */
} smg_hmc_params;      /* HMCParams, src/hmc.h:107-119 (no preconditioner) */
typedef struct {
    int accepted;
    double dH, dKE, dSG, dSF;
    int total_cg_iters;
    double dLD;        /* log-det change (even-odd with clover, else 0) */
    double max_true_residual;   /* no reference counterpart: the largest FP64 true residual
                                   |phi - A X|/|phi| of the trajectory's solves (a solve that stopped
                                   at its iteration budget shows here) */
} smg_hmc_result;      /* HMCResult, src/hmc.h:120-128 */
/* the operator's current gauge field (host, 2V complex) */
int smg_dirac_get_gauge(smg_dirac* D, double* U);
/* plain_leapfrog_evolve, src/hmc_trajectory.cpp:243-409 (no preconditioner, no
 * tracking): the operator's gauge field and the momenta pi (host, in/out) evolved
 * over tau with the pseudofermion phi (host, 2V complex) fixed.  use_eo: the
 * even-odd trajectory's leapfrog (src/hmc_trajectory.cpp:135-170), phi = phi_o
 * (V complex, odd sites in eo_index order) */
int smg_hmc_evolve(smg_dirac* D, const smg_hmc_params* p, double* pi, const double* phi, int* cg_iters);
/* the Hamiltonian of src/hmc_trajectory.cpp:87-103 (use_eo: 115-127) on the current
 * gauge field: out4 = {KE = 0.5 sum pi^2, S_g = -beta sum Re P, S_f = Re phi^dag
 * A^-1 phi with A = D^dag D (M^dag M), LD = -2 log det D_ee (even-odd with clover, else 0)} */
int smg_hmc_hamiltonian(smg_dirac* D, int use_eo, double beta, const double* pi, const double* phi, int cg_maxiter,
                        double cg_tol, double* out4, int* cg_iters);
/* hmc_trajectory, src/hmc_trajectory.cpp:38-236: momenta pi0 and the Gaussian eta
 * of phi = D^dag eta (use_eo: phi_o = M^dag eta_o, V complex) come from the caller
 * (drawn from its generator in that order); u is the uniform of the Metropolis test,
 * consulted only when dH >= 0 and force_accept == 0.  A rejected trajectory restores
 * the gauge field. */
int smg_hmc_trajectory(smg_dirac* D, const smg_hmc_params* p, const double* pi0, const double* eta, double u,
                       smg_hmc_result* res);

/* ------------------------------------------------------------------ multi-timescale MG HMC
 * hmc_trajectory_mg_multiscale (src/hmc_multiscale.cpp:340-888) without eigenspace
 * tracking, forecasting or prolongator perturbation: the outer integrator (leapfrog,
 * or Omelyan with lambda = 0.1932) kicks with F_g + F_f - F_low, F_f from the CG solve
 * of D^dag D X = phi (preconditioned by mg's FP64 cycle when mg is given); the inner
 * leapfrog, n_inner steps per outer half-step, kicks with F_low =
 * coarse_lowmode_force.  The deflation state (the reference's CoarseDeflState) is the
 * deflation space of an smg_coarse handle on P's coarse space (smg_coarse_setup_deflation
 * or smg_coarse_set_deflation); every defl_refresh-th inner step it is Rayleigh-Ritz
 * evolved on the Galerkin operator of P on the current gauge field and stays evolved,
 * as in the reference, also when the trajectory is rejected. */
typedef struct {
    double beta, tau;
    int n_outer, n_inner;
    int omelyan;       /* outer integrator: 0 leapfrog, 1 Omelyan */
    int inner_smooth;  /* MR sweeps around the low-mode solution (0: the raw projection) */
    int defl_refresh;  /* Rayleigh-Ritz refresh every N inner steps (0: never) */
    int cg_maxiter;
    double cg_tol;
    int force_accept;
} smg_ms_params;       /* MGMultiScaleParams, src/hmc.h:346-365 (the used fields) */
typedef struct {
    int accepted;
    double dH, dKE, dSG, dSF;
    int highmode_cg_iters;   /* all outer and Hamiltonian solves */
    int lowmode_force_evals; /* inner-force evaluations */
} smg_ms_result;             /* MGMultiScaleResult, src/hmc.h:367-376 (without the timings) */
/* coarse_lowmode_force (src/hmc_multiscale.cpp:237-276): F = -dS/dtheta of the low-mode
 * solution of D^dag D x = phi through P and defl's deflation space (smooth_iters MR
 * sweeps before and after, 0 for the raw projection); F[mu*V + s] */
int smg_coarse_lowmode_force(smg_dirac* D, smg_prolongator* P, smg_coarse* defl, const double* phi, int smooth_iters,
                             double* F);
/* evolve_coarse_deflation without forecast (src/hmc_multiscale.cpp:278-338): defl's
 * space rotated by rr_evolve (src/eigen_evolve.cpp:48-108) on A_new */
int smg_coarse_evolve_deflation(smg_coarse* defl, smg_coarse* A_new);
/* the molecular dynamics from momenta pi (2V reals, in/out) and phi (2V complex); the
 * gauge field of D evolves in place */
int smg_hmc_ms_evolve(smg_dirac* D, smg_mg* mg, smg_prolongator* P, smg_coarse* defl, const smg_ms_params* p,
                      double* pi, const double* phi, int* cg_iters, int* lowmode_evals);
/* the trajectory: pi0 and eta are the momenta and the Gaussian of the pseudofermion
 * phi = D^dag eta (the reference draws two pseudofermions and keeps the second; eta is
 * that second draw), u the Metropolis uniform */
int smg_hmc_ms_trajectory(smg_dirac* D, smg_mg* mg, smg_prolongator* P, smg_coarse* defl, const smg_ms_params* p,
                          const double* pi0, const double* eta, double u, smg_ms_result* res);

/* ------------------------------------------------------------------ nested integrator, gauge force on the fine scale
 * No reference counterpart: the reference's multi-timescale split
 * (src/hmc_multiscale.cpp:424-506) puts the gauge force on the outer scale; this one
 * puts the fermion force alone on the outer scale (n_outer steps over tau, the leapfrog
 * or Omelyan sequences of plain_leapfrog_evolve, src/hmc_trajectory.cpp:243-409) and
 * makes each outer position update Q(h) n_gauge steps of h / n_gauge of the inner
 * integrator, which kicks with the gauge force alone.  The solves of D^dag D X = phi
 * (every outer kick and both Hamiltonians) are FP64 CG, preconditioned by mg's FP64
 * cycle when mg is given (solver = SMG_F64), or mixed-precision GCR(10) with mg's FP32
 * cycle inside FP64 refinement on the FP64 true residual (solver = SMG_F32).  With
 * mg_refresh > 0 the hierarchy follows the evolving field: after every mg_refresh-th
 * outer position update it is refreshed as by smg_mg_refresh(mg, mg_refresh_iters), and
 * a rejected trajectory restores the near-null vectors it started from and rebuilds the
 * aggregations and Galerkin operators on the restored field. */
typedef struct {
    double beta, tau;
    int n_outer;        /* outer steps (fermion force) */
    int n_gauge;        /* inner steps per outer position update (gauge force) */
    int omelyan_outer;  /* 0 leapfrog, 1 Omelyan 2MN (lambda = 0.1932) */
    int omelyan_inner;
    int solver;         /* SMG_F64 or SMG_F32 (needs mg) */
    int cg_maxiter;
    double cg_tol;
    double inner_tol;   /* inner tolerance of the mixed-precision solver */
    int force_accept;
    int mg_refresh;     /* refresh the hierarchy every N outer position updates (0: never) */
    int mg_refresh_iters;
    int chrono_guess;   /* 1: every solve after the first starts from the previous solution */
} smg_nested_params;
/* the molecular dynamics from momenta pi (2V reals, in/out) and phi (2V complex); the
 * gauge field of D evolves in place; *solver_iters = the iterations of all solves */
int smg_hmc_nested_evolve(smg_dirac* D, smg_mg* mg, const smg_nested_params* p, double* pi, const double* phi,
                          int* solver_iters);
/* the trajectory: momenta pi0, phi = D^dag eta, H = KE + S_g + S_f before and after,
 * Metropolis with the caller's uniform u (accept if force_accept, dH < 0 or
 * u < exp(-dH)); a rejected trajectory restores the gauge field.  total_cg_iters = the
 * iterations of every solve, dLD = 0 */
int smg_hmc_nested_trajectory(smg_dirac* D, smg_mg* mg, const smg_nested_params* p, const double* pi0,
                              const double* eta, double u, smg_hmc_result* res);

/* ------------------------------------------------------------------ many lattices in one launch
 * No reference counterpart (the reference runs one lattice at a time): G Wilson-clover
 * lattices (1 <= G <= 1023) of different sizes (L[g] >= 2; links U of lattice g follow
 * those of g-1, each U[mu*V + s]; masses mass[g]; r, csw, mu_t shared) whose operators
 * and CG solves run as one launch per kernel for all of them.  Vectors are the lattices'
 * vectors one after the other (2 V_g complex each, psi[2s + spin]); smg_lattices_size
 * gives the total. */
int smg_lattices_create(smg_lattices** out, int G, const int* L, const double* U, const double* mass, double r,
                        double csw, double mu_t);
int smg_lattices_destroy(smg_lattices* H);
int smg_lattices_size(smg_lattices* H, size_t* n);
/* kind SMG_D, SMG_DDAG, SMG_DDAGD or SMG_G5D on every lattice (DiracOp::apply and
 * friends, src/dirac.h:93-152, per lattice); _dev: FP64 device arrays */
int smg_lattices_apply(smg_lattices* H, int kind, const double* src, double* dst);
int smg_lattices_apply_dev(smg_lattices* H, int kind, const void* src, void* dst, void* stream);
/* cg_solve on D^dag D (src/solvers.cpp:16-72) for every lattice at once from x0 = 0:
 * per-lattice scalars and stopping rule |r_g|/|b_g| < tol (a converged lattice takes no
 * further steps); iters[g] and the FP64 true residuals true_res[g] (either may be NULL) */
int smg_lattices_cg(smg_lattices* H, const double* b, double* x, int maxit, double tol, int* iters,
                    double* true_res);
/* Multigrid over the lattices of H (H must outlive it): every lattice's hierarchy is
 * the one smg_mg_setup builds for that lattice alone (build_mg_hierarchy,
 * src/multigrid.cpp:595-687; starts: every lattice's start vectors in smg_mg_setup's
 * layout, lattice after lattice), then held in one concatenated storage per level.
 * Every kernel of the cycle runs once for all lattices, with FP64 vectors and arithmetic;
 * coarse_half != 0 reads the coarse links in FP16 (the single-lattice FP16 copies),
 * else FP64.  k[l] must be a power of two in 2..64, aggregates square, the coarsest
 * solve CG (coarse_solver 0); use_graph replays the cycle from a captured hipGraph.  */
int smg_lattices_mg_setup(smg_lattices_mg** out, smg_lattices* H, const smg_mg_params* prm, const double* starts);
int smg_lattices_mg_destroy(smg_lattices_mg* M);
/* MGHierarchy::precondition (src/multigrid.cpp:155-157) on every lattice: one V-cycle
 * x = M b on the concatenated vectors; _dev: FP64 device arrays */
int smg_lattices_mg_precondition(smg_lattices_mg* M, const double* b, double* x);
int smg_lattices_mg_precondition_dev(smg_lattices_mg* M, const void* b, void* x, void* stream);
/* fcg_solve_mg (truncated GCR, src/solvers.cpp:223-300, 803-811) with the cycle on
 * every lattice at once from x0 = 0: one window of trunc <= 7 directions, per-lattice
 * projections and steps, a lattice stopping on |r_g|/|b_g| < tol; iters[g] and the
 * FP64 true residuals true_res[g] (either may be NULL); total = iterations run */
int smg_lattices_gcr(smg_lattices_mg* M, const double* b, double* x, int maxit, double tol, int trunc, int* iters,
                     double* true_res, int* total);
int smg_lattices_gcr_dev(smg_lattices_mg* M, const void* b, void* x, int maxit, double tol, int trunc, int* iters,
                         double* true_res, int* total, void* stream);

/* ------------------------------------------------------------------ U(1) heatbath
 * No counterpart in the reference, which thermalises by HMC: the thermalisation of the
 * quenched configurations (Wilson action S = -beta sum_x Re P(x)).  n_sweeps sweeps of
 * the checkerboard heatbath (per sweep: mu = 0, 1, each parity 0, 1), every link drawn
 * from exp(beta Re(U A)), A its staple sum, by the Best-Fisher von Mises sampler on
 * Philox4x32-10 numbers keyed by seed and counted by (link, sweep): the field depends
 * only on (U, beta, seed, sweep numbers).  Sweeps are numbered from first_sweep, so
 * n sweeps then m more equal n + m at once.  L even. */
int smg_dirac_heatbath(smg_dirac* D, double beta, int n_sweeps, unsigned long long seed, unsigned first_sweep);
/* the same on host links U[mu*V + s] (in/out) */
int smg_heatbath(int L, double* U, double beta, int n_sweeps, unsigned long long seed, unsigned first_sweep);

/* ------------------------------------------------------------------ gauge field files (host I/O)
 * Links are U[mu*V + s] (2V complex).  Host-only: no GPU is touched. */
/* GaugeField::save / load, src/gauge.h:38-57: int32 V, then U[0] and U[1] as complex
 * doubles (host byte order); reading fails if V != L*L */
int smg_gauge_write_raw(const char* path, int L, const double* U);
int smg_gauge_read_raw(const char* path, int L, double* U);
/* GaugeHeader, src/hmc.h:19-25 (32 bytes, host byte order) */
typedef struct {
    uint32_t magic; /* 0x53574E47 = "SWNG" */
    uint32_t L;
    double beta;
    double mass;
    double avg_plaq;
} smg_gauge_header;
/* save_gauge / load_gauge, src/hmc_forces.cpp:18-62: the header (avg_plaq summed in
 * site order) then the links; reading checks the magic number and L and renormalises
 * every link to |U| = 1; hdr (optional) receives the header */
int smg_gauge_write(const char* path, int L, const double* U, double beta, double mass);
int smg_gauge_read(const char* path, int L, double* U, smg_gauge_header* hdr);
/* therm_cfg_path, src/hmc_utils.cpp:7-11: "gauge_L<L>_b<beta as %f, 4 chars>_t<n>.bin" */
int smg_therm_cfg_path(int L, double beta, int n_therm, char* out, size_t cap);

/* ------------------------------------------------------------------ launch-shape autotuner
 * No reference counterpart.  A per-process table keeps, per kernel family and shape,
 * the variant and the reducing-launch workgroup count a timing search picked.  The
 * workgroup count fixes a reduction's summation order, so runs with the same table
 * give bit-identical results; shapes without an entry use the built-in defaults.
 * Families (shape words a..e, choice):
 *   SMG_TUNE_COARSE_BF16: the FP16-link matrix-core coarse apply; (nx, ny, k, nrhs,
 *     4*bf16_vectors + epilogue[0 none, 1 residual, 2 dots, 3 the CG's two-partner dots]);
 *     choice = herm + 2*g, herm = 1
 *     reads 5 of the 9 link blocks of a site (Hermitian Galerkin operators), g indexes the
 *     workgroup counts {128, 256, 512, 1024, 2048} of the launch with dot products;
 *   SMG_TUNE_DSLASH_DOTS: the batched stencil with dot products; (L, nrhs, bytes per
 *     real, 0, 0); choice = g;
 *   SMG_TUNE_DSLASH_STORE: the batched stencil's stores; (L, nrhs, bytes per real,
 *     epilogue, 0); choice 1 = non-temporal stores, 0 = ordinary (bit-identical results);
 *   SMG_TUNE_CG_UPDATE: the coarsest BF16-vector CG's vector update (r, x and p in one
 *     pass, no reduction); (coarse unknowns per RHS, nrhs, 0, 0, 0); choice = g;
 *   SMG_TUNE_TRANSFER: the batched restriction / prolongation; (nx, ny, 128*nd + k of the
 *     finer level, nrhs, 2*FP64 + prolong); choice 1 = a workgroup per aggregate with its
 *     blocks staged in LDS, 0 = one work-item per output (the same sums). */
#define SMG_TUNE_COARSE_BF16 0
#define SMG_TUNE_DSLASH_DOTS 1
#define SMG_TUNE_DSLASH_STORE 2
#define SMG_TUNE_CG_UPDATE 3
#define SMG_TUNE_TRANSFER 4
/* times the candidates of every kernel of mg's batched cycle for nrhs (16, 32 or 64)
 * right-hand sides and keeps the fastest; *entries = table entries added (may be NULL) */
int smg_tune_mg(smg_mg* mg, int nrhs, int* entries);
/* write / read the table (text, one "family a b c d e choice" line per entry; reading
 * adds to the table, a malformed line is an error and adds nothing) */
int smg_tune_save(const char* path);
int smg_tune_load(const char* path);
int smg_tune_clear(void);
/* the choice stored for (family, shape[5]) in *choice, -1 if none; set one by hand */
int smg_tune_get(int family, const int* shape, int* choice);
int smg_tune_set(int family, const int* shape, int choice);

#ifdef __cplusplus
}
#endif
#endif
