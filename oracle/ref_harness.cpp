// ref_harness.cpp -- extern "C" entry points into the reference's own code.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile from the reference
// sources where they lie (never copied) into oracle/_ref/libref.so; used to
// pin the oracle and to write the golden fixtures under tests/golden/.  The
// reference's FEAST paths are never referenced here, and the linker's section
// garbage collection drops them, so no FEAST or LAPACK is needed.
#include "dirac.h"
#include "prolongator.h"
#include "coarse_op.h"
#include "smoother.h"
#include "solvers.h"
#include "multigrid.h"
#include "hmc.h"
#include "eo_multigrid.h"
#include <cstring>
#include <iostream>
#include <deque>
#include <random>

#define API extern "C" __attribute__((visibility("default")))

namespace {
Vec to_vec(const double* p, size_t n) {
    Vec v(n);
    std::memcpy(v.data(), p, n * sizeof(cx));
    return v;
}
void from_vec(const Vec& v, double* p) { std::memcpy(p, v.data(), v.size() * sizeof(cx)); }
void load_gauge_arr(GaugeField& g, const double* U) {
    int V = g.lat.V;
    g.U[0] = to_vec(U, V);
    g.U[1] = to_vec(U + 2 * V, V);
}
struct Ctx {
    Lattice lat;
    GaugeField g;
    DiracOp D;
    Ctx(int L, const double* U, double m, double r, double csw, double mu)
        : lat(L), g(lat), D(lat, (load_gauge_arr(g, U), g), m, r, csw, mu) {}
};
void force_out(const std::array<RVec, 2>& F, double* out, int V) {
    std::memcpy(out, F[0].data(), V * sizeof(double));
    std::memcpy(out + V, F[1].data(), V * sizeof(double));
}
}  // namespace

API void ref_random_vec(int n, unsigned seed, double* out) {
    std::mt19937 rng(seed);
    from_vec(random_vec(n, rng), out);
}

API void ref_eo_maps(int L, int* even, int* odd, int* eoi) {
    Lattice lat(L);
    std::memcpy(even, lat.even_sites.data(), lat.V_half * sizeof(int));
    std::memcpy(odd, lat.odd_sites.data(), lat.V_half * sizeof(int));
    std::memcpy(eoi, lat.eo_index.data(), lat.V * sizeof(int));
}

API void ref_clover(int L, const double* U, double* F01) {
    Ctx c(L, U, 0.1, 1.0, 1.0, 0.0);
    std::memcpy(F01, c.D.clover_field.data(), c.lat.V * sizeof(double));
}

API double ref_avg_plaq(int L, const double* U) {
    Lattice lat(L); GaugeField g(lat); load_gauge_arr(g, U);
    return g.avg_plaq();
}

// kind: 0 D, 1 D^dag, 2 D^dag D, 3 g5 D
API void ref_dirac_apply(int L, const double* U, double m, double r, double csw, double mu,
                         int kind, const double* src, double* dst) {
    Ctx c(L, U, m, r, csw, mu);
    Vec s = to_vec(src, c.lat.ndof), d(c.lat.ndof);
    if (kind == 0) c.D.apply(s, d);
    else if (kind == 1) c.D.apply_dag(s, d);
    else if (kind == 2) c.D.apply_DdagD(s, d);
    else c.D.apply_g5D(s, d);
    from_vec(d, dst);
}

API void ref_dirac_delta(int L, const double* U, double m, double r, double csw, double mu,
                         const double* src, double* dst, const double* pi, double dt) {
    Ctx c(L, U, m, r, csw, mu);
    Vec s = to_vec(src, c.lat.ndof), d(c.lat.ndof);
    std::array<RVec, 2> p;
    p[0].assign(pi, pi + c.lat.V);
    p[1].assign(pi + c.lat.V, pi + 2 * c.lat.V);
    c.D.apply_delta_D(s, d, p, dt);
    from_vec(d, dst);
}

// kind: 0 hop e->o, 1 hop o->e, 2 ee_inv, 3 oo_inv, 4 oo, 5 M, 6 M^dag, 7 M^dag M
API void ref_eo_apply(int L, const double* U, double m, double r, double csw, double mu,
                      int kind, const double* src, double* dst) {
    Ctx c(L, U, m, r, csw, mu);
    EvenOddDiracOp eo(c.D);
    int nh = 2 * c.lat.V_half;
    Vec s = to_vec(src, nh), d(nh);
    switch (kind) {
    case 0: eo.apply_eo(s, d); break;
    case 1: eo.apply_oe(s, d); break;
    case 2: eo.apply_ee_inv(s, d); break;
    case 3: eo.apply_oo_inv(s, d); break;
    case 4: eo.apply_oo(s, d); break;
    case 5: eo.apply_schur(s, d); break;
    case 6: eo.apply_schur_dag(s, d); break;
    default: eo.apply_schur_dag_schur(s, d); break;
    }
    from_vec(d, dst);
}

API void ref_eo_reconstruct(int L, const double* U, double m, double r, double csw, double mu,
                            const double* x_o, const double* phi, double* out) {
    Ctx c(L, U, m, r, csw, mu);
    EvenOddDiracOp eo(c.D);
    Vec xo = to_vec(x_o, 2 * c.lat.V_half);
    Vec res = phi ? eo.reconstruct(xo, to_vec(phi, c.lat.ndof)) : eo.reconstruct_full(xo);
    from_vec(res, out);
}

API double ref_eo_logdet(int L, const double* U, double m, double r, double csw, double mu) {
    Ctx c(L, U, m, r, csw, mu);
    EvenOddDiracOp eo(c.D);
    return eo.log_det_ee();
}

// Geometric prolongator on the fine lattice: block vectors out [nblocks][k][2*b*b]
API void ref_prolong_build(int L, int b, int k, const double* nulls, double* bv) {
    Lattice lat(L);
    Prolongator P(lat, b, b, k);
    std::vector<Vec> nv(k);
    for (int i = 0; i < k; i++) nv[i] = to_vec(nulls + 2 * (size_t)i * lat.ndof, lat.ndof);
    P.build_from_vectors(nv);
    size_t bd = P.block_fine_dim();
    for (int bb = 0; bb < P.nblocks; bb++)
        for (int i = 0; i < k; i++)
            from_vec(P.block_vecs[bb][i], bv + 2 * ((size_t)bb * k + i) * bd);
}

// direction: 0 prolong (coarse -> fine), 1 restrict (fine -> coarse)
API void ref_prolong_apply(int L, int b, int k, const double* nulls, int direction,
                           const double* in, double* out) {
    Lattice lat(L);
    Prolongator P(lat, b, b, k);
    std::vector<Vec> nv(k);
    for (int i = 0; i < k; i++) nv[i] = to_vec(nulls + 2 * (size_t)i * lat.ndof, lat.ndof);
    P.build_from_vectors(nv);
    if (direction == 0) from_vec(P.prolong(to_vec(in, P.coarse_dim)), out);
    else from_vec(P.restrict_vec(to_vec(in, lat.ndof)), out);
}

// Linear-block prolongator of the coarse levels: block vectors [nblocks][k][block_k]
API void ref_coarse_prolong_build(int fine_dim, int block_k, int k, const double* nulls, double* bv) {
    CoarseProlongator P(fine_dim, block_k, k);
    std::vector<Vec> nv(k);
    for (int i = 0; i < k; i++) nv[i] = to_vec(nulls + 2 * (size_t)i * fine_dim, fine_dim);
    P.build_from_vectors(nv);
    for (int bb = 0; bb < P.nblocks; bb++)
        for (int i = 0; i < k; i++)
            from_vec(P.block_vecs[bb][i], bv + 2 * ((size_t)bb * k + i) * block_k);
}

// Galerkin A_c = P^dag (D^dag D) P: sparse 9-point links [nb][9][k*k] and/or dense columns
API void ref_galerkin(int L, const double* U, double m, double r, double csw, double mu,
                      int b, int k, const double* nulls, double* links, double* dense_cols) {
    Ctx c(L, U, m, r, csw, mu);
    Prolongator P(c.lat, b, b, k);
    std::vector<Vec> nv(k);
    for (int i = 0; i < k; i++) nv[i] = to_vec(nulls + 2 * (size_t)i * c.lat.ndof, c.lat.ndof);
    P.build_from_vectors(nv);
    if (links) {
        SparseCoarseOp S;
        S.build(P, c.D.as_DdagD_op(), c.lat.ndof);
        size_t kk = (size_t)k * k;
        for (int bb = 0; bb < S.nblocks; bb++)
            for (int d = 0; d < 9; d++)
                from_vec(S.link[bb][d], links + 2 * ((size_t)bb * 9 + d) * kk);
    }
    if (dense_cols) {
        CoarseOp A;
        A.build(c.D, P);
        for (int j = 0; j < A.dim; j++) from_vec(A.mat[j], dense_cols + 2 * (size_t)j * A.dim);
    }
}

API void ref_dense_solve(int dim, const double* cols, const double* b, double* x) {
    CoarseOp A;
    A.dim = dim;
    A.mat.resize(dim);
    for (int j = 0; j < dim; j++) A.mat[j] = to_vec(cols + 2 * (size_t)j * dim, dim);
    from_vec(A.solve(to_vec(b, dim)), x);
}

API void ref_mr_smooth(int L, const double* U, double m, double r, double csw, double mu,
                       double* x, const double* b, int niter) {
    Ctx c(L, U, m, r, csw, mu);
    Vec xv = to_vec(x, c.lat.ndof);
    mr_smooth(c.D, xv, to_vec(b, c.lat.ndof), niter);
    from_vec(xv, x);
}

API double ref_lambda_max(int L, const double* U, double m, double r, double csw, double mu, int niter) {
    Ctx c(L, U, m, r, csw, mu);
    return estimate_lambda_max(c.D.as_DdagD_op(), c.lat.ndof, niter);
}

// solver: 0 CG, 1 GCR (fcg_solve), 2 FGMRES; operator D^dag D; identity preconditioner
API int ref_krylov(int L, const double* U, double m, double r, double csw, double mu, int solver,
                   const double* rhs, int maxit, double tol, int trunc, double* x, double* res) {
    Ctx c(L, U, m, r, csw, mu);
    OpApply A = c.D.as_DdagD_op();
    Vec b = to_vec(rhs, c.lat.ndof);
    auto ident = [](const Vec& v) { return v; };
    if (solver == 0) {
        auto o = cg_solve(A, c.lat.ndof, b, maxit, tol);
        from_vec(o.solution, x); *res = o.final_residual; return o.iterations;
    } else if (solver == 1) {
        auto o = fcg_solve(A, c.lat.ndof, b, ident, maxit, tol, trunc);
        from_vec(o.solution, x); *res = o.final_residual; return o.iterations;
    }
    auto o = fgmres_solve_generic(A, c.lat.ndof, b, ident, trunc, maxit, tol, 0);
    from_vec(o.solution, x); *res = o.final_residual; return o.iterations;
}

// fgmres_solve_generic on D^dag D, identity preconditioner, with n_ritz_harvest =
// n_ritz: x, the residual, and the harvested pairs (*nr_out of them; vals, vecs
// [.][2V] with room for 2 n_ritz + restart)
API int ref_fgmres_ritz(int L, const double* U, double m, double r, double csw, double mu,
                        const double* rhs, int maxit, double tol, int restart, int n_ritz,
                        double* x, double* res, int* nr_out, double* vals, double* vecs) {
    Ctx c(L, U, m, r, csw, mu);
    OpApply A = c.D.as_DdagD_op();
    auto ident = [](const Vec& v) { return v; };
    auto o = fgmres_solve_generic(A, c.lat.ndof, to_vec(rhs, c.lat.ndof), ident, restart, maxit, tol, n_ritz);
    from_vec(o.solution, x);
    *res = o.final_residual;
    *nr_out = (int)o.ritz_pairs.size();
    for (size_t q = 0; q < o.ritz_pairs.size(); q++) {
        vals[q] = o.ritz_pairs[q].value;
        from_vec(o.ritz_pairs[q].vector, vecs + 2 * q * c.lat.ndof);
    }
    return o.iterations;
}

// Multigrid built from the reference's building blocks exactly as
// build_mg_hierarchy (src/multigrid.cpp:595-687) assembles them, but with every
// level's null-space start vectors supplied (the reference's warm_start).
// Levels >= 1 use CoarseProlongator with block_k = coarse_block.  Then either one
// preconditioner application (mode 0), or a solve: 1 = fcg_solve_mg,
// 2 = CG with the symmetric (Richardson) cycle, 3 = fgmres_solve_mg.
// sym_starts: power-iteration start vectors per non-coarsest level (mode 2).
API int ref_mg(int L, const double* U, double m, double r, double csw, double mu,
               int n_levels, int block, int k, int coarse_block, int null_iters,
               const double* starts, int pre, int post, int w_cycle, int mode,
               const double* b, double* x, int maxit, double tol, int trunc,
               const double* sym_starts, double* res_out) {
    Ctx c(L, U, m, r, csw, mu);
    std::mt19937 rng(1);
    std::vector<MGLevel> levels(n_levels);
    std::deque<Prolongator> gp;
    std::deque<CoarseProlongator> cp;
    std::deque<CoarseOp> acs;
    size_t off = 0;
    levels[0].op = c.D.as_DdagD_op();
    levels[0].dim = c.lat.ndof;
    levels[0].pre_smooth = pre; levels[0].post_smooth = post;
    std::vector<Vec> warm(k);
    for (int i = 0; i < k; i++) warm[i] = to_vec(starts + 2 * (off + (size_t)i * c.lat.ndof), c.lat.ndof);
    off += (size_t)k * c.lat.ndof;
    auto nv = compute_near_null_space(c.D, k, null_iters, rng, &warm);
    gp.emplace_back(c.lat, block, block, k);
    gp.back().build_from_vectors(nv);
    Prolongator& P0 = gp.back();
    levels[0].restrict_fn = [&P0](const Vec& v) { return P0.restrict_vec(v); };
    levels[0].prolong_fn = [&P0](const Vec& v) { return P0.prolong(v); };
    acs.emplace_back();
    acs.back().build(c.D, P0);
    levels[0].Ac = acs.back();
    int prev = P0.coarse_dim;
    for (int l = 1; l < n_levels - 1; l++) {
        CoarseOp& Aab = acs.back();
        OpApply op = [&Aab](const Vec& s, Vec& d) { d = Aab.apply(s); };
        levels[l].op = op; levels[l].dim = prev;
        levels[l].pre_smooth = pre; levels[l].post_smooth = post;
        int blk = coarse_block;
        while (prev % blk != 0 && blk > 1) blk--;
        std::vector<Vec> w(k);
        for (int i = 0; i < k; i++) w[i] = to_vec(starts + 2 * (off + (size_t)i * prev), prev);
        off += (size_t)k * prev;
        auto cn = compute_near_null_space_generic(op, prev, k, null_iters, rng, &w);
        cp.emplace_back(prev, blk, k);
        cp.back().build_from_vectors(cn);
        CoarseProlongator& Pc = cp.back();
        levels[l].restrict_fn = [&Pc](const Vec& v) { return Pc.restrict_vec(v); };
        levels[l].prolong_fn = [&Pc](const Vec& v) { return Pc.prolong(v); };
        acs.emplace_back();
        acs.back().build_generic(op, Pc, prev, Pc.coarse_dim);
        levels[l].Ac = acs.back();
        prev = Pc.coarse_dim;
    }
    {
        CoarseOp& Aab = acs.back();
        levels[n_levels - 1].op = [&Aab](const Vec& s, Vec& d) { d = Aab.apply(s); };
        levels[n_levels - 1].dim = prev;
        levels[n_levels - 1].Ac = Aab;
        levels[n_levels - 1].pre_smooth = 0;
        levels[n_levels - 1].post_smooth = 0;
    }
    if (mode == 2) {   // set_symmetric(0.8) with supplied power-iteration starts
        size_t so = 0;
        for (int l = 0; l < n_levels - 1; l++) {
            int n = levels[l].dim;
            Vec v = to_vec(sym_starts + 2 * so, n);
            so += n;
            double vn = norm(v); scale(v, 1.0 / vn);
            Vec Av(n); double lambda = 0;
            for (int i = 0; i < 30; i++) {
                levels[l].op(v, Av);
                lambda = norm(Av);
                if (lambda < 1e-30) break;
                scale(Av, 1.0 / lambda); v = Av;
            }
            levels[l].richardson_omega = 0.8 / lambda;
        }
    }
    bool wc = w_cycle != 0;
    auto pc = [&levels, wc](const Vec& v) { return mg_cycle(levels, 0, v, wc); };
    Vec bv = to_vec(b, c.lat.ndof);
    if (mode == 0) { from_vec(pc(bv), x); return 0; }
    if (mode == 1) {
        auto o = fcg_solve(levels[0].op, c.lat.ndof, bv, pc, maxit, tol, trunc);
        from_vec(o.solution, x); *res_out = o.final_residual; return o.iterations;
    }
    if (mode == 2) {
        auto o = cg_solve_precond(levels[0].op, c.lat.ndof, bv, pc, maxit, tol);
        from_vec(o.solution, x); *res_out = o.final_residual; return o.iterations;
    }
    auto o = fgmres_solve_generic(levels[0].op, c.lat.ndof, bv, pc, trunc, maxit, tol, 0);
    from_vec(o.solution, x); *res_out = o.final_residual; return o.iterations;
}

// forces (F = -dS/dtheta, layout F[mu*V + s])
API void ref_gauge_force(int L, const double* U, double beta, double* F) {
    Lattice lat(L); GaugeField g(lat); load_gauge_arr(g, U);
    std::array<RVec, 2> f; gauge_force(g, beta, f); force_out(f, F, lat.V);
}
API double ref_gauge_action(int L, const double* U, double beta) {
    Lattice lat(L); GaugeField g(lat); load_gauge_arr(g, U);
    return gauge_action(g, beta);
}
// kind: 0 fermion_force(X), 1 fermion_force_bilinear(chi, X), 2 hopping_force_bilinear(chi, X),
//       3 logdet_ee_force, 4 eo_fermion_force(x_o = X, y_o = chi)
API void ref_fermion_force(int L, const double* U, double m, double r, double csw, double mu,
                           int kind, const double* X, const double* chi, double* F) {
    Ctx c(L, U, m, r, csw, mu);
    std::array<RVec, 2> f;
    int n = c.lat.ndof, nh = 2 * c.lat.V_half;
    if (kind == 0) fermion_force(c.D, to_vec(X, n), f);
    else if (kind == 1) fermion_force_bilinear(c.D, to_vec(chi, n), to_vec(X, n), f);
    else if (kind == 2) hopping_force_bilinear(c.D, to_vec(chi, n), to_vec(X, n), f);
    else if (kind == 3) logdet_ee_force(c.D, f);
    else { EvenOddDiracOp eo(c.D); eo_fermion_force(c.D, eo, to_vec(X, nh), to_vec(chi, nh), f); }
    force_out(f, F, c.lat.V);
}

API double ref_fermion_action(int L, const double* U, double m, double r, double csw, double mu,
                              const double* phi, int maxit, double tol) {
    Ctx c(L, U, m, r, csw, mu);
    return fermion_action(c.D, to_vec(phi, c.lat.ndof), maxit, tol).action;
}

// plain_leapfrog_evolve (src/hmc_trajectory.cpp:243-409): the gauge field U
// (2V complex) and the momenta pi (2V reals, pi[mu*V + s]) are evolved in place
// with the pseudofermion phi fixed; no preconditioner, no tracking.
API int ref_leapfrog_evolve(int L, double* U, double mass, double r, double csw, double beta, double tau,
                            int n_steps, int omelyan, int cg_maxit, double cg_tol, double* pi, const double* phi) {
    Lattice lat(L);
    GaugeField g(lat);
    load_gauge_arr(g, U);
    MomentumField mom(lat);
    std::memcpy(mom.pi[0].data(), pi, lat.V * sizeof(double));
    std::memcpy(mom.pi[1].data(), pi + lat.V, lat.V * sizeof(double));
    HMCParams p{};
    p.beta = beta;
    p.tau = tau;
    p.n_steps = n_steps;
    p.cg_maxiter = cg_maxit;
    p.cg_tol = cg_tol;
    p.use_mg = false;
    p.c_sw = csw;
    p.omelyan = omelyan != 0;
    const int iters = plain_leapfrog_evolve(g, mom, lat, mass, r, to_vec(phi, lat.ndof), p, nullptr, nullptr);
    from_vec(g.U[0], U);
    from_vec(g.U[1], U + 2 * lat.V);
    std::memcpy(pi, mom.pi[0].data(), lat.V * sizeof(double));
    std::memcpy(pi + lat.V, mom.pi[1].data(), lat.V * sizeof(double));
    return iters;
}

// gauge files: GaugeField::save / load (src/gauge.h:38-57), save_gauge / load_gauge
// (src/hmc_forces.cpp:18-62)
API int ref_gauge_save_raw(int L, const double* U, const char* path) {
    Lattice lat(L);
    GaugeField g(lat);
    load_gauge_arr(g, U);
    return g.save(path) ? 0 : -1;
}
API int ref_gauge_load_raw(int L, double* U, const char* path) {
    Lattice lat(L);
    GaugeField g(lat);
    if (!g.load(path)) return -1;
    from_vec(g.U[0], U);
    from_vec(g.U[1], U + 2 * lat.V);
    return 0;
}
API int ref_gauge_save(int L, const double* U, double beta, double mass, const char* path) {
    Lattice lat(L);
    GaugeField g(lat);
    load_gauge_arr(g, U);
    return save_gauge(g, beta, mass, path) ? 0 : -1;
}
API int ref_gauge_load(int L, double* U, double* hdr3, const char* path) {
    Lattice lat(L);
    GaugeField g(lat);
    GaugeHeader h{};
    if (!load_gauge(g, h, path)) return -1;
    from_vec(g.U[0], U);
    from_vec(g.U[1], U + 2 * lat.V);
    hdr3[0] = h.beta;
    hdr3[1] = h.mass;
    hdr3[2] = h.avg_plaq;
    return 0;
}

// the draws of hmc_trajectory (src/hmc_trajectory.cpp:52-53, 62): momenta (pi[0] then
// pi[1]), then the n_eta complex Gaussians of the pseudofermion, from mt19937(seed)
API void ref_hmc_draws(int L, unsigned seed, double* pi, int n_eta, double* eta) {
    Lattice lat(L);
    std::mt19937 rng(seed);
    MomentumField mom(lat);
    mom.randomise(rng);
    std::memcpy(pi, mom.pi[0].data(), lat.V * sizeof(double));
    std::memcpy(pi + lat.V, mom.pi[1].data(), lat.V * sizeof(double));
    from_vec(random_vec(n_eta, rng), eta);
}
// hmc_trajectory (src/hmc_trajectory.cpp:38-236) with force_accept: the final links and
// {dH, dKE, dSG, dSF, dLD, cg iterations}
API void ref_hmc_trajectory(int L, double* U, double mass, double r, double csw, double beta, double tau,
                            int n_steps, int omelyan, int use_eo, int cg_maxit, double cg_tol, unsigned seed,
                            double* out6) {
    Lattice lat(L);
    GaugeField g(lat);
    load_gauge_arr(g, U);
    HMCParams p{};
    p.beta = beta;
    p.tau = tau;
    p.n_steps = n_steps;
    p.cg_maxiter = cg_maxit;
    p.cg_tol = cg_tol;
    p.use_mg = false;
    p.c_sw = csw;
    p.use_eo = use_eo != 0;
    p.omelyan = omelyan != 0;
    p.force_accept = true;
    std::mt19937 rng(seed);
    const HMCResult res = hmc_trajectory(g, lat, mass, r, p, rng, nullptr, nullptr);
    from_vec(g.U[0], U);
    from_vec(g.U[1], U + 2 * lat.V);
    out6[0] = res.dH;
    out6[1] = res.dKE;
    out6[2] = res.dSG;
    out6[3] = res.dSF;
    out6[4] = res.dLD;
    out6[5] = res.total_cg_iters;
}

// lanczos_eigen (src/eigensolver_core.cpp:103-265): A column-major (A[j*n + i]);
// evecs[row*n + col]
API void ref_lanczos_eigen(int n, const double* A, double* evals, double* evecs) {
    std::vector<Vec> cols(n);
    for (int j = 0; j < n; j++) cols[j] = to_vec(A + 2 * (size_t)j * n, n);
    RVec ev;
    std::vector<Vec> vecs;
    lanczos_eigen(cols, n, ev, vecs);
    std::memcpy(evals, ev.data(), n * sizeof(double));
    for (int row = 0; row < n; row++) from_vec(vecs[row], evecs + 2 * (size_t)row * n);
}
// trlm_eigensolver on D^dag D (default start vector: random_vec(n, mt19937(42)))
API int ref_trlm_dirac(int L, const double* U, double m, double r, double csw, double mu, int n_ev, int n_kr,
                       int max_restarts, double tol, double* evals, double* evecs) {
    Ctx c(L, U, m, r, csw, mu);
    TRLMResult res = trlm_eigensolver(c.D.as_DdagD_op(), c.lat.ndof, n_ev, n_kr, max_restarts, tol);
    for (int i = 0; i < n_ev; i++) {
        evals[i] = res.eigvals[i];
        from_vec(res.eigvecs[i], evecs + 2 * (size_t)i * c.lat.ndof);
    }
    return res.converged ? res.num_restarts : -1;
}
// SparseCoarseOp::setup_deflation (TRLM) then solve (src/coarse_op.cpp:10-54,
// src/coarse_op.h:492-547) on the Galerkin operator of b x b aggregates with k vectors
API int ref_coarse_deflated_solve(int L, const double* U, double m, double r, double csw, double mu, int b, int k,
                                  const double* nulls, int n_ev, int n_kr, int max_restarts, double tol,
                                  const double* rhs, int maxit, double cg_tol, double* evals, double* x) {
    Ctx c(L, U, m, r, csw, mu);
    Prolongator P(c.lat, b, b, k);
    std::vector<Vec> nv(k);
    for (int i = 0; i < k; i++) nv[i] = to_vec(nulls + 2 * (size_t)i * c.lat.ndof, c.lat.ndof);
    P.build_from_vectors(nv);
    SparseCoarseOp S;
    S.build(P, c.D.as_DdagD_op(), c.lat.ndof);
    // the TRLM branch of setup_deflation (src/coarse_op.cpp:39-52); the function itself
    // also holds the FEAST branch, which needs the Fortran library this image lacks
    if (n_kr <= 0) n_kr = std::min(2 * n_ev + 10, S.dim);
    if (n_kr > S.dim) n_kr = S.dim;
    if (n_kr < n_ev + 2) n_kr = std::min(n_ev + 6, S.dim);
    TRLMResult tr = trlm_eigensolver(S.as_op(), S.dim, n_ev, n_kr, max_restarts, tol);
    S.defl_vecs = std::move(tr.eigvecs);
    S.defl_vals = std::move(tr.eigvals);
    for (int i = 0; i < (int)S.defl_vals.size(); i++) evals[i] = S.defl_vals[i];
    from_vec(S.solve(to_vec(rhs, S.dim), maxit, cg_tol), x);
    return (int)S.defl_vals.size();
}
// cg_solve_deflated (src/solvers.cpp:111-213) on D^dag D with the identity
// preconditioner and the ndefl deflation vectors vecs[ndefl][ndof]
API int ref_cg_deflated(int L, const double* U, double m, double r, double csw, double mu, int ndefl,
                        const double* vecs, const double* rhs, int maxit, double tol, double* x, double* res) {
    Ctx c(L, U, m, r, csw, mu);
    std::vector<Vec> dv(ndefl);
    for (int i = 0; i < ndefl; i++) dv[i] = to_vec(vecs + 2 * (size_t)i * c.lat.ndof, c.lat.ndof);
    std::vector<double> dl(ndefl, 1.0);   // the eigenvalues are not read by the solver
    auto ident = [](const Vec& v) { return v; };
    CGResult o = cg_solve_deflated(c.D.as_DdagD_op(), c.lat.ndof, to_vec(rhs, c.lat.ndof), ident, dv, dl, maxit, tol);
    from_vec(o.solution, x);
    *res = o.final_residual;
    return o.iterations;
}
// lobpcg_update (src/eigensolver_core.cpp:279-533) on D^dag D, identity
// preconditioner, from k start vectors X0[k][ndof]; returns the iterations
API int ref_lobpcg_dirac(int L, const double* U, double m, double r, double csw, double mu, int k, const double* X0,
                         int max_iter, double tol, double* evals, double* X) {
    Ctx c(L, U, m, r, csw, mu);
    std::vector<Vec> x0(k);
    for (int i = 0; i < k; i++) x0[i] = to_vec(X0 + 2 * (size_t)i * c.lat.ndof, c.lat.ndof);
    auto ident = [](const Vec& v) { return v; };
    LOBPCGResult o = lobpcg_update(c.D.as_DdagD_op(), c.lat.ndof, k, x0, ident, max_iter, tol);
    for (int i = 0; i < k; i++) {
        evals[i] = o.eigvals[i];
        from_vec(o.eigvecs[i], X + 2 * (size_t)i * c.lat.ndof);
    }
    return o.iterations;
}
// hmc_trajectory_mg_multiscale (src/hmc_multiscale.cpp:340-888) with the identity as
// mg_precond, no tracking or forecast; the deflation state is the TRLM space of the
// b x b Galerkin operator of the given null vectors (SparseCoarseOp::setup_deflation,
// n_ev = n_defl).  out = {accepted, dH, dKE, dSG, dSF, highmode CG iterations, low-mode
// force evaluations}; U in/out; the deflation space before (vals0, vecs0) and after
// (vals1, vecs1) the trajectory.
API void ref_hmc_ms_trajectory(int L, double* U, double mass, double r, double csw, double beta, double tau,
                               int n_outer, int n_inner, int omelyan, int inner_smooth, int defl_refresh,
                               int cg_maxit, double cg_tol, int b, int k, const double* nulls, int n_defl,
                               unsigned seed, double* out, double* vals0, double* vecs0, double* vals1,
                               double* vecs1) {
    Lattice lat(L);
    GaugeField g(lat);
    load_gauge_arr(g, U);
    DiracOp D(lat, g, mass, r, csw, 0.0);
    Prolongator P(lat, b, b, k);
    std::vector<Vec> nv(k);
    for (int i = 0; i < k; i++) nv[i] = to_vec(nulls + 2 * (size_t)i * lat.ndof, lat.ndof);
    P.build_from_vectors(nv);
    SparseCoarseOp S;
    S.build(P, D.as_DdagD_op(), lat.ndof);
    int n_kr = std::min(2 * n_defl + 10, S.dim);
    TRLMResult tr = trlm_eigensolver(S.as_op(), S.dim, n_defl, n_kr, 100, 1e-10);
    CoarseDeflState cd;
    cd.eigvecs = tr.eigvecs;
    cd.eigvals = tr.eigvals;
    for (int i = 0; i < n_defl; i++) {
        vals0[i] = cd.eigvals[i];
        from_vec(cd.eigvecs[i], vecs0 + 2 * (size_t)i * S.dim);
    }
    MGMultiScaleParams p;
    p.beta = beta;
    p.tau = tau;
    p.n_outer = n_outer;
    p.n_inner = n_inner;
    p.cg_maxiter = cg_maxit;
    p.cg_tol = cg_tol;
    p.inner_smooth = inner_smooth;
    p.outer_type = omelyan ? OuterIntegrator::Omelyan : OuterIntegrator::Leapfrog;
    p.defl_refresh = defl_refresh;
    p.c_sw = csw;
    std::function<Vec(const Vec&)> ident = [](const Vec& v) { return v; };
    std::mt19937 rng(seed);
    MGMultiScaleResult res = hmc_trajectory_mg_multiscale(g, lat, mass, r, p, cd, P, ident, rng);
    out[0] = res.accepted ? 1.0 : 0.0;
    out[1] = res.dH;
    out[2] = res.dKE;
    out[3] = res.dSG;
    out[4] = res.dSF;
    out[5] = res.highmode_cg_iters;
    out[6] = res.lowmode_force_evals;
    for (int i = 0; i < n_defl; i++) {
        vals1[i] = cd.eigvals[i];
        from_vec(cd.eigvecs[i], vecs1 + 2 * (size_t)i * S.dim);
    }
    from_vec(g.U[0], U);
    from_vec(g.U[1], U + 2 * (size_t)lat.V);
}

// EvenOddMG (src/eo_multigrid.h:229-287): build on the Schur M^dag M with mt19937(seed),
// then one precondition of b; the Richardson weight, the coarse dimension and the
// dense coarse matrix (column j at mat + 2*j*dim) are returned.  The build's report
// line on std::cout is discarded.
API void ref_eo_mg(int L, const double* U, double m, double r, double csw, double mu, int bx, int by, int k,
                   int null_iters, int pre, int post, unsigned seed, const double* b, double* x, double* omega,
                   int* coarse_dim, double* mat) {
    Ctx c(L, U, m, r, csw, mu);
    EvenOddMG mg(c.lat, bx, by, k, pre, post);
    std::mt19937 rng(seed);
    std::streambuf* keep = std::cout.rdbuf(nullptr);
    mg.build(c.D, null_iters, rng);
    std::cout.rdbuf(keep);
    EvenOddDiracOp eoD(c.D);
    OpApply A = [&eoD](const Vec& s, Vec& d) { eoD.apply_schur_dag_schur(s, d); };
    from_vec(mg.precondition(to_vec(b, 2 * c.lat.V_half), A), x);
    *omega = mg.richardson_omega;
    *coarse_dim = mg.Ac.dim;
    if (mat)
        for (int j = 0; j < mg.Ac.dim; j++) from_vec(mg.Ac.mat[j], mat + 2 * (size_t)j * mg.Ac.dim);
}
