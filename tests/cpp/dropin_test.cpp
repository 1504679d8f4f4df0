// A reference-style caller of the C++ drop-in API (include/schwinger_mg.hpp):
// the code below is written against the reference's own signatures
// (src/dirac.h, src/multigrid.h, src/solvers.h, src/hmc.h) and runs on the GPU
// through libsmg.so.  Every result is checked against the oracle (liboracle.so,
// the CPU restatement) on identical inputs; the program prints one line per check
// and exits non-zero on the first failure.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <random>

#include "schwinger_mg.hpp"

using namespace schwinger_mg;

// the oracle's C interface (oracle/oracle.h) with std::complex<double>, which has
// the layout of C99 double complex
extern "C" {
typedef struct orc_dirac orc_dirac;
typedef struct orc_op orc_op;
typedef struct orc_mg orc_mg;
typedef struct {
    int n_levels;
    int bx[4], by[4];
    int k[4];
    int null_iters;
    int pre_smooth, post_smooth;
    int w_cycle;
    int coarse_solver;
    int coarse_maxit;
    double coarse_tol;
    int colored_build;
    int chiral;
} orc_mg_params;
orc_dirac* orc_dirac_new(int L, const cx* U, double mass, double r, double csw, double mu_t);
void orc_dirac_free(orc_dirac* D);
void orc_dirac_apply(const orc_dirac* D, int kind, const cx* src, cx* dst);
void orc_eo_apply(const orc_dirac* D, int kind, const cx* src, cx* dst);
orc_op* orc_op_dirac(const orc_dirac* D, int kind);
void orc_op_free(orc_op* A);
orc_mg* orc_mg_new(const orc_op* fine, int nx, int ny, int nd, const orc_mg_params* prm, const cx* starts);
void orc_mg_free(orc_mg* mg);
void orc_mg_precondition(orc_mg* mg, const cx* b, cx* x);
int orc_cg(const orc_op* A, const cx* rhs, const cx* x0, orc_mg* mg, int maxit, double tol, cx* x, double* res);
void orc_fermion_force(const orc_dirac* D, const cx* X, double* F);
void orc_gauge_force(int L, const cx* U, double beta, double* F);
int orc_hmc_evolve(int L, cx* U, double mass, double r, double csw, double mu_t, double beta, double tau,
                   int n_steps, int omelyan, int cg_maxit, double cg_tol, double* pi, const cx* phi);
int orc_hmc_hamiltonian(int L, const cx* U, double mass, double r, double csw, double mu_t, double beta,
                        int cg_maxit, double cg_tol, const double* pi, const cx* phi, double* out);
int orc_hmc_evolve_eo(int L, cx* U, double mass, double r, double csw, double mu_t, double beta, double tau,
                      int n_steps, int cg_maxit, double cg_tol, double* pi, const cx* phi_o);
int orc_hmc_hamiltonian_eo(int L, const cx* U, double mass, double r, double csw, double mu_t, double beta,
                           int cg_maxit, double cg_tol, const double* pi, const cx* phi_o, double* out);
void orc_lanczos_eigen(int n, const cx* A, double* evals, cx* evecs);
int orc_trlm(const orc_op* A, int n_ev, int n_kr, int max_restarts, double tol, const cx* start, double* evals,
             cx* evecs, int* restarts_out, int* iters_out);
int orc_cg_deflated(const orc_op* A, const cx* rhs, orc_mg* mg, int ndefl, const cx* U, int maxit, double tol, cx* x,
                    double* final_res);
int orc_lobpcg(const orc_op* A, orc_mg* mg, int k, const cx* X0, int max_iter, double tol, cx* Xout, double* evals);
int orc_fgmres_ritz(const orc_op* A, const cx* rhs, orc_mg* mg, int restart, int maxit, double tol, int n_ritz,
                    cx* x, double* final_res, int* n_ritz_out, double* ritz_vals, cx* ritz_vecs);
void orc_heatbath(int L, cx* U, double beta, int n_sweeps, uint64_t seed, uint32_t first_sweep);
typedef struct orc_prolong orc_prolong;
orc_prolong* orc_prolong_new(int nx, int ny, int nd, int bx, int by, int k, int chiral);
void orc_prolong_build(orc_prolong* P, const cx* null_vecs);
void orc_prolong_free(orc_prolong* P);
int orc_hmc_ms_evolve(int L, cx* U, double mass, double r, double csw, double mu_t, double beta, double tau,
                      int n_outer, int n_inner, int omelyan, int inner_smooth, int defl_refresh, int cg_maxit,
                      double cg_tol, const orc_prolong* P, int k, cx* V, double* vals, double* pi, const cx* phi,
                      int* lowmode_evals);
}

static int failures = 0;
static void report(const char* what, double err, double tol) {
    const bool ok = std::isfinite(err) && err <= tol;
    std::printf("%-44s %.3e (tol %.0e) %s\n", what, err, tol, ok ? "ok" : "FAIL");
    std::fflush(stdout);
    if (!ok) failures++;
}
static double relerr(const Vec& a, const Vec& b) {
    double num = 0, den = 0;
    for (size_t i = 0; i < a.size(); i++) {
        num += std::norm(a[i] - b[i]);
        den += std::norm(b[i]);
    }
    return std::sqrt(num / den);
}
static double relerr(const std::array<RVec, 2>& a, const RVec& b, int V) {
    double num = 0, den = 0;
    for (int mu = 0; mu < 2; mu++)
        for (int s = 0; s < V; s++) {
            const double d = a[mu][s] - b[(size_t)mu * V + s];
            num += d * d;
            den += b[(size_t)mu * V + s] * b[(size_t)mu * V + s];
        }
    return std::sqrt(num / den);
}

int main() {
    const int L = 32;
    Lattice lat(L);
    GaugeField gauge(lat);
    std::mt19937 rng(2024);
    gauge.randomise(rng, 0.4);
    DiracOp D(lat, gauge, 0.05, 1.0, 1.0, 0.0);
    orc_dirac* Do = orc_dirac_new(L, gauge.links().data(), 0.05, 1.0, 1.0, 0.0);

    // operators (src/dirac.h:115-152, 399-445)
    const Vec psi = random_vec(lat.ndof, rng);
    Vec got(lat.ndof), want(lat.ndof);
    D.apply(psi, got);
    orc_dirac_apply(Do, 0, psi.data(), want.data());
    report("DiracOp::apply", relerr(got, want), 1e-12);
    OpApply A = D.as_DdagD_op();
    A(psi, got);
    orc_dirac_apply(Do, 2, psi.data(), want.data());
    report("DiracOp::as_DdagD_op", relerr(got, want), 1e-12);
    EvenOddDiracOp eo(D);
    const Vec half = random_vec(2 * lat.V_half, rng);
    Vec hg(half.size()), hw(half.size());
    eo.apply_schur_dag_schur(half, hg);
    orc_eo_apply(Do, 7, half.data(), hw.data());
    report("EvenOddDiracOp::apply_schur_dag_schur", relerr(hg, hw), 1e-12);

    // build_mg_hierarchy with the reference's rng order; the oracle gets the same starts
    std::mt19937 rng_copy = rng;
    MGHierarchy mg = build_mg_hierarchy(D, 3, 4, 4, 16, 10, rng, true, 3, 3, false);
    Vec starts;
    for (int l = 0; l < 2; l++)
        for (int i = 0; i < 4; i++) {
            const Vec v = random_vec(mg.dims[l], rng_copy);
            starts.insert(starts.end(), v.begin(), v.end());
        }
    orc_op* Ao = orc_op_dirac(Do, 2);
    orc_mg_params p{};
    p.n_levels = 3;
    p.bx[0] = p.by[0] = 4;
    p.bx[1] = 4;
    p.by[1] = 1;
    p.k[0] = p.k[1] = 4;
    p.null_iters = 10;
    p.pre_smooth = p.post_smooth = 3;
    p.w_cycle = 1;
    p.coarse_solver = 1;
    p.colored_build = 1;
    orc_mg* mgo = orc_mg_new(Ao, L, L, 2, &p, starts.data());
    const Vec b = random_vec(lat.ndof, rng);
    const Vec xg = mg.precondition(b);
    Vec xo(lat.ndof);
    orc_mg_precondition(mgo, b.data(), xo.data());
    report("MGHierarchy::precondition (3 levels, W)", relerr(xg, xo), 1e-10);

    // build_mg_hierarchy_warm (src/multigrid.cpp:688-701): the level-0 starts are the
    // previous hierarchy's null vectors and take no draws from rng; level 1 draws
    std::mt19937 rng_w(77);
    std::mt19937 rng_w_copy = rng_w;
    MGHierarchy mgw = build_mg_hierarchy_warm(D, 3, 4, 4, 16, 5, rng_w, mg.null_vecs_l0, true, 3, 3, false);
    Vec wstarts;
    for (const Vec& v : mg.null_vecs_l0) wstarts.insert(wstarts.end(), v.begin(), v.end());
    for (int i = 0; i < 4; i++) {
        const Vec v = random_vec(mg.dims[1], rng_w_copy);
        wstarts.insert(wstarts.end(), v.begin(), v.end());
    }
    report("warm start: level 0 takes no rng draws", rng_w() == rng_w_copy() ? 0.0 : 1.0, 0.0);
    p.null_iters = 5;
    orc_mg* mgwo = orc_mg_new(Ao, L, L, 2, &p, wstarts.data());
    p.null_iters = 10;
    const Vec xw = mgw.precondition(b);
    Vec xwo(lat.ndof);
    orc_mg_precondition(mgwo, b.data(), xwo.data());
    report("build_mg_hierarchy_warm precondition", relerr(xw, xwo), 1e-10);
    orc_mg_free(mgwo);

    // solvers (src/solvers.cpp)
    CGResult cg = cg_solve(A, lat.ndof, b, 2000, 1e-10);
    Vec xcg(lat.ndof);
    double res = 0;
    const int ito = orc_cg(Ao, b.data(), nullptr, nullptr, 2000, 1e-10, xcg.data(), &res);
    report("cg_solve iterations vs oracle", std::abs(cg.iterations - ito), 1);
    report("cg_solve solution", relerr(cg.solution, xcg), 1e-8);
    FCGResult g = gcr_solve_mg(A, lat.ndof, b, mg, 200, 1e-10, 10, Precision::Mixed);
    Vec Ax(lat.ndof);
    A(g.solution, Ax);
    double rn = 0, bn = 0;
    for (int i = 0; i < lat.ndof; i++) {
        rn += std::norm(b[i] - Ax[i]);
        bn += std::norm(b[i]);
    }
    report("gcr_solve_mg (mixed) true residual", std::sqrt(rn / bn), 1e-10);
    report("gcr_solve_mg (mixed) vs CG solution", relerr(g.solution, xcg), 1e-8);
    mg.set_symmetric(0.8);
    CGResult pcg = cg_solve_precond(A, lat.ndof, b, mg.as_precond(), 500, 1e-10);
    report("cg_solve_precond (symmetric MG) vs CG", relerr(pcg.solution, xcg), 1e-8);

    // forces (src/hmc_forces.cpp), F = -dS/dtheta
    std::array<RVec, 2> F;
    RVec Fo(2 * (size_t)lat.V);
    fermion_force(D, cg.solution, F);
    orc_fermion_force(Do, cg.solution.data(), Fo.data());
    report("fermion_force", relerr(F, Fo, lat.V), 1e-12);
    gauge_force(gauge, 4.0, F);
    orc_gauge_force(L, gauge.links().data(), 4.0, Fo.data());
    report("gauge_force", relerr(F, Fo, lat.V), 1e-12);
    FermionActionResult fa = fermion_action(D, b, 2000, 1e-10);
    report("fermion_action X vs CG", relerr(fa.X, xcg), 1e-8);

    // HMC (src/hmc_trajectory.cpp): the MD evolution and a trajectory against the oracle
    {
        const int Lh = 8;
        Lattice lh(Lh);
        GaugeField gh(lh);
        std::mt19937 r2(7);
        gh.randomise(r2, 0.3);
        for (int om = 0; om < 2; om++) {
            HMCParams hp{};
            hp.beta = 2.0;
            hp.tau = 0.5;
            hp.n_steps = 4;
            hp.cg_maxiter = 2000;
            hp.cg_tol = 1e-12;
            hp.c_sw = 1.0;
            hp.omelyan = om == 1;
            GaugeField ge(lh);
            ge.U = gh.U;
            MomentumField mom(lh);
            mom.randomise(r2);
            const Vec phi = random_vec(lh.ndof, r2);
            RVec pio = detail::mom_flat(mom);
            Vec Uo = gh.links();
            plain_leapfrog_evolve(ge, mom, lh, 0.1, 1.0, phi, hp);
            orc_hmc_evolve(Lh, Uo.data(), 0.1, 1.0, 1.0, 0.0, 2.0, 0.5, 4, om, 2000, 1e-12, pio.data(), phi.data());
            report(om ? "plain_leapfrog_evolve (Omelyan) links" : "plain_leapfrog_evolve links",
                   relerr(ge.links(), Uo), 1e-11);
            report(om ? "plain_leapfrog_evolve (Omelyan) momenta" : "plain_leapfrog_evolve momenta",
                   relerr(mom.pi, pio, lh.V), 1e-10);
        }
        // hmc_trajectory draws momenta, then eta, from rng: the oracle replays the draws
        HMCParams hp{};
        hp.beta = 2.0;
        hp.tau = 0.5;
        hp.n_steps = 4;
        hp.cg_maxiter = 2000;
        hp.cg_tol = 1e-12;
        hp.c_sw = 1.0;
        hp.force_accept = true;
        GaugeField gt(lh);
        gt.U = gh.U;
        std::mt19937 ra(99), rb(99);
        HMCResult hr = hmc_trajectory(gt, lh, 0.1, 1.0, hp, ra);
        MomentumField m2(lh);
        m2.randomise(rb);
        const Vec eta = random_vec(lh.ndof, rb);
        Vec phi(lh.ndof);
        orc_dirac* Dh = orc_dirac_new(Lh, gh.links().data(), 0.1, 1.0, 1.0, 0.0);
        orc_dirac_apply(Dh, 1, eta.data(), phi.data());
        orc_dirac_free(Dh);
        RVec pio = detail::mom_flat(m2);
        Vec Uo = gh.links();
        double h0[3], h1[3];
        orc_hmc_hamiltonian(Lh, Uo.data(), 0.1, 1.0, 1.0, 0.0, 2.0, 2000, 1e-12, pio.data(), phi.data(), h0);
        orc_hmc_evolve(Lh, Uo.data(), 0.1, 1.0, 1.0, 0.0, 2.0, 0.5, 4, 0, 2000, 1e-12, pio.data(), phi.data());
        orc_hmc_hamiltonian(Lh, Uo.data(), 0.1, 1.0, 1.0, 0.0, 2.0, 2000, 1e-12, pio.data(), phi.data(), h1);
        report("hmc_trajectory dH", std::abs(hr.dH - ((h1[0] + h1[1] + h1[2]) - (h0[0] + h0[1] + h0[2]))), 1e-8);
        report("hmc_trajectory links", relerr(gt.links(), Uo), 1e-11);
        // the even-odd trajectory: momenta, then eta_o (V complex), phi_o = M^dag eta_o
        {
            HMCParams he = hp;
            he.use_eo = true;
            GaugeField ge(lh);
            ge.U = gh.U;
            std::mt19937 ea(31), eb(31);
            HMCResult er = hmc_trajectory(ge, lh, 0.1, 1.0, he, ea);
            MomentumField m3(lh);
            m3.randomise(eb);
            const Vec eta_o = random_vec(lh.V, eb);
            Vec phi_o(lh.V);
            orc_dirac* De = orc_dirac_new(Lh, gh.links().data(), 0.1, 1.0, 1.0, 0.0);
            orc_eo_apply(De, 6, eta_o.data(), phi_o.data());
            orc_dirac_free(De);
            RVec pe = detail::mom_flat(m3);
            Vec Ue = gh.links();
            double e0[4], e1[4];
            orc_hmc_hamiltonian_eo(Lh, Ue.data(), 0.1, 1.0, 1.0, 0.0, 2.0, 2000, 1e-12, pe.data(), phi_o.data(), e0);
            orc_hmc_evolve_eo(Lh, Ue.data(), 0.1, 1.0, 1.0, 0.0, 2.0, 0.5, 4, 2000, 1e-12, pe.data(), phi_o.data());
            orc_hmc_hamiltonian_eo(Lh, Ue.data(), 0.1, 1.0, 1.0, 0.0, 2.0, 2000, 1e-12, pe.data(), phi_o.data(), e1);
            double dHe = 0;
            for (int i = 0; i < 4; i++) dHe += e1[i] - e0[i];
            report("hmc_trajectory (even-odd) dH", std::abs(er.dH - dHe), 1e-8);
            report("hmc_trajectory (even-odd) dLD", std::abs(er.dLD - (e1[3] - e0[3])), 1e-9);
            report("hmc_trajectory (even-odd) links", relerr(ge.links(), Ue), 1e-11);
        }
        GaugeField gr(lh);
        gr.U = gh.U;
        std::mt19937 rr(5);
        ReversibilityResult rv = reversibility_test_plain(gr, lh, 0.1, 1.0, hp, rr);
        report("reversibility_test_plain gauge", rv.gauge_delta, 1e-10);
        report("reversibility_test_plain momenta", rv.mom_delta, 1e-9);
        report("verify_forces (full, clover)", verify_forces(gh, 2.0, 0.1, 1.0, 2000, 1e-12, 1.0) ? 0.0 : 1.0, 0.0);
        report("verify_forces (even-odd, clover)",
               verify_forces(gh, 2.0, 0.1, 1.0, 2000, 1e-12, 1.0, true) ? 0.0 : 1.0, 0.0);
    }

    // eigensolvers and deflation (src/eigensolver.h, src/coarse_op.h, src/solvers.h)
    {
        std::mt19937 r1(12345);
        const Vec rv = random_vec(64, r1);
        Vec rl(64);
        smg_random_vec(64, 12345, reinterpret_cast<double*>(rl.data()));
        report("random_vec (the g++ draw order)", relerr(rv, rl), 0.0);
        const int nh = 10;
        std::vector<Vec> Hc(nh, Vec(nh));
        std::mt19937 rh(7);
        std::normal_distribution<double> nd(0.0, 1.0);
        for (int j = 0; j < nh; j++)
            for (int i = 0; i <= j; i++) {
                const cx z = i == j ? cx(nd(rh), 0.0) : cx(nd(rh), nd(rh));
                Hc[j][i] = z;
                Hc[i][j] = std::conj(z);
            }
        RVec ev;
        std::vector<Vec> evec;
        lanczos_eigen(Hc, nh, ev, evec);
        Vec flat((size_t)nh * nh), ovec((size_t)nh * nh);
        for (int j = 0; j < nh; j++)
            for (int i = 0; i < nh; i++) flat[(size_t)j * nh + i] = Hc[j][i];
        RVec eo_(nh);
        orc_lanczos_eigen(nh, flat.data(), eo_.data(), ovec.data());
        double de = 0;
        for (int i = 0; i < nh; i++) de = std::max(de, std::abs(ev[i] - eo_[i]));
        report("lanczos_eigen eigenvalues", de, 1e-13);
        double dv = 0;
        for (int r = 0; r < nh; r++)
            for (int c = 0; c < nh; c++) dv = std::max(dv, std::abs(evec[r][c] - ovec[(size_t)r * nh + c]));
        report("lanczos_eigen evecs[row][col]", dv, 1e-12);

        const int n_ev = 6, n = lat.ndof;
        TRLMResult tr = trlm_eigensolver(A, n, n_ev, 24, 100, 1e-10);
        std::mt19937 r42(42);
        const Vec st = random_vec(n, r42);
        RVec evo(n_ev);
        Vec vo((size_t)n_ev * n);
        int rs = 0, its = 0;
        orc_trlm(Ao, n_ev, 24, 100, 1e-10, st.data(), evo.data(), vo.data(), &rs, &its);
        de = 0;
        for (int i = 0; i < n_ev; i++) de = std::max(de, std::abs(tr.eigvals[i] - evo[i]) / evo[n_ev - 1]);
        report("trlm_eigensolver eigenvalues", de, 1e-10);
        report("trlm_eigensolver restarts / iterations",
               (tr.converged && tr.num_restarts == rs && tr.iterations == its) ? 0.0 : 1.0, 0.0);

        Vec U((size_t)n_ev * n);
        for (int i = 0; i < n_ev; i++) std::copy(tr.eigvecs[i].begin(), tr.eigvecs[i].end(), U.begin() + (size_t)i * n);
        CGResult cd = cg_solve_deflated(A, n, b, {}, tr.eigvecs, tr.eigvals, 2000, 1e-10);
        Vec xd(n);
        double rd = 0;
        const int itd = orc_cg_deflated(Ao, b.data(), nullptr, n_ev, U.data(), 2000, 1e-10, xd.data(), &rd);
        report("cg_solve_deflated", relerr(cd.solution, xd), 1e-8);
        report("cg_solve_deflated iterations", std::abs(cd.iterations - itd), 1.0);
        // mg holds the symmetric cycle here (set_symmetric above), which the oracle's
        // hierarchy does not: the same solution, in fewer iterations
        CGResult cm = cg_solve_deflated(A, n, b, mg.as_precond(), tr.eigvecs, tr.eigvals, 200, 1e-10);
        report("cg_solve_deflated (symmetric MG) solution", relerr(cm.solution, xd), 1e-8);
        report("cg_solve_deflated (symmetric MG) fewer iterations", cm.iterations < cd.iterations ? 0.0 : 1.0, 0.0);

        // SparseCoarseOp::setup_deflation + solve: x solves A_c x = b_c
        Prolongator P(lat, 4, 4, 4);
        std::vector<Vec> nv;
        for (int i = 0; i < 4; i++) nv.push_back(random_vec(lat.ndof, rng));
        P.build_from_vectors(nv);
        SparseCoarseOp S;
        S.build(P, A, lat.ndof);
        S.setup_deflation(8);
        const Vec bc = random_vec(S.dim, rng);
        const Vec xc = S.solve(bc, 500, 1e-12);
        Vec Axc;
        S.apply_to(xc, Axc);
        report("SparseCoarseOp deflation eigenvalues ascend",
               (S.defl_vals.size() == 8 && S.defl_vals[0] > 0 && S.defl_vals[0] <= S.defl_vals[7]) ? 0.0 : 1.0, 0.0);
        report("SparseCoarseOp::solve (deflated) residual", relerr(Axc, bc), 1e-11);
        TRLMResult tc = trlm_eigensolver(S.as_op(), S.dim, 8, 0, 100, 1e-10);
        de = 0;
        for (int i = 0; i < 8; i++) de = std::max(de, std::abs(tc.eigvals[i] - S.defl_vals[i]));
        report("trlm_eigensolver on SparseCoarseOp::as_op", de, 1e-9 * S.defl_vals[7]);
        report("SparseCoarseOp::validate_deflation", S.validate_deflation() == 8 ? 0.0 : 1.0, 0.0);

        // lobpcg_update (identity preconditioner) against the oracle from the same block
        const int kl = 8;
        std::vector<Vec> X0;
        Vec x0f;
        for (int i = 0; i < kl; i++) {
            X0.push_back(random_vec(n, rng));
            x0f.insert(x0f.end(), X0.back().begin(), X0.back().end());
        }
        LOBPCGResult lr = lobpcg_update(A, n, kl, X0, identity_precond(), 8, 1e-10);
        Vec xl((size_t)kl * n), xg;
        RVec el(kl);
        const int itl = orc_lobpcg(Ao, nullptr, kl, x0f.data(), 8, 1e-10, xl.data(), el.data());
        de = 0;
        for (int i = 0; i < kl; i++) {
            de = std::max(de, std::abs(lr.eigvals[i] - el[i]) / std::abs(el[0]));
            xg.insert(xg.end(), lr.eigvecs[i].begin(), lr.eigvecs[i].end());
        }
        report("lobpcg_update Rayleigh quotients", de, 1e-9);
        report("lobpcg_update vectors", relerr(xg, xl), 1e-7);
        report("lobpcg_update iterations", std::abs(lr.iterations - itl), 0.0);
        // fgmres_solve_generic with the Ritz harvest (identity preconditioner, two cycles)
        {
            const Vec bf = random_vec(n, rng);
            FGMRESResult fr = fgmres_solve_generic(A, n, bf, identity_precond(), 10, 20, 1e-10, 3);
            Vec xo(n), vo((size_t)(2 * 3 + 10) * n);
            RVec lo(2 * 3 + 10);
            double reso = 0;
            int nro = 0;
            const int ito = orc_fgmres_ritz(Ao, bf.data(), nullptr, 10, 20, 1e-10, 3, xo.data(), &reso, &nro,
                                            lo.data(), vo.data());
            report("fgmres_solve_generic (Ritz harvest) solution", relerr(fr.solution, xo), 1e-11);
            report("fgmres_solve_generic (Ritz harvest) iterations", std::abs(fr.iterations - ito), 0.0);
            report("fgmres_solve_generic ritz_pairs count", std::abs((int)fr.ritz_pairs.size() - nro), 0.0);
            double dv = 0;
            Vec vg;
            for (size_t q = 0; q < fr.ritz_pairs.size() && (int)q < nro; q++) {
                dv = std::max(dv, std::abs(fr.ritz_pairs[q].value - lo[q]) / std::abs(lo[nro - 1]));
                vg.insert(vg.end(), fr.ritz_pairs[q].vector.begin(), fr.ritz_pairs[q].vector.end());
            }
            vo.resize((size_t)nro * n);
            report("fgmres_solve_generic Ritz values", dv, 1e-10);
            report("fgmres_solve_generic Ritz vectors", relerr(vg, vo), 1e-8);
        }
        mg.update_coarsest_deflation(8, 3);
        report("MGHierarchy::update_coarsest_deflation",
               mg.coarsest_defl_vals.size() == 8 && mg.coarsest_defl_vecs[0].size() == (size_t)mg.dims.back() ? 0.0
                                                                                                       : 1.0,
               0.0);
    }

    // hmc_trajectory_mg_multiscale (Omelyan outer, smoothed low-mode force, refresh every
    // 2 inner steps) against the oracle's MD + Hamiltonian on the same draws
    {
        const int Lm = 16;
        Lattice lm(Lm);
        GaugeField gm(lm);
        std::mt19937 rg(77);
        gm.randomise(rg, 0.3);
        const GaugeField g0 = gm;
        DiracOp Dm(lm, gm, 0.1, 1.0, 1.0, 0.0);
        Prolongator Pm(lm, 4, 4, 4);
        std::vector<Vec> nv;
        Vec nvf;
        for (int i = 0; i < 4; i++) {
            nv.push_back(random_vec(lm.ndof, rg));
            nvf.insert(nvf.end(), nv.back().begin(), nv.back().end());
        }
        Pm.build_from_vectors(nv);
        SparseCoarseOp Sm;
        Sm.build(Pm, Dm.as_DdagD_op(), lm.ndof);
        Sm.setup_deflation(6);
        CoarseDeflState cd{Sm.defl_vecs, Sm.defl_vals};
        const CoarseDeflState cd0 = cd;
        MGMultiScaleParams mp;
        mp.beta = 2.0;
        mp.tau = 0.3;
        mp.n_outer = 3;
        mp.n_inner = 3;
        mp.cg_tol = 1e-12;
        mp.inner_smooth = 3;
        mp.outer_type = OuterIntegrator::Omelyan;
        mp.defl_refresh = 2;
        mp.c_sw = 1.0;
        std::function<Vec(const Vec&)> idp = identity_precond();
        std::mt19937 rt(5), rc(5);
        MGMultiScaleResult mr = hmc_trajectory_mg_multiscale(gm, lm, 0.1, 1.0, mp, cd, Pm, idp, rt);
        // the oracle from the same draws
        MomentumField mo(lm);
        mo.randomise(rc);
        (void)random_vec(lm.ndof, rc);
        const Vec eta = random_vec(lm.ndof, rc);
        Vec Uo = g0.links(), phi(lm.ndof);
        orc_dirac* Dq = orc_dirac_new(Lm, Uo.data(), 0.1, 1.0, 1.0, 0.0);
        orc_dirac_apply(Dq, 1, eta.data(), phi.data());
        orc_dirac_free(Dq);
        orc_prolong* Po = orc_prolong_new(Lm, Lm, 2, 4, 4, 4, 0);
        orc_prolong_build(Po, nvf.data());
        Vec Vd;
        for (auto& v : cd0.eigvecs) Vd.insert(Vd.end(), v.begin(), v.end());
        RVec vals = cd0.eigvals, pio = detail::mom_flat(mo);
        double e0[3], e1[3];
        orc_hmc_hamiltonian(Lm, Uo.data(), 0.1, 1.0, 1.0, 0.0, 2.0, 500, 1e-12, pio.data(), phi.data(), e0);
        int nl = 0;
        orc_hmc_ms_evolve(Lm, Uo.data(), 0.1, 1.0, 1.0, 0.0, 2.0, 0.3, 3, 3, 1, 3, 2, 500, 1e-12, Po, 6, Vd.data(),
                          vals.data(), pio.data(), phi.data(), &nl);
        orc_hmc_hamiltonian(Lm, Uo.data(), 0.1, 1.0, 1.0, 0.0, 2.0, 500, 1e-12, pio.data(), phi.data(), e1);
        orc_prolong_free(Po);
        const double dHo = (e1[0] + e1[1] + e1[2]) - (e0[0] + e0[1] + e0[2]);
        report("hmc_trajectory_mg_multiscale dH", std::abs(mr.dH - dHo), 1e-8);
        report("hmc_trajectory_mg_multiscale low-mode evaluations", std::abs(mr.lowmode_force_evals - nl), 0.0);
        double dv = 0;
        for (int i = 0; i < 6; i++) dv = std::max(dv, std::abs(cd.eigvals[i] - vals[i]));
        report("hmc_trajectory_mg_multiscale evolved deflation", dv, 1e-10);
        if (mr.accepted) report("hmc_trajectory_mg_multiscale links", relerr(gm.links(), Uo), 1e-10);
        else report("hmc_trajectory_mg_multiscale rejected: links restored", relerr(gm.links(), g0.links()), 0.0);
    }

    {   // the heatbath addition: 10 sweeps at beta = 4 against the oracle
        Lattice lh(16);
        GaugeField gh(lh);
        std::mt19937 rh(3);
        gh.randomise(rh, 0.35);
        Vec uo = gh.links();
        heatbath(gh, 4.0, 10, 99);
        orc_heatbath(16, uo.data(), 4.0, 10, 99, 0);
        report("heatbath (addition) links", relerr(gh.links(), uo), 1e-10);
    }

    orc_mg_free(mgo);
    orc_op_free(Ao);
    orc_dirac_free(Do);
    std::printf("%s\n", failures ? "FAILED" : "ALL OK");
    return failures ? 1 : 0;
}
