// A reference-style caller of the C++ drop-in API (include/schwinger_mg.hpp):
// the code below is written against the reference's own signatures
// (src/dirac.h, src/multigrid.h, src/solvers.h, src/hmc.h) and runs on the GPU
// through libsmg.so.  Every result is checked against the oracle (liboracle.so,
// the CPU restatement) on identical inputs; the program prints one line per check
// and exits non-zero on the first failure.
#include <cmath>
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

    orc_mg_free(mgo);
    orc_op_free(Ao);
    orc_dirac_free(Do);
    std::printf("%s\n", failures ? "FAILED" : "ALL OK");
    return failures ? 1 : 0;
}
