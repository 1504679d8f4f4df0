// Callers written the way the reference's own code calls its solvers: OpApply
// lambdas around DiracOp::apply_DdagD (src/hmc_forces.cpp:113-129, 919-932) and
// preconditioner lambdas around MGHierarchy::precondition (src/mode_mg_hmc.cpp:92-97),
// handed to cg_solve, cg_solve_precond, fcg_solve, fgmres_solve_generic,
// mr_smooth_op, estimate_lambda_max and SparseCoarseOp::build of
// include/schwinger_mg.hpp.  Each result is checked against the oracle
// (liboracle.so) and against the same call with the header's own operator objects,
// which the lambdas reproduce to roundoff: the same kernels run, with a host round
// trip per application, except that a callback operator's dot products are a pass
// of their own instead of being fused into the stencil (another summation order).
// An operator that is not a Dirac operator at all and a callback that throws are
// checked too, and so is EvenOddMG preconditioning with a lambda smoothing operator.  Prints one line per check; exits non-zero on a failure.
#include <cmath>
#include <cstdio>
#include <random>
#include <stdexcept>

#include "schwinger_mg.hpp"

using namespace schwinger_mg;

extern "C" {
typedef struct orc_dirac orc_dirac;
typedef struct orc_op orc_op;
orc_dirac* orc_dirac_new(int L, const cx* U, double mass, double r, double csw, double mu_t);
void orc_dirac_free(orc_dirac* D);
orc_op* orc_op_dirac(const orc_dirac* D, int kind);
void orc_op_free(orc_op* A);
int orc_cg(const orc_op* A, const cx* rhs, const cx* x0, void* mg, int maxit, double tol, cx* x, double* res);
}

static int failures = 0;
static void report(const char* what, double err, double tol) {
    const bool ok = std::isfinite(err) && err <= tol;
    std::printf("%-52s %.3e (tol %.0e) %s\n", what, err, tol, ok ? "ok" : "FAIL");
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

int main() {
    const int L = 16;
    Lattice lat(L);
    GaugeField gauge(lat);
    std::mt19937 rng(77);
    gauge.randomise(rng, 0.4);
    const double mass = 0.1;
    DiracOp D(lat, gauge, mass, 1.0, 1.0, 0.0);
    orc_dirac* Do = orc_dirac_new(L, gauge.links().data(), mass, 1.0, 1.0, 0.0);
    orc_op* Ao = orc_op_dirac(Do, 2);
    const Vec phi = random_vec(lat.ndof, rng);

    // fermion_action, src/hmc_forces.cpp:113-119
    OpApply A = [&D](const Vec& src, Vec& dst) { D.apply_DdagD(src, dst); };
    auto res = cg_solve(A, D.lat.ndof, phi, 2000, 1e-10);
    Vec xo(lat.ndof);
    double ores = 0;
    const int ito = orc_cg(Ao, phi.data(), nullptr, nullptr, 2000, 1e-10, xo.data(), &ores);
    report("cg_solve(lambda) iterations vs oracle", std::abs(res.iterations - ito), 1);
    report("cg_solve(lambda) solution vs oracle", relerr(res.solution, xo), 1e-8);
    const double act = std::real(dot(phi, res.solution)), acto = std::real(dot(phi, xo));
    report("fermion action (lambda) vs oracle", std::abs(act - acto) / std::abs(acto), 1e-10);
    const auto nat = cg_solve(D.as_DdagD_op(), D.lat.ndof, phi, 2000, 1e-10);
    report("cg_solve(lambda) vs cg_solve(as_DdagD_op)", relerr(res.solution, nat.solution), 1e-13);
    report("cg_solve(lambda) iterations == native", std::abs(res.iterations - nat.iterations), 0);

    // preconditioner lambdas, src/mode_mg_hmc.cpp:92-97, and fermion_action_precond,
    // src/hmc_forces.cpp:121-129
    MGHierarchy mg = build_mg_hierarchy(D, 2, 4, 4, 16, 10, rng, true, 3, 3, false);
    mg.set_symmetric(0.8);
    std::function<Vec(const Vec&)> mg_precond = [&mg](const Vec& b) -> Vec { return mg.precondition(b); };
    auto pres = cg_solve_precond(A, D.lat.ndof, phi, mg_precond, 500, 1e-10);
    auto pnat = cg_solve_precond(D.as_DdagD_op(), D.lat.ndof, phi, mg.as_precond(), 500, 1e-10);
    report("cg_solve_precond(lambdas) vs native objects", relerr(pres.solution, pnat.solution), 1e-13);
    report("cg_solve_precond(lambdas) iterations == native", std::abs(pres.iterations - pnat.iterations), 0);
    report("cg_solve_precond(lambdas) vs oracle CG", relerr(pres.solution, xo), 1e-8);
    auto x0 = random_vec(lat.ndof, rng);
    auto r0 = cg_solve_x0_precond(A, D.lat.ndof, phi, x0, mg_precond, 500, 1e-10);
    report("cg_solve_x0_precond(lambdas) vs oracle CG", relerr(r0.solution, xo), 1e-8);

    // fcg_solve and fgmres_solve_generic with lambdas (src/solvers.cpp:223-300, 555-800)
    auto f1 = fcg_solve(A, lat.ndof, phi, mg_precond, 200, 1e-10, 10);
    auto f2 = fcg_solve(D.as_DdagD_op(), lat.ndof, phi, mg.as_precond(), 200, 1e-10, 10);
    report("fcg_solve(lambdas) == native objects", relerr(f1.solution, f2.solution), 0.0);
    report("fcg_solve(lambdas) vs oracle CG", relerr(f1.solution, xo), 1e-8);
    auto g1 = fgmres_solve_generic(A, lat.ndof, phi, mg_precond, 10, 200, 1e-10, 2);
    auto g2 = fgmres_solve_generic(D.as_DdagD_op(), lat.ndof, phi, mg.as_precond(), 10, 200, 1e-10, 2);
    report("fgmres_solve_generic(lambdas) == native objects", relerr(g1.solution, g2.solution), 0.0);
    report("fgmres_solve_generic(lambdas) Ritz pairs", std::abs((double)g1.ritz_pairs.size() - g2.ritz_pairs.size()),
           0);
    report("fgmres_solve_generic(lambdas) vs oracle CG", relerr(g1.solution, xo), 1e-8);

    // smoother, power iteration and the Galerkin build with a lambda operator
    Vec xs1 = random_vec(lat.ndof, rng), xs2 = xs1;
    mr_smooth_op(A, xs1, phi, 5);
    mr_smooth(D, xs2, phi, 5);
    report("mr_smooth_op(lambda) vs mr_smooth(D)", relerr(xs1, xs2), 1e-13);
    report("estimate_lambda_max(lambda) == native",
           std::abs(estimate_lambda_max(A, lat.ndof, 20) - estimate_lambda_max(D.as_DdagD_op(), lat.ndof, 20)), 0.0);
    std::vector<Vec> nulls;
    for (int i = 0; i < 4; i++) nulls.push_back(random_vec(lat.ndof, rng));
    Prolongator P(lat, 4, 4, 4);
    P.build_from_vectors(nulls);
    SparseCoarseOp C1, C2;
    C1.build(P, A, lat.ndof);
    C2.build(P, D.as_DdagD_op(), lat.ndof);
    const Vec c = random_vec(C1.dim, rng);
    Vec y1, y2;
    C1.apply_to(c, y1);
    C2.apply_to(c, y2);
    report("SparseCoarseOp::build(lambda) == build(as_DdagD_op)", relerr(y1, y2), 1e-14);

    // an operator that is no Dirac operator: a positive diagonal, solved exactly
    const int n = 1000;
    Vec d(n), b(n);
    for (int i = 0; i < n; i++) {
        d[i] = cx(1.0 + 0.01 * i, 0.0);
        b[i] = cx(std::cos(i), std::sin(0.5 * i));
    }
    OpApply Dg = [&d](const Vec& s, Vec& o) {
        o.resize(s.size());
        for (size_t i = 0; i < s.size(); i++) o[i] = d[i] * s[i];
    };
    auto dr = cg_solve(Dg, n, b, 200, 1e-12);
    Vec exact(n);
    for (int i = 0; i < n; i++) exact[i] = b[i] / d[i];
    report("cg_solve(diagonal lambda) vs exact", relerr(dr.solution, exact), 1e-11);
    std::function<Vec(const Vec&)> jac = [&d](const Vec& s) -> Vec {   // Jacobi: exact in one step
        Vec o(s.size());
        for (size_t i = 0; i < s.size(); i++) o[i] = s[i] / d[i];
        return o;
    };
    auto jr = cg_solve_precond(Dg, n, b, jac, 200, 1e-12);
    report("cg_solve_precond(diagonal, Jacobi) iterations", jr.iterations, 1);
    report("cg_solve_precond(diagonal, Jacobi) vs exact", relerr(jr.solution, exact), 1e-13);

    // a callback's exception reaches the caller
    OpApply bad = [](const Vec&, Vec&) { throw std::runtime_error("operator refused"); };
    bool caught = false;
    try {
        cg_solve(bad, n, b, 10, 1e-10);
    } catch (const std::runtime_error& e) {
        caught = std::string(e.what()) == "operator refused";
    }
    report("an OpApply's exception reaches the caller", caught ? 0.0 : 1.0, 0.0);

    // EvenOddMG (src/eo_multigrid.h:229-287) as its reference would be used: built from
    // an rng, preconditioning with a lambda around apply_schur_dag_schur; against the
    // C-ABI cycle on the same starts (a copy of the rng) and the setup's own operator
    EvenOddDiracOp eoD(D);
    OpApply AMM = [&eoD](const Vec& s, Vec& d) { eoD.apply_schur_dag_schur(s, d); };
    EvenOddMG emg(lat, 4, 4, 4, 3, 3);
    std::mt19937 rng_e(5), rng_e2 = rng_e;
    emg.build(D, 8, rng_e);
    const Vec bo = random_vec(emg.fine_half_dim, rng);
    const Vec xe = emg.precondition(bo, AMM);
    Vec st;
    for (int i = 0; i < 4; i++) {
        const Vec v = random_vec(emg.fine_half_dim, rng_e2);
        st.insert(st.end(), v.begin(), v.end());
    }
    std::mt19937 rng_l(12345);
    const Vec lam = random_vec(emg.fine_half_dim, rng_l);
    smg_eo_mg* eh = nullptr;
    int erc = smg_eo_mg_setup(&eh, D.handle(), 4, 4, 4, 3, 3, 8, reinterpret_cast<const double*>(st.data()),
                              reinterpret_cast<const double*>(lam.data()));
    Vec xn(emg.fine_half_dim);
    if (erc == 0)
        erc = smg_eo_mg_precondition(eh, nullptr, reinterpret_cast<const double*>(bo.data()),
                                     reinterpret_cast<double*>(xn.data()));
    report("EvenOddMG setup and precondition calls", erc, 0);
    report("EvenOddMG::precondition(lambda) vs the device operator", relerr(xe, xn), 1e-13);
    report("EvenOddMG coarse dimension", std::abs(emg.Ac.dim - 64), 0);
    smg_eo_mg_destroy(eh);

    orc_op_free(Ao);
    orc_dirac_free(Do);
    if (failures) {
        std::printf("%d FAILED\n", failures);
        return 1;
    }
    std::printf("ALL OK\n");
    return 0;
}
