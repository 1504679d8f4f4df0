// eigensolve.cpp -- host sequencing of the eigensolvers and deflated solves (the dense
// problems on the host, every vector operation on the device).
#include "eigen.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <numeric>
#include <random>

namespace smg {

namespace {
using cxd = std::complex<double>;
inline cxd to_std(c64 a) { return cxd(a.re, a.im); }

// y = a x + y on the device (a complex, host)
void axpy_h(c64* y, const c64* x, cxd a, size_t n, hipStream_t st) {
    launch_axpby<double>(y, x, c64{a.real(), a.imag()}, c64{1.0, 0.0}, n, st);
}
// the two-pass classical Gram-Schmidt of r against the m device vectors V[0..m)
// (contiguous, stride n): r -= V (V^dag r), twice; coefficients never leave the device.
// coef holds 64 doubles per started group of 32 vectors.
void cgs2(const c64* V, int m, c64* r, size_t n, DevMem& coef, hipStream_t st) {
    Context& cx = Context::get();
    for (int pass = 0; pass < 2; pass++) {
        for (int c0 = 0; c0 < m; c0 += 32) {   // all dots from the same r first
            const int nv = std::min(32, m - c0);
            VecPtrs<double> P{};
            for (int i = 0; i < nv; i++) P.v[i] = V + (size_t)(c0 + i) * n;
            launch_multi_dot<double>(P, nv, r, n, 0, cx.fin(FIN_STORE, coef.as<double>() + 2 * c0), st);
        }
        for (int c0 = 0; c0 < m; c0 += 32) {
            const int nv = std::min(32, m - c0);
            VecPtrs<double> P{};
            for (int i = 0; i < nv; i++) P.v[i] = V + (size_t)(c0 + i) * n;
            launch_lincomb<double>(r, P, nv, coef.as<double>() + 2 * c0, 2, n, st);
        }
    }
}
// out = sum_j c_j V[j] for m device vectors and host coefficients c
void combine(c64* out, const c64* V, int m, const std::vector<cxd>& c, size_t n, DevMem& coef, hipStream_t st) {
    std::vector<double> h(2 * (size_t)m);
    for (int j = 0; j < m; j++) {
        h[2 * j] = c[j].real();
        h[2 * j + 1] = c[j].imag();
    }
    SMG_HIP(hipMemcpyAsync(coef.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, st));
    for (int c0 = 0; c0 < m; c0 += 32) {
        const int nv = std::min(32, m - c0);
        VecPtrs<double> P{};
        for (int i = 0; i < nv; i++) P.v[i] = V + (size_t)(c0 + i) * n;
        launch_lincomb<double>(out, P, nv, coef.as<double>() + 2 * c0, c0 == 0 ? 0 : 1, n, st);
    }
    SMG_HIP(hipStreamSynchronize(st));   // the host coefficients may go away
}
}  // namespace

void dense_eigh(int n, const c64* A_in, double* evals, c64* evecs_out) {
    if (n <= 0) return;
    auto A = [&](int i, int j) { return to_std(A_in[(size_t)j * n + i]); };
    if (n == 1) {
        evals[0] = A(0, 0).real();
        evecs_out[0] = c64{1.0, 0.0};
        return;
    }
    auto dot = [&](const std::vector<cxd>& a, const std::vector<cxd>& b) {
        cxd s = 0.0;
        for (int i = 0; i < n; i++) s += std::conj(a[i]) * b[i];
        return s;
    };
    auto nrm = [&](const std::vector<cxd>& a) { return std::sqrt(std::real(dot(a, a))); };
    auto matvec = [&](const std::vector<cxd>& x, std::vector<cxd>& y) {
        std::fill(y.begin(), y.end(), 0.0);
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++) y[i] += A(i, j) * x[j];
    };
    auto orth = [&](std::vector<std::vector<cxd>>& Q, int upto, std::vector<cxd>& v) {
        for (int pass = 0; pass < 2; pass++)
            for (int i = 0; i < upto; i++) {
                const cxd pr = dot(Q[i], v);
                for (int k = 0; k < n; k++) v[k] -= pr * Q[i][k];
            }
    };
    // Lanczos tridiagonalisation from e_0 with two MGS passes
    std::vector<std::vector<cxd>> Q(n, std::vector<cxd>(n, 0.0));
    std::vector<double> al(n, 0.0), be(n + 1, 0.0);
    std::vector<cxd> w(n, 0.0);
    Q[0][0] = 1.0;
    matvec(Q[0], w);
    al[0] = std::real(dot(Q[0], w));
    for (int i = 0; i < n; i++) w[i] -= al[0] * Q[0][i];
    for (int j = 1; j < n; j++) {
        be[j] = nrm(w);
        if (be[j] < 1e-14) {   // invariant subspace: an orthogonal unit vector
            bool found = false;
            for (int t = 0; t < n && !found; t++) {
                std::fill(Q[j].begin(), Q[j].end(), 0.0);
                Q[j][t] = 1.0;
                orth(Q, j, Q[j]);
                const double nq = nrm(Q[j]);
                if (nq > 1e-10) {
                    for (auto& q : Q[j]) q /= nq;
                    found = true;
                }
            }
            be[j] = 0.0;
        } else {
            for (int i = 0; i < n; i++) Q[j][i] = w[i] / be[j];
        }
        orth(Q, j, Q[j]);
        const double nq = nrm(Q[j]);
        if (nq > 1e-14)
            for (auto& q : Q[j]) q /= nq;
        matvec(Q[j], w);
        al[j] = std::real(dot(Q[j], w));
        for (int i = 0; i < n; i++) w[i] -= al[j] * Q[j][i] + be[j] * Q[j - 1][i];
        orth(Q, j + 1, w);
    }
    // implicit QL with the Wilkinson shift on the real tridiagonal
    std::vector<double> d = al, e(n, 0.0);
    for (int i = 0; i < n - 1; i++) e[i] = be[i + 1];
    std::vector<double> Z((size_t)n * n, 0.0);   // Z[k*n + i]
    for (int i = 0; i < n; i++) Z[(size_t)i * n + i] = 1.0;
    for (int l = 0; l < n; l++) {
        int iter = 0;
        for (;;) {
            int m;
            for (m = l; m < n - 1; m++) {
                const double dd = std::abs(d[m]) + std::abs(d[m + 1]);
                if (std::abs(e[m]) <= 1e-15 * dd) break;
            }
            if (m == l || ++iter > 200) break;
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = std::hypot(g, 1.0);
            g = d[m] - d[l] + e[l] / (g + std::copysign(r, g));
            double s = 1.0, c = 1.0, p = 0.0;
            bool early = false;
            for (int i = m - 1; i >= l; i--) {
                const double f = s * e[i], b = c * e[i];
                r = std::hypot(f, g);
                e[i + 1] = r;
                if (r < 1e-30) {
                    d[i + 1] -= p;
                    e[m] = 0.0;
                    early = true;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * s + 2.0 * c * b;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - b;
                for (int k = 0; k < n; k++) {
                    const double t = Z[(size_t)k * n + i + 1];
                    Z[(size_t)k * n + i + 1] = s * Z[(size_t)k * n + i] + c * t;
                    Z[(size_t)k * n + i] = c * Z[(size_t)k * n + i] - s * t;
                }
            }
            if (!early) {
                d[l] -= p;
                e[l] = g;
                e[m] = 0.0;
            }
        }
    }
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return d[a] < d[b]; });
    for (int col = 0; col < n; col++) {
        evals[col] = d[idx[col]];
        for (int row = 0; row < n; row++) {
            cxd v = 0.0;
            for (int j = 0; j < n; j++) v += Q[j][row] * Z[(size_t)j * n + idx[col]];
            evecs_out[(size_t)row * n + col] = c64{v.real(), v.imag()};
        }
    }
}

Eigenpairs trlm(LinOp<double>& A, int n_ev, int n_kr, int max_restarts, double tol, const c64* start,
                hipStream_t st) {
    const size_t n = A.n();
    SMG_REQUIRE(n_ev >= 1 && (size_t)n_ev < n, "need 1 <= n_ev < dimension");
    if (n_kr <= 0) n_kr = (int)std::min<size_t>(std::max(2 * n_ev + 6, n_ev + 32), n);
    if ((size_t)n_kr > n) n_kr = (int)n;
    if (n_kr < n_ev + 6) n_kr = std::min<int>(n_ev + 6, (int)n);
    Field<double> K((size_t)(n_kr + 1) * n), r(n), rot;
    // the multi-dot's finish stores all 64 doubles of its 32-vector group, so the
    // coefficient buffer holds whole groups
    DevMem coef(64 * (size_t)((n_kr + 1 + 31) / 32) * sizeof(double));
    auto kv = [&](int j) { return K.p((size_t)j * n); };
    launch_zero<double>(K.p(), (size_t)(n_kr + 1) * n, st);   // zeros, as the reference's kSpace
    launch_copy<double>(kv(0), start, n, st);
    {
        const double nv = dev_norm(kv(0), n, st);
        launch_axpby<double>(kv(0), kv(0), c64{0.0, 0.0}, c64{1.0 / nv, 0.0}, n, st);
    }
    std::vector<double> alpha(n_kr, 0.0), beta(n_kr, 0.0), resid(n_kr, 0.0);
    std::vector<int> order(n_kr + 1);   // kSpace slot -> vector (the final sort permutes slots)
    std::iota(order.begin(), order.end(), 0);
    int num_locked = 0, num_converged = 0, num_keep = 0, restart = 0, iters = 0;
    bool converged = false;
    for (restart = 0; restart < max_restarts && !converged; restart++) {
        iters += n_kr - num_keep;
        for (int j = num_keep; j < n_kr; j++) {   // lanczosStep (src/eigensolver_core.cpp:815-848)
            A.apply(r.p(), kv(j), st);
            alpha[j] = dev_dot<double>(kv(j), r.p(), n, st).re;
            axpy_h(r.p(), kv(j), -alpha[j], n, st);
            for (int i = (j > num_keep ? j - 1 : 0); i < j; i++) axpy_h(r.p(), kv(i), -beta[i], n, st);
            cgs2(K.p(), j + 1, r.p(), n, coef, st);
            beta[j] = dev_norm(r.p(), n, st);
            if (beta[j] > 1e-14)
                launch_axpby<double>(kv(j + 1), r.p(), c64{1.0 / beta[j], 0.0}, c64{0.0, 0.0}, n, st);
        }
        // the arrow matrix of the thick restart and its eigenpairs
        const int dim = n_kr - num_locked, arrow = num_keep - num_locked;
        std::vector<c64> M((size_t)dim * dim, c64{0.0, 0.0}), Z((size_t)dim * dim);
        std::vector<double> ev(dim);
        for (int i = 0; i < dim; i++) M[(size_t)i * dim + i].re = alpha[i + num_locked];
        for (int i = 0; i < arrow; i++) {
            M[(size_t)arrow * dim + i].re = beta[i + num_locked];
            M[(size_t)i * dim + arrow].re = beta[i + num_locked];
        }
        for (int i = arrow; i < dim - 1; i++) {
            M[(size_t)(i + 1) * dim + i].re = beta[i + num_locked];
            M[(size_t)i * dim + i + 1].re = beta[i + num_locked];
        }
        dense_eigh(dim, M.data(), ev.data(), Z.data());
        for (int i = 0; i < dim; i++) {
            resid[i + num_locked] = std::abs(beta[n_kr - 1] * Z[(size_t)(dim - 1) * dim + i].re);
            alpha[i + num_locked] = ev[i];
        }
        int it_locked = 0;
        for (int i = 1; i < dim; i++) {
            if (resid[i + num_locked] < 1e-15 * std::abs(alpha[i + num_locked])) it_locked = i;
            else break;
        }
        int it_conv = it_locked;
        for (int i = it_locked + 1; i < dim; i++) {
            if (resid[i + num_locked] < tol * std::abs(alpha[i + num_locked])) it_conv = i;
            else break;
        }
        int keep = std::min(it_conv + (n_kr - num_converged) / 2, n_kr - num_locked - 12);
        keep = std::max(keep, 1);
        keep = std::min(keep, dim - 1);
        // rotate the Krylov space onto the kept Ritz vectors
        rot = Field<double>((size_t)keep * n);
        std::vector<cxd> c(dim);
        for (int i = 0; i < keep; i++) {
            for (int j = 0; j < dim; j++) c[j] = to_std(Z[(size_t)j * dim + i]);
            combine(rot.p((size_t)i * n), kv(num_locked), dim, c, n, coef, st);
        }
        launch_copy<double>(kv(num_locked), rot.p(), (size_t)keep * n, st);
        launch_copy<double>(kv(num_locked + keep), kv(n_kr), n, st);
        for (int i = 0; i < keep; i++) beta[i + num_locked] = beta[n_kr - 1] * Z[(size_t)(dim - 1) * dim + i].re;
        num_converged = num_locked + it_conv;
        num_keep = num_locked + keep;
        num_locked += it_locked;
        if (num_converged >= n_ev) {
            converged = true;
            for (int i = 1; i < n_kr; i++) {   // ascending (insertion sort, as the reference)
                int j = i;
                while (j > 0 && alpha[j - 1] > alpha[j]) {
                    std::swap(alpha[j], alpha[j - 1]);
                    std::swap(order[j], order[j - 1]);
                    j--;
                }
            }
        }
    }
    Eigenpairs out;
    out.converged = converged;
    out.restarts = restart;
    out.iterations = iters;
    out.vals.assign(alpha.begin(), alpha.begin() + n_ev);
    out.vecs = Field<double>((size_t)n_ev * n);
    for (int i = 0; i < n_ev; i++) launch_copy<double>(out.vecs.p((size_t)i * n), kv(order[i]), n, st);
    SMG_HIP(hipStreamSynchronize(st));
    return out;
}

void lanczos_extremes(LinOp<double>& A, int m, const c64* start, double& lo, double& hi, hipStream_t st) {
    const size_t n = A.n();
    SMG_REQUIRE(m >= 2, "Lanczos bounds: at least two steps");
    if ((size_t)m > n) m = (int)n;
    Field<double> K((size_t)(m + 1) * n), r(n);
    DevMem coef(64 * (size_t)((m + 1 + 31) / 32) * sizeof(double));
    auto kv = [&](int j) { return K.p((size_t)j * n); };
    launch_copy<double>(kv(0), start, n, st);
    {
        const double nv = dev_norm(kv(0), n, st);
        SMG_REQUIRE(nv > 0.0, "Lanczos bounds: zero start vector");
        launch_axpby<double>(kv(0), kv(0), c64{0.0, 0.0}, c64{1.0 / nv, 0.0}, n, st);
    }
    std::vector<double> alpha, beta;
    for (int j = 0; j < m; j++) {
        A.apply(r.p(), kv(j), st);
        alpha.push_back(dev_dot<double>(kv(j), r.p(), n, st).re);
        axpy_h(r.p(), kv(j), -alpha[j], n, st);
        if (j > 0) axpy_h(r.p(), kv(j - 1), -beta[j - 1], n, st);
        cgs2(K.p(), j + 1, r.p(), n, coef, st);
        beta.push_back(dev_norm(r.p(), n, st));
        if (j + 1 == m || beta[j] <= 1e-14 * std::abs(alpha[j])) break;   // invariant subspace
        launch_axpby<double>(kv(j + 1), r.p(), c64{1.0 / beta[j], 0.0}, c64{0.0, 0.0}, n, st);
    }
    const int k = (int)alpha.size();
    std::vector<c64> T((size_t)k * k, c64{0.0, 0.0}), Z((size_t)k * k);
    std::vector<double> ev(k);
    for (int i = 0; i < k; i++) {
        T[(size_t)i * k + i].re = alpha[i];
        if (i + 1 < k) T[(size_t)i * k + i + 1].re = T[(size_t)(i + 1) * k + i].re = beta[i];
    }
    dense_eigh(k, T.data(), ev.data(), Z.data());
    lo = ev.front();
    hi = ev.back();
}

int deflated_coarse_cg(LinOp<double>& A, const Eigenpairs& E, const c64* b, c64* x, int maxit, double tol,
                       hipStream_t st) {
    const size_t n = A.n();
    launch_zero<double>(x, n, st);
    const double bnorm = dev_norm(b, n, st);
    if (bnorm < 1e-30) return 0;
    for (size_t i = 0; i < E.vals.size(); i++)   // x0 = sum (v_i^dag b / lambda_i) v_i
        if (E.vals[i] > 1e-14) {
            const cxd c = to_std(dev_dot<double>(E.vecs.p(i * n), b, n, st)) / E.vals[i];
            axpy_h(x, E.vecs.p(i * n), c, n, st);
        }
    Field<double> r(n), p(n), Ap(n);
    A.apply(Ap.p(), x, st);
    launch_copy<double>(r.p(), b, n, st);
    axpy_h(r.p(), Ap.p(), -1.0, n, st);
    launch_copy<double>(p.p(), r.p(), n, st);
    cxd rr = to_std(dev_dot<double>(r.p(), r.p(), n, st));
    int it = 0;
    while (it < maxit) {
        A.apply(Ap.p(), p.p(), st);
        const cxd pAp = to_std(dev_dot<double>(p.p(), Ap.p(), n, st));
        if (std::abs(pAp) < 1e-30) break;
        const cxd a = rr / pAp;
        axpy_h(x, p.p(), a, n, st);
        axpy_h(r.p(), Ap.p(), -a, n, st);
        it++;
        const double rn = dev_norm(r.p(), n, st);
        if (rn / bnorm < tol || std::isnan(rn) || std::isinf(rn)) break;
        const cxd rn2 = to_std(dev_dot<double>(r.p(), r.p(), n, st));
        const cxd be = rn2 / rr;
        rr = rn2;
        launch_axpby<double>(p.p(), r.p(), c64{1.0, 0.0}, c64{be.real(), be.imag()}, n, st);
    }
    return it;
}

SolveResult cg_deflated(LinOp<double>& A, MG* mg, const c64* U, int nd, const c64* b, c64* x, int maxit, double tol,
                        hipStream_t st) {
    const size_t n = A.n();
    SolveResult res;
    const double bnorm = dev_norm(b, n, st);
    launch_zero<double>(x, n, st);
    if (bnorm < 1e-30) return res;
    // AU and E = U^dag A U, solved on the host
    Field<double> AU((size_t)std::max(nd, 1) * n);
    for (int j = 0; j < nd; j++) A.apply(AU.p((size_t)j * n), U + (size_t)j * n, st);
    std::vector<cxd> E((size_t)nd * nd);   // E[i*nd + j] = <u_i, A u_j>
    for (int i = 0; i < nd; i++)
        for (int j = 0; j < nd; j++)
            E[(size_t)i * nd + j] = to_std(dev_dot<double>(U + (size_t)i * n, AU.p((size_t)j * n), n, st));
    // E c = t by partial-pivot elimination and back substitution (CoarseOp::solve,
    // src/coarse_op.h:101-137)
    auto esolve = [&](const std::vector<cxd>& t) {
        std::vector<std::vector<cxd>> a(nd, std::vector<cxd>(nd + 1));
        for (int i = 0; i < nd; i++) {
            for (int j = 0; j < nd; j++) a[i][j] = E[(size_t)i * nd + j];
            a[i][nd] = t[i];
        }
        for (int k = 0; k < nd; k++) {
            double best = 0.0;
            int br = k;
            for (int i = k; i < nd; i++)
                if (std::abs(a[i][k]) > best) {
                    best = std::abs(a[i][k]);
                    br = i;
                }
            std::swap(a[k], a[br]);
            if (best < 1e-30) continue;
            for (int i = k + 1; i < nd; i++) {
                const cxd f = a[i][k] / a[k][k];
                for (int j = k; j <= nd; j++) a[i][j] -= f * a[k][j];
            }
        }
        std::vector<cxd> c(nd, 0.0);
        for (int i = nd - 1; i >= 0; i--) {
            cxd s = a[i][nd];
            for (int j = i + 1; j < nd; j++) s -= a[i][j] * c[j];
            c[i] = std::abs(a[i][i]) > 1e-30 ? s / a[i][i] : 0.0;
        }
        return c;
    };
    // x0 = U E^-1 U^dag b
    std::vector<cxd> utb(nd);
    for (int i = 0; i < nd; i++) utb[i] = to_std(dev_dot<double>(U + (size_t)i * n, b, n, st));
    const std::vector<cxd> c0 = esolve(utb);
    for (int i = 0; i < nd; i++) axpy_h(x, U + (size_t)i * n, c0[i], n, st);
    Field<double> r(n), z(n), p(n), Ap(n);
    A.apply(Ap.p(), x, st);
    launch_copy<double>(r.p(), b, n, st);
    axpy_h(r.p(), Ap.p(), -1.0, n, st);
    // z = M^-1 r - U E^-1 (AU)^dag M^-1 r
    auto project = [&]() {
        if (mg) mg->precondition(r.p(), z.p(), st);
        else launch_copy<double>(z.p(), r.p(), n, st);
        std::vector<cxd> u(nd);
        for (int i = 0; i < nd; i++) u[i] = to_std(dev_dot<double>(AU.p((size_t)i * n), z.p(), n, st));
        const std::vector<cxd> c = esolve(u);
        for (int i = 0; i < nd; i++) axpy_h(z.p(), U + (size_t)i * n, -c[i], n, st);
    };
    project();
    launch_copy<double>(p.p(), z.p(), n, st);
    cxd rz = to_std(dev_dot<double>(r.p(), z.p(), n, st));
    int it = 0;
    double rn = dev_norm(r.p(), n, st);
    while (it < maxit) {
        A.apply(Ap.p(), p.p(), st);
        const cxd pAp = to_std(dev_dot<double>(p.p(), Ap.p(), n, st));
        const cxd a = rz / pAp;
        axpy_h(x, p.p(), a, n, st);
        axpy_h(r.p(), Ap.p(), -a, n, st);
        it++;
        rn = dev_norm(r.p(), n, st);
        if (rn / bnorm < tol) break;
        project();
        const cxd rz2 = to_std(dev_dot<double>(r.p(), z.p(), n, st));
        const cxd be = rz2 / rz;
        rz = rz2;
        launch_axpby<double>(p.p(), z.p(), c64{1.0, 0.0}, c64{be.real(), be.imag()}, n, st);
    }
    res.iterations = it;
    res.final_residual = rn / bnorm;
    return res;
}


// ----------------------------------------------------------------- Rayleigh-Ritz, low modes
namespace {
// host copies of <V_i, w> for the m device vectors V (stride n)
std::vector<cxd> dots_h(const c64* V, int m, const c64* w, size_t n, hipStream_t st) {
    DevMem out(64 * (size_t)((m + 31) / 32) * sizeof(double));
    Context& cx = Context::get();
    for (int c0 = 0; c0 < m; c0 += 32) {
        const int nv = std::min(32, m - c0);
        VecPtrs<double> P{};
        for (int i = 0; i < nv; i++) P.v[i] = V + (size_t)(c0 + i) * n;
        launch_multi_dot<double>(P, nv, w, n, 0, cx.fin(FIN_STORE, out.as<double>() + 2 * c0), st);
    }
    std::vector<double> h(64 * (size_t)((m + 31) / 32));
    SMG_HIP(hipMemcpyAsync(h.data(), out.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    SMG_HIP(hipStreamSynchronize(st));
    std::vector<cxd> d(m);
    for (int i = 0; i < m; i++) d[i] = cxd(h[2 * i], h[2 * i + 1]);
    return d;
}
}  // namespace

void rr_evolve(LinOp<double>& A, Eigenpairs& E, hipStream_t st) {
    const int k = (int)E.vals.size();
    if (k == 0) return;
    const size_t n = A.n();
    Field<double> AV((size_t)k * n), W((size_t)k * n);
    for (int j = 0; j < k; j++) A.apply(AV.p((size_t)j * n), E.vecs.p((size_t)j * n), st);
    std::vector<c64> M((size_t)k * k);   // columns: M[j*k + i] = M_ij
    for (int j = 0; j < k; j++) {
        const std::vector<cxd> d = dots_h(E.vecs.p(), j + 1, AV.p((size_t)j * n), n, st);
        for (int i = 0; i <= j; i++) {
            M[(size_t)j * k + i] = c64{d[i].real(), d[i].imag()};
            if (i != j) M[(size_t)i * k + j] = c64{d[i].real(), -d[i].imag()};
        }
    }
    std::vector<double> ev(k);
    std::vector<c64> Z((size_t)k * k);   // Z[row*k + col]
    dense_eigh(k, M.data(), ev.data(), Z.data());
    DevMem coef(64 * (size_t)((k + 31) / 32) * sizeof(double));
    std::vector<cxd> c(k);
    for (int i = 0; i < k; i++) {   // v_i <- sum_j Z[j][i] v_j
        for (int j = 0; j < k; j++) c[j] = to_std(Z[(size_t)j * k + i]);
        combine(W.p((size_t)i * n), E.vecs.p(), k, c, n, coef, st);
    }
    E.vecs = std::move(W);
    E.vals = ev;
}

void lowmode_solution(const Aggregation& P, const Eigenpairs& E, const c64* v, c64* out, hipStream_t st) {
    const size_t nc = P.coarse_n();
    const int k = (int)E.vals.size();
    Field<double> pc(nc), xc(nc);
    launch_restrict<double>(P.A, P.bv<double>(), v, pc.p(), st);
    launch_zero<double>(xc.p(), nc, st);
    if (k > 0) {
        std::vector<cxd> c = dots_h(E.vecs.p(), k, pc.p(), nc, st);
        for (int i = 0; i < k; i++) c[i] = E.vals[i] > 1e-14 ? c[i] / E.vals[i] : cxd(0.0);
        DevMem coef(64 * (size_t)((k + 31) / 32) * sizeof(double));
        combine(xc.p(), E.vecs.p(), k, c, nc, coef, st);
    }
    launch_prolong<double>(P.A, P.bv<double>(), xc.p(), out, 0, st);
}

// ----------------------------------------------------------------- LOBPCG
namespace {
// Block-vector workspace of the LOBPCG iteration and its host-side Gram helper.
struct BlockWork {
    size_t n;
    int N, k;   // block width and the number of live columns
    DevMem part, gout, coef;
    BlockWork(size_t n_, int N_, int k_)
        : n(n_), N(N_), k(k_),
          part((size_t)blk_gram_grid(n_) * 2 * kGramPairsMax * sizeof(double)),
          gout(2 * (size_t)kGramPairsMax * sizeof(double)),
          coef(2 * 3 * (size_t)N_ * N_ * sizeof(c64)) {}
    static constexpr int kGramPairsMax = 4096;
    // G[a][b] = <L_a, R_b>, a over the nl*N left columns, b over the nr*N right ones
    std::vector<cxd> gram(const std::vector<const c64*>& L, const std::vector<const c64*>& R, hipStream_t st) {
        const int nl = (int)L.size(), nr = (int)R.size(), CL = nl * N, CR = nr * N;
        std::vector<cxd> G((size_t)CL * CR);
        BlockPtrs lp{};
        for (int i = 0; i < nl; i++) lp.p[i] = L[i];
        int chunk = std::max(1, std::min(nr, kGramPairsMax / (CL * N)));
        while (chunk > 1 && (nl + chunk) * N > 3072) chunk--;
        std::vector<double> h;
        for (int r0 = 0; r0 < nr; r0 += chunk) {
            const int nc = std::min(chunk, nr - r0);
            BlockPtrs rp{};
            for (int i = 0; i < nc; i++) rp.p[i] = R[r0 + i];
            launch_blk_gram(lp, nl, rp, nc, N, n, part.as<double>(), gout.as<double>(), st);
            h.resize(2 * (size_t)CL * nc * N);
            SMG_HIP(hipMemcpyAsync(h.data(), gout.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, st));
            SMG_HIP(hipStreamSynchronize(st));
            for (int a = 0; a < CL; a++)
                for (int b = 0; b < nc * N; b++)
                    G[(size_t)a * CR + r0 * N + b] = cxd(h[2 * ((size_t)a * nc * N + b)], h[2 * ((size_t)a * nc * N + b) + 1]);
        }
        return G;
    }
    // out0 = sum_s S_s C0[s], out1 = sum_s S_s C1[s] (host coefficient matrices
    // (ns*N) x N, row-major; C1 empty for one output)
    void combine(c64* out0, c64* out1, const std::vector<const c64*>& S, const std::vector<cxd>& C0,
                 const std::vector<cxd>& C1, hipStream_t st) {
        const int ns = (int)S.size();
        const size_t m = (size_t)ns * N * N;
        std::vector<c64> h(2 * m);
        for (size_t i = 0; i < m; i++) {
            h[i] = c64{C0[i].real(), C0[i].imag()};
            if (!C1.empty()) h[m + i] = c64{C1[i].real(), C1[i].imag()};
        }
        SMG_HIP(hipMemcpyAsync(coef.p, h.data(), (C1.empty() ? m : 2 * m) * sizeof(c64), hipMemcpyHostToDevice, st));
        BlockPtrs sp{};
        for (int i = 0; i < ns; i++) sp.p[i] = S[i];
        launch_blk_combine(out0, out1, sp, ns, N, k, coef.as<c64>(), C1.empty() ? nullptr : coef.as<c64>() + m, n,
                           st);
        SMG_HIP(hipStreamSynchronize(st));   // the host coefficients go out of scope
    }
    // per-column y_j = a_j x_j + b_j y_j
    void cols(c64* y, const c64* x, const std::vector<cxd>& a, const std::vector<cxd>& b, hipStream_t st) {
        std::vector<c64> h(2 * (size_t)N);
        for (int j = 0; j < N; j++) {
            h[j] = c64{a[j].real(), a[j].imag()};
            h[N + j] = c64{b[j].real(), b[j].imag()};
        }
        SMG_HIP(hipMemcpyAsync(coef.p, h.data(), h.size() * sizeof(c64), hipMemcpyHostToDevice, st));
        launch_blk_cols(y, x, coef.as<c64>(), coef.as<c64>() + N, N, n, st);
        SMG_HIP(hipStreamSynchronize(st));
    }
};
}  // namespace

LobpcgResult lobpcg(LinOpB<double>& A, MG* mg, int k, const c64* X0, int max_iter, double tol, hipStream_t st,
                    bool exact_ax) {
    const size_t n = A.n();
    const int N = A.N;
    SMG_REQUIRE(k >= 1 && k <= N && N <= 32 && (N & (N - 1)) == 0, "lobpcg: 1 <= k <= N <= 32, N a power of two");
    SMG_REQUIRE((size_t)k <= n, "lobpcg: k exceeds the dimension");
    BlockWork bw(n, N, k);
    const size_t bn = n * N;
    Field<double> X(bn), AX(bn), W(bn), AW(bn), P(bn), AP(bn), Xn(bn), AXn(bn), Pn(bn), APn(bn), R(bn), sep(bn);
    launch_zero<double>(sep.p(), bn, st);
    launch_copy<double>(sep.p(), X0, (size_t)k * n, st);
    launch_b_pack<double, double>(X.p(), sep.p(), n, N, st);
    std::vector<cxd> one(N, 1.0), zero(N, 0.0);
    auto ident = [&](int ns) {   // (ns*N) x N zero matrix
        return std::vector<cxd>((size_t)ns * N * N, 0.0);
    };
    // X <- X R^-1 with R from the Cholesky factor of X^dag X (the reference's modified
    // Gram-Schmidt, src/eigensolver_core.cpp:300-307 and :510-523); with AX, also
    // AX <- X R^-1 + (AX - X) diag(1/R_ii), the reference's update of AX
    auto chol_qr = [&](bool with_ax) {
        std::vector<cxd> G = bw.gram({X.p()}, {X.p()}, st);
        std::vector<cxd> Rm((size_t)k * k, 0.0);
        for (int i = 0; i < k; i++) {
            cxd d = G[(size_t)i * N + i];
            for (int p = 0; p < i; p++) d -= std::norm(Rm[(size_t)p * k + i]);
            const double rii = std::sqrt(std::max(d.real(), 0.0));
            Rm[(size_t)i * k + i] = rii > 1e-14 ? rii : 1.0;   // the reference leaves such a vector unscaled
            for (int j = i + 1; j < k; j++) {
                cxd s = G[(size_t)i * N + j];
                for (int p = 0; p < i; p++) s -= std::conj(Rm[(size_t)p * k + i]) * Rm[(size_t)p * k + j];
                Rm[(size_t)i * k + j] = s / Rm[(size_t)i * k + i];
            }
        }
        std::vector<cxd> Ri((size_t)k * k, 0.0);   // R^-1, upper triangular
        for (int c = 0; c < k; c++) {
            Ri[(size_t)c * k + c] = 1.0 / Rm[(size_t)c * k + c];
            for (int i = c - 1; i >= 0; i--) {
                cxd s = 0.0;
                for (int p = i + 1; p <= c; p++) s += Rm[(size_t)i * k + p] * Ri[(size_t)p * k + c];
                Ri[(size_t)i * k + c] = -s / Rm[(size_t)i * k + i];
            }
        }
        std::vector<cxd> C0 = ident(with_ax ? 2 : 1), C1;
        for (int j = 0; j < k; j++)
            for (int c = 0; c < k; c++) C0[(size_t)j * N + c] = Ri[(size_t)j * k + c];
        if (with_ax) {
            C1 = ident(2);
            if (exact_ax) {   // AX <- AX R^-1
                for (int j = 0; j < k; j++)
                    for (int c = 0; c < k; c++) C1[((size_t)N + j) * N + c] = Ri[(size_t)j * k + c];
            } else {
                for (int j = 0; j < k; j++)
                    for (int c = 0; c < k; c++) C1[(size_t)j * N + c] = Ri[(size_t)j * k + c];
                for (int c = 0; c < k; c++) {
                    const double inv = 1.0 / Rm[(size_t)c * k + c].real();
                    C1[(size_t)c * N + c] -= inv;
                    C1[((size_t)N + c) * N + c] = inv;
                }
            }
            bw.combine(Xn.p(), AXn.p(), {X.p(), AX.p()}, C0, C1, st);
            std::swap(X, Xn);
            std::swap(AX, AXn);
        } else {
            bw.combine(Xn.p(), nullptr, {X.p()}, C0, {}, st);
            std::swap(X, Xn);
        }
    };
    chol_qr(false);
    A.apply(AX.p(), X.p(), st);
    bool have_P = false;
    int iter;
    for (iter = 0; iter < max_iter; iter++) {
        std::vector<cxd> G = bw.gram({X.p()}, {AX.p()}, st);
        std::vector<cxd> mth(N, 0.0);
        std::vector<double> theta(k);
        for (int i = 0; i < k; i++) {
            theta[i] = G[(size_t)i * N + i].real();
            mth[i] = -theta[i];
        }
        launch_copy<double>(R.p(), AX.p(), bn, st);
        bw.cols(R.p(), X.p(), mth, one, st);
        std::vector<cxd> GR = bw.gram({R.p()}, {R.p()}, st);
        double max_r = 0.0;
        for (int i = 0; i < k; i++)
            max_r = std::max(max_r, std::sqrt(GR[(size_t)i * N + i].real()) / std::max(std::abs(theta[i]), 1e-30));
        if (max_r < tol) break;
        if (mg) mg_precondition_b<double>(*mg, R.p(), W.p(), N, st);
        else launch_copy<double>(W.p(), R.p(), bn, st);
        A.apply(AW.p(), W.p(), st);
        const int nb = have_P ? 3 : 2, sdim = nb * k;
        std::vector<const c64*> S = {X.p(), W.p(), P.p()}, AS = {AX.p(), AW.p(), AP.p()};
        S.resize(nb);
        AS.resize(nb);
        std::vector<cxd> GM = bw.gram(S, S, st), GH = bw.gram(S, AS, st);
        const int CL = nb * N;
        auto at = [&](const std::vector<cxd>& Gm, int i, int j) {   // <S_i, T_j> in the reference's indexing
            return Gm[(size_t)((i / k) * N + i % k) * CL + (j / k) * N + j % k];
        };
        // H_cols[j][i] = <S_i, A S_j>, M_cols[j][i] = <S_i, S_j>, then Hermitised
        std::vector<cxd> H((size_t)sdim * sdim), M((size_t)sdim * sdim);
        for (int j = 0; j < sdim; j++)
            for (int i = 0; i < sdim; i++) {
                H[(size_t)j * sdim + i] = at(GH, i, j);
                M[(size_t)j * sdim + i] = at(GM, i, j);
            }
        std::vector<cxd> xAw(k), wAw(k);
        for (int i = 0; i < k; i++) {
            xAw[i] = H[(size_t)(k + i) * sdim + i];
            wAw[i] = H[(size_t)(k + i) * sdim + k + i];
        }
        for (int i = 0; i < sdim; i++)
            for (int j = i + 1; j < sdim; j++) {
                const cxd ah = 0.5 * (H[(size_t)j * sdim + i] + std::conj(H[(size_t)i * sdim + j]));
                H[(size_t)j * sdim + i] = ah;
                H[(size_t)i * sdim + j] = std::conj(ah);
                const cxd am = 0.5 * (M[(size_t)j * sdim + i] + std::conj(M[(size_t)i * sdim + j]));
                M[(size_t)j * sdim + i] = am;
                M[(size_t)i * sdim + j] = std::conj(am);
            }
        // M = L L^dag, L[j][i] (i <= j) stored Lc[j*sdim + i]
        std::vector<cxd> Lc((size_t)sdim * sdim, 0.0);
        bool chol_ok = true;
        for (int j = 0; j < sdim && chol_ok; j++) {
            for (int i = 0; i < j; i++) {
                cxd s = M[(size_t)j * sdim + i];
                for (int p = 0; p < i; p++) s -= Lc[(size_t)i * sdim + p] * std::conj(Lc[(size_t)j * sdim + p]);
                Lc[(size_t)j * sdim + i] = s / Lc[(size_t)i * sdim + i];
            }
            cxd s = M[(size_t)j * sdim + j];
            for (int p = 0; p < j; p++) s -= Lc[(size_t)j * sdim + p] * std::conj(Lc[(size_t)j * sdim + p]);
            if (s.real() < 1e-14) chol_ok = false;
            else Lc[(size_t)j * sdim + j] = std::sqrt(s.real());
        }
        if (!chol_ok) {   // one preconditioned steepest-descent step per vector (:396-423)
            std::vector<cxd> al(N, 0.0);
            for (int i = 0; i < k; i++)
                if (std::abs(wAw[i]) > 1e-30) al[i] = -xAw[i] / wAw[i];
            bw.cols(X.p(), W.p(), al, one, st);
            std::vector<cxd> GX = bw.gram({X.p()}, {X.p()}, st);
            std::vector<cxd> sc(N, 1.0);
            for (int i = 0; i < k; i++) {
                const double nv = std::sqrt(GX[(size_t)i * N + i].real());
                if (nv > 1e-14) sc[i] = 1.0 / nv;
            }
            bw.cols(X.p(), nullptr, zero, sc, st);
            chol_qr(false);
            A.apply(AX.p(), X.p(), st);
            have_P = false;
            continue;
        }
        // Ht = L^-1 H L^-dag, Hermitised, diagonalised
        std::vector<cxd> Y((size_t)sdim * sdim), Ht((size_t)sdim * sdim), z(sdim);
        for (int col = 0; col < sdim; col++) {
            for (int i = 0; i < sdim; i++) {
                cxd s = H[(size_t)col * sdim + i];
                for (int p = 0; p < i; p++) s -= Lc[(size_t)i * sdim + p] * Y[(size_t)col * sdim + p];
                Y[(size_t)col * sdim + i] = s / Lc[(size_t)i * sdim + i];
            }
            for (int i = sdim - 1; i >= 0; i--) {
                cxd s = Y[(size_t)col * sdim + i];
                for (int p = i + 1; p < sdim; p++) s -= std::conj(Lc[(size_t)p * sdim + i]) * z[p];
                z[i] = s / Lc[(size_t)i * sdim + i];
            }
            for (int i = 0; i < sdim; i++) Ht[(size_t)col * sdim + i] = z[i];
        }
        for (int i = 0; i < sdim; i++)
            for (int j = i + 1; j < sdim; j++) {
                const cxd a = 0.5 * (Ht[(size_t)j * sdim + i] + std::conj(Ht[(size_t)i * sdim + j]));
                Ht[(size_t)j * sdim + i] = a;
                Ht[(size_t)i * sdim + j] = std::conj(a);
            }
        std::vector<c64> Hc((size_t)sdim * sdim), Ev((size_t)sdim * sdim);
        for (size_t i = 0; i < Hc.size(); i++) Hc[i] = c64{Ht[i].real(), Ht[i].imag()};
        std::vector<double> ev(sdim);
        dense_eigh(sdim, Hc.data(), ev.data(), Ev.data());   // Ev[row*sdim + col]
        // y = L^-dag z for the k lowest
        std::vector<cxd> Yb((size_t)k * sdim, 0.0);
        for (int col = 0; col < k; col++)
            for (int i = sdim - 1; i >= 0; i--) {
                cxd s = to_std(Ev[(size_t)i * sdim + col]);
                for (int p = i + 1; p < sdim; p++) s -= std::conj(Lc[(size_t)p * sdim + i]) * Yb[(size_t)col * sdim + p];
                Yb[(size_t)col * sdim + i] = s / Lc[(size_t)i * sdim + i];
            }
        // X_new = S y, P = X_new - X (and their A images), one launch each for S and AS
        std::vector<cxd> CX = ident(nb), CP;
        for (int c = 0; c < k; c++)
            for (int j = 0; j < sdim; j++) CX[((size_t)(j / k) * N + j % k) * N + c] = Yb[(size_t)c * sdim + j];
        CP = CX;
        for (int c = 0; c < k; c++) CP[(size_t)c * N + c] -= 1.0;
        bw.combine(Xn.p(), Pn.p(), S, CX, CP, st);
        bw.combine(AXn.p(), APn.p(), AS, CX, CP, st);
        std::vector<cxd> GP = bw.gram({Pn.p()}, {Pn.p()}, st);
        std::vector<cxd> sc(N, 1.0);
        for (int i = 0; i < k; i++) {
            const double np = std::sqrt(GP[(size_t)i * N + i].real());
            if (np > 1e-14) sc[i] = 1.0 / np;
        }
        bw.cols(Pn.p(), nullptr, zero, sc, st);
        bw.cols(APn.p(), nullptr, zero, sc, st);
        std::swap(X, Xn);
        std::swap(AX, AXn);
        std::swap(P, Pn);
        std::swap(AP, APn);
        have_P = true;
        chol_qr(true);
    }
    LobpcgResult out;
    out.iterations = iter;
    std::vector<cxd> G = bw.gram({X.p()}, {AX.p()}, st);
    out.vals.resize(k);
    for (int i = 0; i < k; i++) out.vals[i] = G[(size_t)i * N + i].real();
    launch_b_unpack<double, double>(sep.p(), X.p(), n, N, st);
    out.vecs = Field<double>((size_t)k * n);
    launch_copy<double>(out.vecs.p(), sep.p(), (size_t)k * n, st);
    SMG_HIP(hipStreamSynchronize(st));
    return out;
}

// ----------------------------------------------------------------- FGMRES Ritz harvest
RitzHarvest::RitzHarvest(int n_ritz_, int restart, size_t n_)
    : n_ritz(n_ritz_), n(n_), vecs((size_t)3 * std::max(n_ritz_, 1) * n_),
      coef(64 * (size_t)((std::max(3 * n_ritz_, restart) + 31) / 32) * sizeof(double)) {}

void RitzHarvest::deflate(const c64* v, c64* z, hipStream_t st) {
    const int m = count();
    if (m == 0) return;
    std::vector<cxd> c = dots_h(vecs.p(), m, v, n, st);
    std::vector<double> h(2 * (size_t)m);
    for (int q = 0; q < m; q++) {
        const cxd a = vals[q] > 1e-14 ? c[q] / vals[q] : cxd(0.0);
        h[2 * q] = a.real();
        h[2 * q + 1] = a.imag();
    }
    SMG_HIP(hipMemcpyAsync(coef.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, st));
    for (int c0 = 0; c0 < m; c0 += 32) {
        const int nv = std::min(32, m - c0);
        VecPtrs<double> P{};
        for (int i = 0; i < nv; i++) P.v[i] = vecs.p((size_t)(c0 + i) * n);
        launch_lincomb<double>(z, P, nv, coef.as<double>() + 2 * c0, 1, n, st);
    }
    SMG_HIP(hipStreamSynchronize(st));
}

void RitzHarvest::harvest(const std::vector<c64>& Hraw, int ld, int kk, const c64* Z, hipStream_t st) {
    if (n_ritz <= 0 || kk < n_ritz) return;
    auto hr = [&](int p, int i) { return to_std(Hraw[(size_t)p * ld + i]); };
    std::vector<cxd> G((size_t)kk * kk, 0.0);   // G[j*kk + i] = sum_p conj(H~_pi) H~_pj
    for (int j = 0; j < kk; j++)
        for (int i = 0; i < kk; i++) {
            cxd s = 0.0;
            for (int p = 0; p <= kk; p++) s += std::conj(hr(p, i)) * hr(p, j);
            G[(size_t)j * kk + i] = s;
        }
    for (int i = 0; i < kk; i++)
        for (int j = i + 1; j < kk; j++) {
            const cxd av = 0.5 * (G[(size_t)j * kk + i] + std::conj(G[(size_t)i * kk + j]));
            G[(size_t)j * kk + i] = av;
            G[(size_t)i * kk + j] = std::conj(av);
        }
    std::vector<c64> Gc((size_t)kk * kk), E((size_t)kk * kk);
    for (size_t i = 0; i < Gc.size(); i++) Gc[i] = c64{G[i].real(), G[i].imag()};
    std::vector<double> ev(kk);
    dense_eigh(kk, Gc.data(), ev.data(), E.data());
    const int nh = std::min(n_ritz, kk);
    std::vector<cxd> c(kk);
    for (int p = 0; p < nh; p++) {
        const int q = count();
        c64* out = vecs.p((size_t)q * n);
        for (int j = 0; j < kk; j++) c[j] = to_std(E[(size_t)j * kk + p]);
        combine(out, Z, kk, c, n, coef, st);
        const double nv = dev_norm(out, n, st);
        if (nv > 1e-14) launch_axpby<double>(out, out, c64{0, 0}, c64{1.0 / nv, 0}, n, st);
        vals.push_back(ev[p]);
    }
    if (count() > 2 * n_ritz) {   // keep the n_ritz smallest, in ascending order
        std::vector<int> idx(count());
        std::iota(idx.begin(), idx.end(), 0);
        std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return vals[a] < vals[b]; });
        Field<double> kept((size_t)3 * n_ritz * n);
        std::vector<double> kv(n_ritz);
        for (int q = 0; q < n_ritz; q++) {
            launch_copy<double>(kept.p((size_t)q * n), vecs.p((size_t)idx[q] * n), n, st);
            kv[q] = vals[idx[q]];
        }
        vecs = std::move(kept);
        vals = kv;
    }
    SMG_HIP(hipStreamSynchronize(st));
}

}  // namespace smg
