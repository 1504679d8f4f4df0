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

}  // namespace smg
