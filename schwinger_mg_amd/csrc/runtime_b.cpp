// runtime_b.cpp -- sequencing of the multi-right-hand-side solve path.
#include "runtime_b.h"
#include "tune.h"

#include <algorithm>
#include <cmath>
#include <deque>
#include <functional>

namespace smg {

namespace {
FinB finb(int op, double* out, int ostride, double* aux = nullptr, int astride = 16) {
    const Fin f = Context::get().fin(op, out, aux);
    FinB b;
    b.op = op;
    b.out = out;
    b.ostride = ostride;
    b.aux = aux;
    b.astride = astride;
    b.part = f.part;
    b.counter = f.counter;
    return b;
}
struct ScalarsB {
    DevMem m;
    explicit ScalarsB(size_t n) : m(n * sizeof(double)) {
        SMG_HIP(hipMemsetAsync(m.p, 0, n * sizeof(double), Context::get().stream));
    }
    double* p(size_t i = 0) const { return m.as<double>() + i; }
};
template <typename T>
void bcopy(cplx<T>* d, const cplx<T>* s, size_t n, hipStream_t st) {
    launch_copy<T>(d, s, n, st);
}
}  // namespace

template <typename T>
void DdagDOpB<T>::apply(cplx<T>* dst, const cplx<T>* src, Epi epi, const cplx<T>* b, const cplx<T>* aux,
                        const FinB& fin, hipStream_t st) {
    const DiracView<T> v = D.view<T>();
    const T mu = (T)D.mu_t;
    if (launch_b_ddagd<T>(v, dst, src, mu, epi, b, aux, fin, this->N, st)) return;
    launch_b_dslash<T>(v, tmp.p(), src, mu, 0, 0, EPI_NONE, nullptr, nullptr, FinB(), this->N, st);
    launch_b_dslash<T>(v, dst, tmp.p(), -mu, 1, 1, epi, b, aux, fin, this->N, st);
}
template <typename T>
void DdagDOpB<T>::cheb_step(cplx<T>* d_new, const cplx<T>* d, const cplx<T>* r, cplx<T>* x, const FinB& f,
                            hipStream_t st) {
    SMG_REQUIRE(fused_cheb(), "Chebyshev step: the lattice or N does not fit the fused D^dag D");
    const DiracView<T> v = D.view<T>();
    SMG_REQUIRE(launch_b_ddagd<T>(v, d_new, d, (T)D.mu_t, EPI_CHEB, r, x, f, this->N, st),
                "Chebyshev step: fused D^dag D refused");
}
template struct DdagDOpB<double>;
template struct DdagDOpB<float>;

template <typename T, typename TL>
void CoarseLinOpB<T, TL>::apply(cplx<T>* dst, const cplx<T>* src, Epi epi, const cplx<T>* b, const cplx<T>* aux,
                                const FinB& fin, hipStream_t st) {
    const TL* link;
    if constexpr (std::is_same<TL, c64>::value) link = C.l64.p();
    else if constexpr (std::is_same<TL, c32>::value) link = C.l32.p();
    else link = C.l16.template as<ch16>();
    const int herm = std::is_same<TL, ch16>::value ? C.herm16 : 0;
    launch_b_coarse_apply<T, TL>(CoarseView<TL>{link, C.nx, C.ny, C.k, herm}, dst, src, epi, b, aux, fin, this->N,
                                 st);
}
template <typename T, typename TL>
void CoarseLinOpB<T, TL>::cheb_step(cplx<T>* d_new, const cplx<T>* d, const cplx<T>* r, cplx<T>* x, const FinB& f,
                                    hipStream_t st) {
    SMG_REQUIRE(fused_cheb(), "Chebyshev step: the coarse operator's fused step needs FP16 links, FP32 vectors");
    apply(d_new, d, EPI_CHEB, r, x, f, st);
}
template struct CoarseLinOpB<double, c64>;
template struct CoarseLinOpB<float, c32>;
template struct CoarseLinOpB<float, ch16>;

// ------------------------------------------------------------------ Galerkin build (batched probing)
void CoarseOperator::build_b(LinOpB<double>& A, const Aggregation& P, hipStream_t st) {
    SMG_REQUIRE(P.A.nbx == nx && P.A.nby == ny && P.A.k == k, "coarse operator / aggregation mismatch");
    hermitian = true;
    SMG_REQUIRE(A.n() == P.fine_n(), "operator / aggregation mismatch");
    const int N = A.N;
    const size_t nc = n(), nf = P.fine_n();
    launch_zero<double>(l64.p(), l64.n, st);
    Field<double> sep(nc * N), eb(nc * N), fb(nf * N), afb(nf * N), cb(nc * N);
    launch_zero<double>(sep.p(), nc * N, st);   // columns past the last probe stay zero
    const int cx = color_period(nx), cy = color_period(ny);
    const int total = cx * cy * k;   // probes (class, kj), class-major as in build
    for (int p0 = 0; p0 < total; p0 += N) {
        const int m = std::min(N, total - p0);
        for (int q = 0; q < m; q++)
            launch_probe_vector(nx, ny, k, cx, cy, (p0 + q) / k, (p0 + q) % k, sep.p((size_t)q * nc), st);
        launch_b_pack<double, double>(eb.p(), sep.p(), nc, N, st);
        launch_b_prolong<double>(P.A, P.bv64.p(), eb.p(), fb.p(), 0, N, st);
        A.apply(afb.p(), fb.p(), st);
        launch_b_restrict<double>(P.A, P.bv64.p(), afb.p(), cb.p(), N, st);
        launch_b_unpack<double, double>(sep.p(), cb.p(), nc, N, st);
        for (int q = 0; q < m; q++)
            launch_probe_extract(nx, ny, k, cx, cy, (p0 + q) / k, (p0 + q) % k, sep.p((size_t)q * nc), l64.p(), st);
    }
    refresh_low_precision(st);
    SMG_HIP(hipStreamSynchronize(st));
}

// ------------------------------------------------------------------ smoothers (batched)
// MR, src/smoother.cpp:26-40, with the residual carried recursively; per-RHS omega.
// From x = 0 the first sweep reads b in place of r = b (no zero fill, no copy); with
// need_r false (the post-smoothing, whose residual is not read again) the last sweep
// updates x only.  On return r = b - A x when need_r.
template <typename T>
static void mr_b(LinOpB<T>& A, cplx<T>* x, const cplx<T>* b, cplx<T>* r, cplx<T>* Ar, int niter, bool zero_init,
                 bool need_r, double* w, hipStream_t st) {
    const size_t n = A.n(), nN = n * A.N;
    const FinB f = finb(FIN_MR, w, 2);
    int it = 0;
    if (zero_init) {
        if (niter == 0) {
            launch_zero<T>(x, nN, st);
            bcopy(r, b, nN, st);
            return;
        }
        A.apply(Ar, b, EPI_DOTS, nullptr, b, f, st);
        launch_b_mr_update<T>(x, r, Ar, w, n, A.N, st, 1, b);   // x = w b, r = b - w Ar
        it = 1;
    } else {
        A.apply(r, x, EPI_RESID, b, nullptr, FinB(), st);
    }
    for (; it < niter; it++) {
        A.apply(Ar, r, EPI_DOTS, nullptr, r, f, st);
        launch_b_mr_update<T>(x, r, Ar, w, n, A.N, st, it + 1 == niter && !need_r ? 2 : 0);
    }
}

// Richardson, src/smoother.cpp:63-73; on return r = b - A x when need_r
template <typename T>
static void rich_b(LinOpB<T>& A, cplx<T>* x, const cplx<T>* b, cplx<T>* r, int niter, double omega, bool zero_init,
                   bool need_r, hipStream_t st) {
    const size_t nN = A.n() * A.N;
    if (zero_init) launch_zero<T>(x, nN, st);
    for (int it = 0; it < niter; it++) {
        if (it == 0 && zero_init) bcopy(r, b, nN, st);
        else A.apply(r, x, EPI_RESID, b, nullptr, FinB(), st);
        launch_b_axpby_real<T>(x, r, omega, 1.0, nN, st);
    }
    if (!need_r) return;
    if (niter == 0 && zero_init) bcopy(r, b, nN, st);
    else A.apply(r, x, EPI_RESID, b, nullptr, FinB(), st);
}

// Chebyshev smoothing on [lo, hi] (the three-term recurrence of the coarsest Chebyshev
// solve, cheb_h): niter applications of A.  From x = 0 (zero_init) r starts as b, else
// r = b - A x; with need_r, r = b - A x on return.  Returns 0, or s != 0 when x was left
// unwritten and is s b (a single fused step from x = 0 when the caller allows it with
// implicit_x; the caller then forms x itself).
template <typename T>
static double cheb_smooth_b(LinOpB<T>& A, cplx<T>* x, const cplx<T>* b, cplx<T>* r, cplx<T>* d, cplx<T>* q,
                            int niter, bool zero_init, bool need_r, double lo, double hi, bool fused, hipStream_t st,
                            bool implicit_x = false) {
    const size_t nN = A.n() * A.N;
    const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma = theta / delta;
    double rho = 1.0 / sigma;
    if (niter > 0 && fused && A.fused_cheb() && zero_init == need_r) {
        // Each step one launch of A with the updates in its epilogue.  The first one also
        // does the initial pass: its direction and residual are b (from x = 0) or the
        // residual r0 = b - A x0 (in r), scaled by 1 / theta, and its new residual goes to
        // r (from x = 0, where r stays) or q (r0 is still read there).  The directions
        // take whichever of r, d, q holds neither the residual nor the current direction.
        const cplx<T>* dc = b;
        const cplx<T>* rc = b;
        cplx<T>* rn = r;
        if (!zero_init) {
            A.apply(r, x, EPI_RESID, b, nullptr, FinB(), st);
            dc = rc = r;
            rn = q;
        }
        cplx<T>* const buf[3] = {r, d, q};
        // from x0 with two steps, the first leaves x alone and the second adds r0 / theta
        // and both directions (r0 stays in r: the last step writes no direction)
        const bool defer_x = !zero_init && niter == 2;
        const bool skip_x = zero_init && niter == 1 && implicit_x;   // x = b / theta, unwritten
        for (int k = 1; k <= niter; k++) {
            const double rho_k = 1.0 / (2.0 * sigma - rho);
            FinB f;
            f.cheb_c1 = rho_k * rho;
            f.cheb_c2 = 2.0 * rho_k / delta;
            f.cheb_mode = k < niter ? 0 : (need_r ? 1 : 2);
            if (k == 1) {
                f.cheb_start = zero_init ? 1 : 2;
                f.cheb_s = 1.0 / theta;
                f.cheb_rout = rn;
                if (defer_x || skip_x) f.cheb_xacc = 1;
            } else if (k == 2 && defer_x) {
                f.cheb_xacc = 2;
                f.cheb_xadd = r;
                f.cheb_xadd_s = 1.0 / theta;
            }
            cplx<T>* dn = nullptr;
            for (cplx<T>* c : buf)
                if (c != rn && c != dc && (k == 1 || c != rc)) {
                    dn = c;
                    break;
                }
            A.cheb_step(dn, dc, rc, x, f, st);
            if (k == 1) rc = rn;
            dc = dn;
            rho = rho_k;
        }
        return skip_x ? 1.0 / theta : 0.0;
    }
    if (zero_init) {
        launch_zero<T>(x, nN, st);
        bcopy(r, b, nN, st);
    } else {
        A.apply(r, x, EPI_RESID, b, nullptr, FinB(), st);
    }
    if (niter == 0) return 0.0;
    launch_zero<T>(d, nN, st);
    launch_b_axpby_real<T>(d, r, 1.0 / theta, 1.0, nN, st);   // d_0 = r / theta
    launch_b_axpby_real<T>(x, d, 1.0, 1.0, nN, st);           // x_1 = x_0 + d_0
    for (int k = 1; k <= niter; k++) {
        A.apply(q, d, st);
        launch_b_axpby_real<T>(r, q, -1.0, 1.0, nN, st);      // r_k = r_{k-1} - A d_{k-1}
        if (k == niter && need_r) break;
        const double rho_k = 1.0 / (2.0 * sigma - rho);
        launch_b_axpby_real<T>(d, r, 2.0 * rho_k / delta, rho_k * rho, nN, st);
        launch_b_axpby_real<T>(x, d, 1.0, 1.0, nN, st);
        rho = rho_k;
    }
    return 0.0;
}

// CG per RHS (src/coarse_op.h:492-547 without deflation); converged RHS take zero steps
template <typename T>
static void cg_b(LinOpB<T>& A, const cplx<T>* b, cplx<T>* x, cplx<T>* r, cplx<T>* p, cplx<T>* Ap, int maxit,
                 double tol, double* sc, hipStream_t st) {
    const size_t n = A.n(), nN = n * A.N;
    const int N = A.N;
    double* alpha = sc;             // [N][2]
    double* beta = sc + 2 * 64;     // [N][2]
    double* aux = sc + 4 * 64;      // [N][16]
    launch_zero<T>(x, nN, st);
    bcopy(r, b, nN, st);
    bcopy(p, b, nN, st);
    // once every RHS has converged the remaining iterations' kernels return at
    // entry (device-side flag, no host synchronisation)
    int* done = reinterpret_cast<int*>(aux + 16 * 64);
    int* late = reinterpret_cast<int*>(aux + 16 * 64 + 1);
    launch_b_cg_init(aux, 16, N, tol * tol, done, late, st);
    launch_b_norm2<T>(r, n, finb(FIN_CG_INIT, nullptr, 0, aux), N, st);
    FinB fa = finb(FIN_CG_ALPHA, alpha, 2, aux), fbt = finb(FIN_CG_BETA, beta, 2, aux);
    fa.skip = done;
    fbt.skip = done;
    fbt.done = done;
    for (int it = 0; it < maxit; it++) {
        A.apply(Ap, p, EPI_DOTS, nullptr, p, fa, st);
        launch_b_cg_r<T>(r, Ap, alpha, n, fbt, N, late, st);
        launch_b_cg_xp<T>(x, p, r, alpha, beta, n, N, late, st);
    }
}

// The same CG on the FP16-link coarse operator with BF16 vectors (k_h_* kernels):
// half the vector traffic of cg_b<float>, for the loose coarsest-level tolerance, and
// two launches per iteration: the apply q = A p with the sums <q,p>, <q,r>, |q|^2,
// whose finish computes alpha, the new |r|^2 (from those sums), the stop test and
// beta (FIN_CG_STEP), then one pass r -= alpha q, x += alpha p, p = r + beta p.  The
// iterates are those of cg_b in exact arithmetic.
static void cg_h(const CoarseOperator& C, const c32* b, c32* x, cb16* r, cb16* p, cb16* Ap, cb16* xh, int N,
                 int maxit, double tol, double* sc, hipStream_t st) {
    const size_t n = C.n();
    double* alpha = sc;
    double* aux = sc + 4 * 64;
    int* done = reinterpret_cast<int*>(aux + 16 * 64);
    int* late = reinterpret_cast<int*>(aux + 16 * 64 + 1);
    launch_b_cg_init(aux, 16, N, tol * tol, done, late, st);
    launch_h_cg_init(b, r, p, xh, n, finb(FIN_CG_INIT, nullptr, 0, aux), N, st);
    FinB fs = finb(FIN_CG_STEP, alpha, 2, aux);
    fs.skip = done;
    fs.done = done;
    fs.late = late;
    const CoarseView<ch16> v = C.view16();
    for (int it = 0; it < maxit; it++) {
        launch_b_coarse_apply_h(v, Ap, p, EPI_CG, p, fs, N, st, r);
        launch_h_cg_step(xh, p, r, Ap, alpha, aux, n, N, late, st);
    }
    launch_h_to_f32(x, xh, n, N, st);
}

// The coarsest solve as a Chebyshev iteration on the BF16 vectors (cfg.coarse_solver 2):
// on the spectral interval [lo, hi] of the coarsest operator, found at setup, the
// polynomial needs no inner products, so a step has no reduction and no host read, and
// every step after the first is one coarse apply with its vector updates fused into the
// epilogue (EPI_CHEB).  The three-term recurrence of Saad, Iterative Methods for Sparse
// Linear Systems (2nd ed.), Algorithm 12.1, from x = 0:
//   theta = (hi + lo)/2, delta = (hi - lo)/2, sigma = theta/delta, rho_0 = 1/sigma,
//   r_0 = b, d_0 = b/theta, x_1 = d_0;
//   step k >= 1: r_k = r_{k-1} - A d_{k-1}, rho_k = 1/(2 sigma - rho_{k-1}),
//                d_k = rho_k rho_{k-1} d_{k-1} + (2 rho_k/delta) r_k, x_{k+1} = x_k + d_k.
// steps coarse applies; the directions alternate between d0 and d1.
static void cheb_h(const CoarseOperator& C, const c32* b, c32* x, cb16* r, cb16* d0, cb16* d1, cb16* xh, int N,
                   int steps, double lo, double hi, hipStream_t st) {
    SMG_REQUIRE(hi > lo && lo > 0.0, "Chebyshev coarsest solve: no spectral interval");
    const size_t n = C.n();
    const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma = theta / delta;
    launch_h_cheb_init(b, r, d0, xh, n, N, theta, st);
    const CoarseView<ch16> v = C.view16();
    double rho = 1.0 / sigma;
    cb16* dold = d0;
    cb16* dnew = d1;
    for (int k = 1; k <= steps; k++) {
        const double rho_k = 1.0 / (2.0 * sigma - rho);
        FinB f;
        f.cheb_c1 = rho_k * rho;
        f.cheb_c2 = 2.0 * rho_k / delta;
        // x is read and written every other step (x += d_{k-1} + d_k), and at the last
        f.cheb_xacc = (k % 2 == 1 && k < steps) ? 1 : (k % 2 == 0 ? 2 : 0);
        launch_b_coarse_apply_h(v, dnew, dold, EPI_CHEB, xh, f, N, st, r);
        std::swap(dold, dnew);
        rho = rho_k;
    }
    launch_h_to_f32(x, xh, n, N, st);
}

// ------------------------------------------------------------------ cycle (batched)
template <typename T>
struct CycleB {
    MG& mg;
    int N;
    std::vector<LinOpB<T>*> ops;
    std::vector<std::unique_ptr<LinOpB<T>>> own;
    std::vector<Field<T>> xs, bs, rs, Ars, ts, e2s, ps, cds;   // cds: Chebyshev smoother directions
    std::vector<std::unique_ptr<ScalarsB>> sc;
    int nlev;
    Field<T> dwork;   // column-slice partials of the matrix-core dense coarsest product
    bool hcg = false;  // coarsest CG on BF16 vectors (FP32 cycle, cfg.coarse_half == 2)
    bool hcheb = false;   // coarsest Chebyshev on BF16 vectors (the same, cfg.coarse_solver == 2)
    DevMem hr, hp, hAp, hx;
    // the cycle captured as a hipGraph (cfg.use_graph) on fixed input/output buffers
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    Field<T> gin, gout;
    ~CycleB() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
    }
    CycleB(MG& m, int N_)
        : mg(m), N(N_), nlev(m.cfg.n_levels), dwork(std::max<size_t>(1, b_dense_work(m.dims[m.cfg.n_levels - 1], N_))) {
        SMG_REQUIRE(N >= 1 && N <= 64 && (N & (N - 1)) == 0, "batched solves take a power-of-two count <= 64");
        if constexpr (std::is_same<T, float>::value) {
            const int kc = m.C.empty() ? 0 : m.C.back()->k;
            if (m.cfg.coarse_half == 2 && (m.cfg.coarse_solver == 0 || m.cfg.coarse_solver == 2) && N >= 16 &&
                (kc == 8 || kc == 16 || kc == 32)) {
                hcg = m.cfg.coarse_solver == 0;
                hcheb = m.cfg.coarse_solver == 2;
                const size_t bytes = m.dims[nlev - 1] * N * sizeof(cb16);
                hr.alloc(bytes);
                hp.alloc(bytes);
                hAp.alloc(bytes);
                hx.alloc(bytes);
            }
        }
        own.emplace_back(new DdagDOpB<T>(*m.D, N));
        ops.push_back(own.back().get());
        for (int l = 1; l < nlev; l++) {
            if constexpr (std::is_same<T, double>::value)
                own.emplace_back(new CoarseLinOpB<double, c64>(*m.C[l - 1], N));
            else if (m.cfg.coarse_half)
                own.emplace_back(new CoarseLinOpB<float, ch16>(*m.C[l - 1], N));
            else
                own.emplace_back(new CoarseLinOpB<float, c32>(*m.C[l - 1], N));
            ops.push_back(own.back().get());
        }
        for (int l = 0; l < nlev; l++) {
            const size_t nN = m.dims[l] * N;
            const bool last = (l == nlev - 1);
            xs.emplace_back(nN);
            bs.emplace_back(nN);
            rs.emplace_back(nN);
            Ars.emplace_back(nN);
            ts.emplace_back(l == 0 ? 1 : nN);
            e2s.emplace_back(l == 0 ? 1 : nN);
            ps.emplace_back(last ? nN : 1);
            cds.emplace_back(!last && m.cheb_smooth_ratio > 0.0 ? nN : 1);
            sc.emplace_back(new ScalarsB(8 * 64 + 16 * 64));
        }
    }
    // src/multigrid.cpp:43-102 on N right-hand sides
    void run(int lev, const cplx<T>* b, cplx<T>* x, hipStream_t st) {
        LinOpB<T>& A = *ops[lev];
        const int last = nlev - 1;
        if (lev == last) {
            if (mg.cfg.coarse_solver == 1) {
                const cplx<T>* inv;
                if constexpr (std::is_same<T, double>::value) inv = mg.inv64.p();
                else inv = mg.inv32.p();
                launch_b_dense_apply<T>(inv, b, x, mg.dims[lev], N,
                                        b_dense_work(mg.dims[lev], N) ? dwork.p() : nullptr, st);
            } else if constexpr (std::is_same<T, float>::value) {
                if (hcg)
                    cg_h(*mg.C[lev - 1], b, x, hr.as<cb16>(), hp.as<cb16>(), hAp.as<cb16>(), hx.as<cb16>(), N,
                         mg.cfg.coarse_maxit, mg.cfg.coarse_tol, sc[lev]->p(), st);
                else if (hcheb)
                    cheb_h(*mg.C[lev - 1], b, x, hr.as<cb16>(), hp.as<cb16>(), hAp.as<cb16>(), hx.as<cb16>(), N,
                           mg.cfg.coarse_maxit, mg.cheb_lo, mg.cheb_hi, st);
                else
                    cg_b<T>(A, b, x, rs[lev].p(), ps[lev].p(), Ars[lev].p(), mg.cfg.coarse_maxit, mg.cfg.coarse_tol,
                            sc[lev]->p(), st);
            } else {
                cg_b<T>(A, b, x, rs[lev].p(), ps[lev].p(), Ars[lev].p(), mg.cfg.coarse_maxit, mg.cfg.coarse_tol,
                        sc[lev]->p(), st);
            }
            return;
        }
        const double om = mg.omega[lev];
        const bool cheb = mg.cheb_smooth_ratio > 0.0;
        const double chi = cheb ? mg.cheb_smooth_hi[lev] : 0.0, clo = cheb ? chi / mg.cheb_smooth_ratio : 0.0;
        // a single fused pre-smoothing step leaves x = s b unwritten: the prolongation forms it
        double ximp = 0.0;
        if (cheb)
            ximp = cheb_smooth_b<T>(A, x, b, rs[lev].p(), cds[lev].p(), Ars[lev].p(), mg.cfg.pre, true, true, clo, chi,
                                  mg.cheb_smooth_fused, st, true);
        else if (om > 0) rich_b<T>(A, x, b, rs[lev].p(), mg.cfg.pre, om, true, true, st);
        else mr_b<T>(A, x, b, rs[lev].p(), Ars[lev].p(), mg.cfg.pre, true, true, sc[lev]->p(), st);
        const Aggregation& P = *mg.P[lev];
        launch_b_restrict<T>(P.A, P.bv<T>(), rs[lev].p(), bs[lev + 1].p(), N, st);
        run(lev + 1, bs[lev + 1].p(), xs[lev + 1].p(), st);
        if (mg.cfg.w_cycle && lev + 1 < last) {
            ops[lev + 1]->apply(ts[lev + 1].p(), xs[lev + 1].p(), EPI_RESID, bs[lev + 1].p(), nullptr, FinB(), st);
            run(lev + 1, ts[lev + 1].p(), e2s[lev + 1].p(), st);
            launch_b_axpby_real<T>(xs[lev + 1].p(), e2s[lev + 1].p(), 1.0, 1.0, mg.dims[lev + 1] * N, st);
        }
        if (ximp != 0.0) launch_b_prolong<T>(P.A, P.bv<T>(), xs[lev + 1].p(), x, 2, N, st, b, ximp);
        else launch_b_prolong<T>(P.A, P.bv<T>(), xs[lev + 1].p(), x, 1, N, st);
        if (cheb) cheb_smooth_b<T>(A, x, b, rs[lev].p(), cds[lev].p(), Ars[lev].p(), mg.cfg.post, false, false, clo,
                                    chi, mg.cheb_smooth_fused, st);
        else if (om > 0) rich_b<T>(A, x, b, rs[lev].p(), mg.cfg.post, om, false, false, st);
        else mr_b<T>(A, x, b, rs[lev].p(), Ars[lev].p(), mg.cfg.post, false, false, sc[lev]->p(), st);
    }
};

namespace {
std::map<std::pair<MG*, int>, std::unique_ptr<CycleB<double>>>& cyc64_map() {
    static auto* m = new std::map<std::pair<MG*, int>, std::unique_ptr<CycleB<double>>>();
    return *m;
}
std::map<std::pair<MG*, int>, std::unique_ptr<CycleB<float>>>& cyc32_map() {
    static auto* m = new std::map<std::pair<MG*, int>, std::unique_ptr<CycleB<float>>>();
    return *m;
}
}  // namespace

template <>
CycleB<double>& batched_cycle<double>(MG& mg, int N) {
    auto& m = cyc64_map();
    auto& c = m[{&mg, N}];
    if (!c) c.reset(new CycleB<double>(mg, N));
    return *c;
}
template <>
CycleB<float>& batched_cycle<float>(MG& mg, int N) {
    auto& m = cyc32_map();
    auto& c = m[{&mg, N}];
    if (!c) c.reset(new CycleB<float>(mg, N));
    return *c;
}
void coarsest_solve_b32(MG& mg, const c32* b, c32* x, int N, hipStream_t st) {
    CycleB<float>& c = batched_cycle<float>(mg, N);
    c.run(c.nlev - 1, b, x, st);
}
void release_batched_cycles(MG* mg) {
    for (auto it = cyc64_map().begin(); it != cyc64_map().end();)
        it = (it->first.first == mg) ? cyc64_map().erase(it) : std::next(it);
    for (auto it = cyc32_map().begin(); it != cyc32_map().end();)
        it = (it->first.first == mg) ? cyc32_map().erase(it) : std::next(it);
}

// With cfg.use_graph the cycle is captured into a hipGraph on its first use, on fixed
// input/output buffers, and replayed: every kernel of the cycle reads its scalars on
// the device, none synchronises, allocates or clears memory with a memset, so it
// captures as is.  A hierarchy rebuilt, refreshed or retuned drops its batched cycles,
// and so their graphs.
template <typename T>
void mg_precondition_b(MG& mg, const cplx<T>* b, cplx<T>* x, int N, hipStream_t st) {
    CycleB<T>& c = batched_cycle<T>(mg, N);
    if (!mg.cfg.use_graph) {
        c.run(0, b, x, st);
        return;
    }
    const size_t nN = mg.dims[0] * N;
    if (!c.gexec) {
        c.gin = Field<T>(nN);
        c.gout = Field<T>(nN);
        SMG_HIP(hipStreamSynchronize(st));
        SMG_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        try {
            c.run(0, c.gin.p(), c.gout.p(), st);
        } catch (...) {
            hipGraph_t tmp = nullptr;
            (void)hipStreamEndCapture(st, &tmp);
            if (tmp) (void)hipGraphDestroy(tmp);
            throw;
        }
        SMG_HIP(hipStreamEndCapture(st, &c.graph));
        SMG_HIP(hipGraphInstantiate(&c.gexec, c.graph, nullptr, nullptr, 0));
    }
    launch_copy<T>(c.gin.p(), b, nN, st);
    SMG_HIP(hipGraphLaunch(c.gexec, st));
    launch_copy<T>(x, c.gout.p(), nN, st);
}
template void mg_precondition_b<double>(MG&, const c64*, c64*, int, hipStream_t);
template void mg_precondition_b<float>(MG&, const c32*, c32*, int, hipStream_t);

// ------------------------------------------------------------------ GCR (batched)
// fcg_solve (src/solvers.cpp:223-300) with the textbook projection, N RHS in lockstep
template <typename T>
int gcr_solve_b(LinOpB<T>& A, const cplx<T>* b, cplx<T>* x, MG* mg, int maxit, double tol, int trunc, double* res,
                hipStream_t st) {
    const size_t n = A.n(), nN = n * A.N;
    const int N = A.N;
    SMG_REQUIRE(trunc >= 1 && trunc <= 16, "batched GCR truncation window must be 1..16");
    Context& cx = Context::get();
    ScalarsB sc(64 * 2 + 64 + 64 * 32 + 17 * 64 + 17 * 128);
    double* rr = sc.p(128);                    // [N]
    double* dots = sc.p(192);                  // [N][32]
    double* apap = sc.p(192 + 64 * 32);        // [slot][N]
    double* alphas = sc.p(192 + 64 * 32 + 17 * 64);   // [slot][N][2]
    std::vector<double> bn(N), rn(N);
    launch_b_norm2<T>(b, n, finb(FIN_STORE, rr, 1), N, st);
    cx.read(rr, bn.data(), N);
    for (int j = 0; j < N; j++) {
        bn[j] = std::sqrt(bn[j]);
        rn[j] = bn[j];
    }
    Field<T> r(nN);
    bcopy(r.p(), b, nN, st);
    std::vector<Field<T>> Pd, APd;
    for (int i = 0; i <= trunc; i++) {
        Pd.emplace_back(nN);
        APd.emplace_back(nN);
    }
    std::deque<int> order;
    std::vector<int> freeslots;
    for (int i = trunc; i >= 0; i--) freeslots.push_back(i);
    // x is formed from the directions and their alphas, x = sum alpha_i p_i, at the end and
    // before a slot whose direction is not yet in x is reused, instead of at every step
    std::vector<int> pend;
    bool xset = false;
    auto flush = [&]() {
        for (size_t i0 = 0; i0 < pend.size(); i0 += 16) {   // at most 16 directions per pass
            const int m = (int)std::min<size_t>(16, pend.size() - i0);
            VecPtrs<T> Pv;
            const double* co[16];
            for (int i = 0; i < m; i++) {
                Pv.v[i] = Pd[pend[i0 + i]].p();
                co[i] = alphas + (size_t)pend[i0 + i] * 128;
            }
            launch_b_combine<T>(x, Pv, co, m, xset ? 1 : 0, n, N, st);
            xset = true;
        }
        pend.clear();
    };
    auto done = [&]() {
        for (int j = 0; j < N; j++)
            if (bn[j] > 1e-30 && rn[j] / bn[j] >= tol) return false;
        return true;
    };
    int iter = 0;
    while (iter < maxit && !done()) {
        const int slot = freeslots.back();
        freeslots.pop_back();
        cplx<T>* p = Pd[slot].p();
        cplx<T>* Ap = APd[slot].p();
        if (mg) batched_cycle<T>(*mg, N).run(0, r.p(), p, st);
        else bcopy(p, r.p(), nN, st);
        const int nd = (int)order.size();
        // alpha = <Ap, r> / |Ap|^2 from the sums of the pass that forms Ap last: the apply
        // (no earlier directions) or the orthogonalisation
        double* const alpha = alphas + (size_t)slot * 128;
        const FinB fa = finb(FIN_GCR_ALPHA, alpha, 2, apap + (size_t)slot * N, 1);
        if (nd == 0) A.apply(Ap, p, EPI_DOTS, nullptr, r.p(), fa, st);
        else A.apply(Ap, p, st);
        if (nd > 0) {
            VecPtrs<T> Pv, APv;
            const double* ap[16];
            for (int i = 0; i < nd; i++) {
                Pv.v[i] = Pd[order[i]].p();
                APv.v[i] = APd[order[i]].p();
                ap[i] = apap + (size_t)order[i] * N;
            }
            launch_b_multi_dot<T>(APv, nd, Ap, n, finb(FIN_STORE, dots, 32), N, st);
            launch_b_gcr_orth<T>(p, Ap, Pv, APv, nd, dots, ap, n, N, st, r.p(), &fa);
        }
        launch_b_update<T>(nullptr, r.p(), p, Ap, alpha, n, finb(FIN_STORE, rr, 1), N, st);
        pend.push_back(slot);
        cx.read(rr, rn.data(), N);
        for (int j = 0; j < N; j++) rn[j] = std::sqrt(rn[j]);
        iter++;
        if ((int)order.size() >= trunc) {
            if (std::find(pend.begin(), pend.end(), order.front()) != pend.end()) flush();
            freeslots.push_back(order.front());
            order.pop_front();
        }
        order.push_back(slot);
    }
    flush();
    if (!xset) launch_zero<T>(x, nN, st);   // no step taken
    for (int j = 0; j < N; j++) res[j] = bn[j] > 1e-30 ? rn[j] / bn[j] : 0.0;
    return iter;
}
template int gcr_solve_b<double>(LinOpB<double>&, const c64*, c64*, MG*, int, double, int, double*, hipStream_t);
template int gcr_solve_b<float>(LinOpB<float>&, const c32*, c32*, MG*, int, double, int, double*, hipStream_t);

BatchedSolveResult gcr_solve_batched(const Dirac& D, MG* mg, int prec, const c64* B, c64* X, int N, int maxit,
                                     double tol, int trunc, double inner_tol, hipStream_t st) {
    SMG_REQUIRE(N >= 1 && N <= 32 && (N & (N - 1)) == 0, "batched solves take a power-of-two count <= 32");
    BatchedSolveResult out;
    out.true_residual.assign(N, 0.0);
    Context& cx = Context::get();
    const size_t n = 2 * (size_t)D.V, nN = n * N;
    DdagDOpB<double> A64(D, N);
    Field<double> b64(nN), x64(nN), r64(nN);
    ScalarsB sc(64);
    launch_b_pack<double, double>(b64.p(), B, n, N, st);
    std::vector<double> bn(N), rn(N), res(N);
    launch_b_norm2<double>(b64.p(), n, finb(FIN_STORE, sc.p(), 1), N, st);
    cx.read(sc.p(), bn.data(), N);
    for (int j = 0; j < N; j++) bn[j] = std::sqrt(bn[j]);
    if (prec == 0) {
        out.iterations = gcr_solve_b<double>(A64, b64.p(), x64.p(), mg, maxit, tol, trunc, res.data(), st);
    } else {
        DdagDOpB<float> A32(D, N);
        Field<float> r32(nN), e32(nN);
        const c64* rsrc = b64.p();   // the residual of x = 0 is b; then r64
        rn = bn;
        int total = 0;
        while (total < maxit) {
            bool conv = true;
            double worst = 0.0;
            for (int j = 0; j < N; j++)
                if (bn[j] > 1e-30) {
                    worst = std::max(worst, rn[j] / bn[j]);
                    if (rn[j] / bn[j] >= tol) conv = false;
                }
            if (conv) break;
            launch_convert<float, double>(r32.p(), rsrc, nN, st);
            const double itol = std::max(inner_tol, 0.5 * tol / worst);
            const int it = gcr_solve_b<float>(A32, r32.p(), e32.p(), mg, maxit - total, itol, trunc, res.data(), st);
            if (rsrc == b64.p()) launch_convert<double, float>(x64.p(), e32.p(), nN, st);   // x = e
            else launch_b_accum_mixed(x64.p(), e32.p(), n, N, st);
            total += std::max(it, 1);
            rsrc = r64.p();
            A64.apply(r64.p(), x64.p(), EPI_RESID, b64.p(), nullptr, FinB(), st);
            launch_b_norm2<double>(r64.p(), n, finb(FIN_STORE, sc.p(), 1), N, st);
            cx.read(sc.p(), rn.data(), N);
            for (int j = 0; j < N; j++) rn[j] = std::sqrt(rn[j]);
            out.restarts++;
        }
        if (rsrc == b64.p()) launch_zero<double>(x64.p(), nN, st);   // no refinement step: x = 0
        out.iterations = total;
    }
    if (prec == 0) {   // FP64 true residuals (the refinement loop above ends with them in rn)
        A64.apply(r64.p(), x64.p(), EPI_RESID, b64.p(), nullptr, FinB(), st);
        launch_b_norm2<double>(r64.p(), n, finb(FIN_STORE, sc.p(), 1), N, st);
        cx.read(sc.p(), rn.data(), N);
        for (int j = 0; j < N; j++) rn[j] = std::sqrt(rn[j]);
    }
    for (int j = 0; j < N; j++) out.true_residual[j] = bn[j] > 1e-30 ? rn[j] / bn[j] : 0.0;
    launch_b_unpack<double, double>(X, x64.p(), n, N, st);
    return out;
}


// ------------------------------------------------------------------ autotuner searches
namespace {
// average duration of one call of f, from HIP events around `reps` calls after two
// untimed calls (every call on the library stream st)
double time_us(const std::function<void()>& f, hipStream_t st, int reps = 6) {
    hipEvent_t e0, e1;
    SMG_HIP(hipEventCreate(&e0));
    SMG_HIP(hipEventCreate(&e1));
    f();
    f();
    SMG_HIP(hipEventRecord(e0, st));
    for (int i = 0; i < reps; i++) f();
    SMG_HIP(hipEventRecord(e1, st));
    SMG_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    SMG_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return 1e3 * ms / reps;
}
// times every candidate choice of key with launch(), keeps the fastest in the table
int pick(const TuneKey& key, const std::vector<int>& cands, const std::function<void()>& launch, hipStream_t st) {
    double best = 1e300;
    int arg = cands.front();
    for (int c : cands) {
        tune_store(key, c);
        const double t = time_us(launch, st);
        if (t < best) {
            best = t;
            arg = c;
        }
    }
    tune_store(key, arg);
    return arg;
}
}  // namespace

// The launch shapes of the batched cycle for N right-hand sides: for every coarse
// level's FP16-link matrix-core apply (each epilogue, and the BF16-vector form of the
// coarsest CG) the link-read variant and the workgroup count of its reducing launch,
// and the workgroup count of the fine stencil's reducing launch (FP32 and FP64).
int tune_mg(MG& mg, int N, hipStream_t st) {
    SMG_REQUIRE(N == 16 || N == 32 || N == 64, "tuning: N must be 16, 32 or 64");
    const int before = (int)tune_size();
    Field<double> out((size_t)5 * 64);
    const FinB fin = finb(FIN_STORE, reinterpret_cast<double*>(out.p()), 3);
    for (size_t l = 0; l < mg.C.size(); l++) {
        const CoarseOperator& C = *mg.C[l];
        if (C.k != 8 && C.k != 16 && C.k != 32) continue;
        const size_t nN = C.n() * N;
        Field<float> src(nN), dst(nN), aux(nN);
        launch_zero<float>(src.p(), nN, st);
        launch_zero<float>(aux.p(), nN, st);
        CoarseLinOpB<float, ch16> op(C, N);
        const int sites = C.nx * C.ny, waves_per_wg = kBlock / 64;
        std::vector<int> herm = {0};
        if (C.herm16 && C.k % 16 == 0) herm.push_back(1);
        std::vector<int> dots;
        for (int h : herm)
            for (int g = 0; g < kTuneNGrids; g++)
                if (g == 0 || kTuneGrids[g] <= (sites + waves_per_wg - 1) / waves_per_wg) dots.push_back(h + 2 * g);
        for (int epi : {(int)EPI_NONE, (int)EPI_RESID, (int)EPI_DOTS}) {
            const TuneKey key{TUNE_COARSE_BF16, C.nx, C.ny, C.k, N, epi};
            pick(key, epi == EPI_DOTS ? dots : herm, [&] {
                op.apply(dst.p(), src.p(), (Epi)epi, aux.p(), aux.p(), fin, st);
            }, st);
        }
        const bool last = l + 1 == mg.C.size();
        if (last && mg.cfg.coarse_half == 2 && mg.cfg.coarse_solver == 2) {   // BF16-vector coarsest Chebyshev
            DevMem hs(nN * sizeof(cb16)), hd(nN * sizeof(cb16)), hr(nN * sizeof(cb16)), hx(nN * sizeof(cb16));
            for (DevMem* m : {&hs, &hd, &hr, &hx}) launch_zero<float>(m->as<c32>(), nN / 2, st);   // BF16 pairs
            FinB f;
            pick(TuneKey{TUNE_COARSE_BF16, C.nx, C.ny, C.k, N, 4 + (int)EPI_CHEB}, herm, [&] {
                launch_b_coarse_apply_h(C.view16(), hd.as<cb16>(), hs.as<cb16>(), EPI_CHEB, hx.as<cb16>(), f, N, st,
                                        hr.as<cb16>());
            }, st);
        }
        if (last && mg.cfg.coarse_half == 2 && mg.cfg.coarse_solver == 0) {   // BF16-vector coarsest CG
            DevMem hs(nN * sizeof(cb16)), hd(nN * sizeof(cb16));
            SMG_HIP(hipMemsetAsync(hs.p, 0, nN * sizeof(cb16), st));
            const FinB fin5 = finb(FIN_STORE, reinterpret_cast<double*>(out.p()), 5);
            for (int epi : {(int)EPI_DOTS, (int)EPI_CG})
                pick(TuneKey{TUNE_COARSE_BF16, C.nx, C.ny, C.k, N, 4 + epi}, dots, [&] {
                    launch_b_coarse_apply_h(C.view16(), hd.as<cb16>(), hs.as<cb16>(), (Epi)epi, hs.as<cb16>(),
                                            epi == EPI_CG ? fin5 : fin, N, st, hs.as<cb16>());
                }, st);
            // the CG's vector update (zero coefficients leave the vectors as they are)
            Field<double> a0((size_t)16 * 64);
            launch_zero<double>(a0.p(), (size_t)16 * 64, st);
            std::vector<int> vgrids;
            for (int g = 0; g < kTuneNGrids; g++) vgrids.push_back(g);
            DevMem hx(nN * sizeof(cb16)), hq(nN * sizeof(cb16));
            SMG_HIP(hipMemsetAsync(hx.p, 0, nN * sizeof(cb16), st));
            SMG_HIP(hipMemsetAsync(hq.p, 0, nN * sizeof(cb16), st));
            SMG_HIP(hipMemsetAsync(hd.p, 0, nN * sizeof(cb16), st));
            pick(TuneKey{TUNE_CG_UPDATE, (int)C.n(), N, 0, 0, 0}, vgrids, [&] {
                launch_h_cg_step(hx.as<cb16>(), hs.as<cb16>(), hd.as<cb16>(), hq.as<cb16>(),
                                 reinterpret_cast<const double*>(a0.p()), reinterpret_cast<const double*>(a0.p()),
                                 C.n(), N, nullptr, st);
            }, st);
        }
    }
    // the transfers of every level, FP32 (the mixed cycle) and FP64: the per-output kernel
    // or the workgroup per aggregate with LDS-staged blocks
    for (size_t l = 0; l < mg.P.size(); l++) {
        const AggView& A = mg.P[l]->A;
        auto transfers = [&](auto zero) {
            using T = decltype(zero);
            const size_t nf = mg.P[l]->fine_n() * N, nc = mg.P[l]->coarse_n() * N;
            Field<T> f(nf), c(nc);
            launch_zero<T>(f.p(), nf, st);
            launch_zero<T>(c.p(), nc, st);
            const int prec = sizeof(T) == 8 ? 1 : 0;
            const bool mfma = sizeof(T) == 4 && b_transfer_mfma_fits(A, N);
            for (int pro = 0; pro < 2; pro++) {
                std::vector<int> cands{0};
                if (b_transfer_lds_bytes(A, N, sizeof(cplx<T>), pro == 1)) cands.push_back(1);
                if (mfma) cands.push_back(2);
                if (cands.size() < 2) continue;
                pick(TuneKey{TUNE_TRANSFER, A.nx, A.ny, 128 * A.nd + A.k, N, 2 * prec + pro}, cands, [&] {
                    if (pro) launch_b_prolong<T>(A, mg.P[l]->bv<T>(), c.p(), f.p(), 1, N, st);
                    else launch_b_restrict<T>(A, mg.P[l]->bv<T>(), f.p(), c.p(), N, st);
                }, st);
            }
        };
        transfers(0.0f);
        transfers(0.0);
    }
    // the fine stencil with fused dots (the second pass of D^dag D in the smoother and CG)
    const Dirac& D = *mg.D;
    const size_t nf = 2 * (size_t)D.V * N;
    std::vector<int> grids;
    for (int g = 0; g < kTuneNGrids; g++) grids.push_back(g);
    // the stencil's store hint per epilogue (the first pass of D^dag D, the residual and
    // the pass with dots), then the grid of the pass with dots
    auto stencil = [&](auto zero) {
        using T = decltype(zero);
        Field<T> src(nf), dst(nf);
        launch_zero<T>(src.p(), nf, st);
        const DiracView<T> v = D.view<T>();
        for (int epi : {(int)EPI_NONE, (int)EPI_RESID, (int)EPI_DOTS})
            pick(TuneKey{TUNE_DSLASH_STORE, D.L, N, (int)sizeof(T), epi, 0}, {0, 1}, [&] {
                launch_b_dslash<T>(v, dst.p(), src.p(), zero, 1, 1, (Epi)epi, src.p(), src.p(), fin, N, st);
            }, st);
        pick(TuneKey{TUNE_DSLASH_DOTS, D.L, N, (int)sizeof(T), 0, 0}, grids, [&] {
            launch_b_dslash<T>(v, dst.p(), src.p(), zero, 1, 1, EPI_DOTS, nullptr, src.p(), fin, N, st);
        }, st);
    };
    stencil(0.0f);
    stencil(0.0);
    // D^dag D per epilogue: the two stencil launches (their choices above) or the fused
    // launch with ordinary or non-temporal stores (and, with dots, its workgroup count)
    auto ddagd = [&](auto zero) {
        using T = decltype(zero);
        if (!b_ddagd_fits(D.L, N, sizeof(T))) return;
        DdagDOpB<T> A(D, N);
        Field<T> src(nf), dst(nf);
        launch_zero<T>(src.p(), nf, st);
        for (int epi : {(int)EPI_NONE, (int)EPI_RESID, (int)EPI_DOTS}) {
            std::vector<int> cands = {0, 1, 2};
            if (epi == EPI_DOTS)
                for (int g = 1; g < kTuneNGrids; g++) {
                    cands.push_back(1 + 2 * g);
                    cands.push_back(2 + 2 * g);
                }
            pick(TuneKey{TUNE_DDAGD, D.L, N, (int)sizeof(T), epi, 0}, cands, [&] {
                A.apply(dst.p(), src.p(), (Epi)epi, src.p(), src.p(), fin, st);
            }, st);
        }
    };
    ddagd(0.0f);
    ddagd(0.0);
    mg.reset_graphs();   // captured cycles hold the old launch shapes
    release_batched_cycles(&mg);
    SMG_HIP(hipStreamSynchronize(st));
    return (int)tune_size() - before;
}

}  // namespace smg
