// capi.cpp -- the extern "C" boundary (include/smg.h) over the C++ runtime.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>

#include "../../include/smg.h"
#include "eigen.h"
#include "runtime.h"
#include "smg_lattices.h"
#include "runtime_b.h"
#include "runtime_lat.h"
#include "tune.h"

using namespace smg;

struct smg_dirac_s {
    std::unique_ptr<Dirac> D;
    std::unique_ptr<DdagDOp<double>> op64;
    std::unique_ptr<DdagDOp<float>> op32;
    std::unique_ptr<SchurOp<double>> schur64;
    std::unique_ptr<SchurOp<float>> schur32;
};
struct smg_prolongator_s {
    std::unique_ptr<Aggregation> P;
};
struct smg_coarse_s {
    std::unique_ptr<CoarseOperator> C;
    Eigenpairs defl;   // SparseCoarseOp::defl_vecs / defl_vals (empty: no deflation)
};
struct smg_lattices_s {
    std::vector<std::unique_ptr<Dirac>> D;   // per lattice: links, clover, parameters
    std::vector<LatDesc> h;                  // host copy of the table
    DevMem d;                                // the table on the device
    size_t sites = 0;
    int nwg = 0;
};
struct smg_mg_s {
    std::unique_ptr<MG> mg;
    smg_dirac* D;
};
struct smg_eo_mg_s {
    std::unique_ptr<EOMG> mg;
};
struct smg_lattices_mg_s {
    std::unique_ptr<LatticeMG> mg;
    smg_lattices* H;
};

namespace {
thread_local std::string g_err;

template <class F>
int guard(F&& f) {
    try {
        f();
        return 0;
    } catch (const Error& e) {
        g_err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -3;
    } catch (...) {
        g_err = "unknown error";
        return -3;
    }
}

hipStream_t S() { return Context::get().stream; }

Field<double> up(const double* h, size_t n) {
    SMG_REQUIRE(h != nullptr, "null input array");
    Field<double> f(n);
    SMG_HIP(hipMemcpyAsync(f.p(), h, n * sizeof(c64), hipMemcpyHostToDevice, S()));
    return f;
}
void down(double* h, const void* d, size_t bytes) {
    SMG_REQUIRE(h != nullptr, "null output array");
    SMG_HIP(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, S()));
    SMG_HIP(hipStreamSynchronize(S()));
}
void downc(double* h, const c64* d, size_t n) { down(h, d, n * sizeof(c64)); }

// Orders the library stream after the caller's stream on entry, and the caller's
// stream after the library's work on exit.
// Orders the library's stream after the caller's stream on entry and the
// caller's stream after the library's on exit.  A caller that passes the
// library's own stream needs neither.  The two events are reused (a wait takes
// the event's state at the time of the call), one pair per host thread.
struct StreamJoin {
    hipStream_t user;
    bool same;
    static hipEvent_t event(int i) {
        thread_local hipEvent_t ev[2] = {nullptr, nullptr};
        if (!ev[i]) SMG_HIP(hipEventCreateWithFlags(&ev[i], hipEventDisableTiming));
        return ev[i];
    }
    explicit StreamJoin(void* s) : user(reinterpret_cast<hipStream_t>(s)), same(user == S()) {
        if (same) return;
        hipEvent_t ev = event(0);
        SMG_HIP(hipEventRecord(ev, user));
        SMG_HIP(hipStreamWaitEvent(S(), ev, 0));
    }
    ~StreamJoin() {
        if (same) return;
        hipEvent_t ev = event(1);
        if (hipEventRecord(ev, S()) == hipSuccess) (void)hipStreamWaitEvent(user, ev, 0);
    }
};

void check_prec(int prec) { SMG_REQUIRE(prec == SMG_F64 || prec == SMG_F32, "prec must be SMG_F64 or SMG_F32"); }
}  // namespace

// Every smg_* definition below has C linkage from its declaration in smg.h.

int smg_init(int device) {
    return guard([&] {
        SMG_HIP(hipSetDevice(device));
        Context::get();
        noalloc_scratch();   // this device's address of the allocation-free entry's scratch
    });
}
const char* smg_last_error(void) { return g_err.c_str(); }
int smg_version(void) { return 1; }
int smg_synchronize(void) { return guard([] { Context::get().sync(); }); }
int smg_get_stream(void** stream) {
    return guard([&] {
        SMG_REQUIRE(stream, "null argument");
        *stream = reinterpret_cast<void*>(Context::get().stream);
    });
}

int smg_eo_maps(int L, int* even, int* odd, int* eoi) {
    return guard([&] {
        SMG_REQUIRE(L >= 2 && even && odd && eoi, "bad arguments");
        int ne = 0, no = 0;
        for (int s = 0; s < L * L; s++) {
            if (((s % L) + (s / L)) % 2 == 0) { eoi[s] = ne; even[ne++] = s; }
            else { eoi[s] = no; odd[no++] = s; }
        }
    });
}

// a one-component FIN_STORE reduction target and its host read-back
struct Scalar1 {
    DevMem m{8 * sizeof(double)};
    Fin fin() { return Context::get().fin(FIN_STORE, m.as<double>()); }
    double read() {
        double h;
        Context::get().read(m.as<double>(), &h, 1);
        return h;
    }
};

int smg_avg_plaq(int L, const double* U, double* out) {
    return guard([&] {
        Field<double> u = up(U, 2 * (size_t)L * L);
        Scalar1 r;
        launch_plaq_sum(L, u.p(), r.fin(), S());
        *out = r.read() / ((double)L * L);
    });
}
int smg_gauge_action(int L, const double* U, double beta, double* out) {
    return guard([&] {
        Field<double> u = up(U, 2 * (size_t)L * L);
        Scalar1 r;
        launch_plaq_sum(L, u.p(), r.fin(), S());
        *out = -beta * r.read();
    });
}
int smg_gauge_force(int L, const double* U, double beta, double* F) {
    return guard([&] {
        Field<double> u = up(U, 2 * (size_t)L * L);
        DevMem f(2 * (size_t)L * L * sizeof(double));
        launch_gauge_force(L, u.p(), beta, f.as<double>(), S());
        down(F, f.p, 2 * (size_t)L * L * sizeof(double));
    });
}

// ------------------------------------------------------------------ Dirac
int smg_dirac_create(smg_dirac** out, int L, const double* U, double mass, double r, double csw, double mu_t) {
    return guard([&] {
        SMG_REQUIRE(out && U, "null argument");
        SMG_REQUIRE(L >= 2, "L must be >= 2");
        auto* h = new smg_dirac_s();
        try {
            h->D.reset(new Dirac(L, reinterpret_cast<const c64*>(U), mass, r, csw, mu_t));
            h->op64.reset(new DdagDOp<double>(*h->D));
            h->op32.reset(new DdagDOp<float>(*h->D));
            if (L % 2 == 0) {
                h->schur64.reset(new SchurOp<double>(*h->D));
                h->schur32.reset(new SchurOp<float>(*h->D));
            }
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}
int smg_dirac_destroy(smg_dirac* D) {
    return guard([&] { delete D; });
}
int smg_dirac_set_gauge(smg_dirac* D, const double* U) {
    return guard([&] {
        SMG_REQUIRE(D && U, "null argument");
        D->D->set_gauge(reinterpret_cast<const c64*>(U), false);
    });
}
int smg_dirac_clover(smg_dirac* D, double* F01) {
    return guard([&] {
        SMG_REQUIRE(D, "null handle");
        down(F01, D->D->F01.p, D->D->V * sizeof(double));
    });
}

template <typename T>
static void dirac_kind(const Dirac& d, LinOp<T>& op, int kind, const cplx<T>* src, cplx<T>* dst, int nrhs,
                       hipStream_t st) {
    const DiracView<T> v = d.view<T>();
    const size_t n = 2 * (size_t)d.V;
    const T mu = (T)d.mu_t;
    switch (kind) {
    case SMG_D:
        launch_dslash<T>(v, dst, src, mu, 0, 0, EPI_NONE, nullptr, nullptr, Fin(), nrhs, n, st);
        break;
    case SMG_DDAG:
        launch_dslash<T>(v, dst, src, -mu, 1, 1, EPI_NONE, nullptr, nullptr, Fin(), nrhs, n, st);
        break;
    case SMG_G5D:
        launch_dslash<T>(v, dst, src, mu, 0, 1, EPI_NONE, nullptr, nullptr, Fin(), nrhs, n, st);
        break;
    case SMG_DDAGD:
        for (int i = 0; i < nrhs; i++) op.apply(dst + i * n, src + i * n, st);
        break;
    default:
        SMG_REQUIRE(false, "unknown Dirac kind");
    }
}

int smg_dirac_apply(smg_dirac* D, int kind, const double* src, double* dst) {
    return guard([&] {
        SMG_REQUIRE(D, "null handle");
        const size_t n = 2 * (size_t)D->D->V;
        Field<double> s = up(src, n), d(n);
        dirac_kind<double>(*D->D, *D->op64, kind, s.p(), d.p(), 1, S());
        downc(dst, d.p(), n);
    });
}
int smg_dirac_apply_dev(smg_dirac* D, int kind, int prec, const void* src, void* dst, int nrhs, void* stream) {
    return guard([&] {
        SMG_REQUIRE(D && src && dst && nrhs >= 1, "bad arguments");
        check_prec(prec);
        StreamJoin j(stream);
        if (prec == SMG_F64)
            dirac_kind<double>(*D->D, *D->op64, kind, (const c64*)src, (c64*)dst, nrhs, S());
        else
            dirac_kind<float>(*D->D, *D->op32, kind, (const c32*)src, (c32*)dst, nrhs, S());
    });
}
int smg_dirac_delta(smg_dirac* D, const double* src, double* dst, const double* pi, double dt) {
    return guard([&] {
        SMG_REQUIRE(D && pi, "null argument");
        const size_t n = 2 * (size_t)D->D->V;
        Field<double> s = up(src, n), d(n);
        DevMem p(n * sizeof(double));
        SMG_HIP(hipMemcpyAsync(p.p, pi, n * sizeof(double), hipMemcpyHostToDevice, S()));
        launch_delta_dirac(D->D->v64(), d.p(), s.p(), p.as<double>(), dt, S());
        downc(dst, d.p(), n);
    });
}

template <typename T>
static void eo_kind(smg_dirac* D, int kind, const cplx<T>* s, cplx<T>* d, hipStream_t st) {
    const Dirac& dd = *D->D;
    SMG_REQUIRE(dd.L % 2 == 0, "even-odd operators need an even L");
    const DiracView<T> v = dd.view<T>();
    const T mu = (T)dd.mu_t;
    SchurOp<T>* sop;
    if constexpr (std::is_same<T, double>::value) sop = D->schur64.get();
    else sop = D->schur32.get();
    switch (kind) {
    case SMG_EO_HOP_EO:
        launch_eo_hop<T>(v, d, s, 1, 0, mu, 0, 0, nullptr, EPI_NONE, nullptr, nullptr, Fin(), st);
        break;
    case SMG_EO_HOP_OE:
        launch_eo_hop<T>(v, d, s, 0, 0, mu, 0, 0, nullptr, EPI_NONE, nullptr, nullptr, Fin(), st);
        break;
    case SMG_EO_EE_INV: launch_eo_diag<T>(v, d, s, 0, 1, mu, st); break;
    case SMG_EO_OO_INV: launch_eo_diag<T>(v, d, s, 1, 1, mu, st); break;
    case SMG_EO_OO: launch_eo_diag<T>(v, d, s, 1, 0, mu, st); break;
    case SMG_EO_SCHUR: sop->apply_m(d, s, 0, st); break;
    case SMG_EO_SCHUR_DAG: sop->apply_m(d, s, 1, st); break;
    case SMG_EO_SCHUR_DAG_SCHUR: sop->apply(d, s, st); break;
    default: SMG_REQUIRE(false, "unknown even-odd kind");
    }
}
int smg_eo_apply(smg_dirac* D, int kind, const double* src, double* dst) {
    return guard([&] {
        SMG_REQUIRE(D, "null handle");
        const size_t n = (size_t)D->D->V;
        Field<double> s = up(src, n), d(n);
        eo_kind<double>(D, kind, s.p(), d.p(), S());
        downc(dst, d.p(), n);
    });
}
int smg_eo_apply_dev(smg_dirac* D, int kind, int prec, const void* src, void* dst, void* stream) {
    return guard([&] {
        SMG_REQUIRE(D && src && dst, "bad arguments");
        check_prec(prec);
        StreamJoin j(stream);
        if (prec == SMG_F64) eo_kind<double>(D, kind, (const c64*)src, (c64*)dst, S());
        else eo_kind<float>(D, kind, (const c32*)src, (c32*)dst, S());
    });
}
int smg_eo_reconstruct(smg_dirac* D, const double* x_o, const double* phi, double* out) {
    return guard([&] {
        SMG_REQUIRE(D, "null handle");
        const Dirac& dd = *D->D;
        SMG_REQUIRE(dd.L % 2 == 0, "even-odd operators need an even L");
        const size_t nh = (size_t)dd.V, n = 2 * nh;
        Field<double> xo = up(x_o, nh), h(nh), te(nh), xe(nh), full(n);
        const DiracView<double> v = dd.v64();
        launch_eo_hop<double>(v, h.p(), xo.p(), 0, 0, dd.mu_t, 0, 0, nullptr, EPI_NONE, nullptr, nullptr, Fin(),
                              S());
        if (phi) {
            Field<double> ph = up(phi, n);
            launch_eo_gather<double>(dd.L, te.p(), ph.p(), 0, S());
            launch_axpby<double>(te.p(), h.p(), c64{-1, 0}, c64{1, 0}, nh, S());   // phi_e - H x_o
        } else {
            launch_axpby<double>(te.p(), h.p(), c64{-1, 0}, c64{0, 0}, nh, S());   // -H x_o
        }
        launch_eo_diag<double>(v, xe.p(), te.p(), 0, 1, dd.mu_t, S());
        launch_eo_scatter<double>(dd.L, full.p(), xe.p(), 0, S());
        launch_eo_scatter<double>(dd.L, full.p(), xo.p(), 1, S());
        downc(out, full.p(), n);
    });
}
int smg_eo_logdet(smg_dirac* D, double* out) {
    return guard([&] {
        SMG_REQUIRE(D && out, "null argument");
        SMG_REQUIRE(D->D->L % 2 == 0, "even-odd operators need an even L");
        Scalar1 r;
        launch_logdet(D->D->v64(), D->D->mu_t, r.fin(), S());
        *out = r.read();
    });
}

// ------------------------------------------------------------------ BLAS-1
int smg_dot_dev(const void* a, const void* b, size_t n, int prec, double* out2, void* stream) {
    return guard([&] {
        check_prec(prec);
        StreamJoin j(stream);
        c64 d = prec == SMG_F64 ? dev_dot<double>((const c64*)a, (const c64*)b, n, S())
                                : dev_dot<float>((const c32*)a, (const c32*)b, n, S());
        out2[0] = d.re;
        out2[1] = d.im;
    });
}
int smg_dot_dev_noalloc(const void* a, const void* b, size_t n, int prec, double* out2, void* stream) {
    return guard([&] {
        SMG_REQUIRE(a && b && out2, "null argument");
        check_prec(prec);
        noalloc_dot(a, b, n, prec == SMG_F64 ? 1 : 0, out2, reinterpret_cast<hipStream_t>(stream));
    });
}
int smg_norm_dev(const void* a, size_t n, int prec, double* out, void* stream) {
    return guard([&] {
        check_prec(prec);
        StreamJoin j(stream);
        c64 d = prec == SMG_F64 ? dev_dot<double>((const c64*)a, (const c64*)a, n, S())
                                : dev_dot<float>((const c32*)a, (const c32*)a, n, S());
        *out = std::sqrt(d.re);
    });
}

// ------------------------------------------------------------------ prolongator
int smg_prolongator_create(smg_prolongator** out, int nx, int ny, int nd, int bx, int by, int k, int chiral) {
    return guard([&] {
        SMG_REQUIRE(out && nx > 0 && ny > 0 && nd > 0, "bad arguments");
        auto* h = new smg_prolongator_s();
        try { h->P.reset(new Aggregation(nx, ny, nd, bx, by, k, chiral)); } catch (...) { delete h; throw; }
        *out = h;
    });
}
int smg_prolongator_destroy(smg_prolongator* P) { return guard([&] { delete P; }); }
int smg_prolongator_build(smg_prolongator* P, const double* null_vecs) {
    return guard([&] {
        SMG_REQUIRE(P, "null handle");
        Field<double> nv = up(null_vecs, P->P->fine_n() * P->P->nvec());
        P->P->build(nv.p(), S());
        Context::get().sync();
    });
}
int smg_prolongator_get(smg_prolongator* P, double* bv) {
    return guard([&] { SMG_REQUIRE(P, "null handle"); downc(bv, P->P->bv64.p(), P->P->bv64.n); });
}
int smg_prolongator_set(smg_prolongator* P, const double* bv) {
    return guard([&] {
        SMG_REQUIRE(P, "null handle");
        Field<double> b = up(bv, P->P->bv64.n);
        P->P->set_block_vecs(b.p(), S());
        Context::get().sync();
    });
}
int smg_prolong(smg_prolongator* P, const double* coarse, double* fine) {
    return guard([&] {
        SMG_REQUIRE(P, "null handle");
        Field<double> c = up(coarse, P->P->coarse_n()), f(P->P->fine_n());
        launch_prolong<double>(P->P->A, P->P->bv64.p(), c.p(), f.p(), 0, S());
        downc(fine, f.p(), f.n);
    });
}
int smg_restrict(smg_prolongator* P, const double* fine, double* coarse) {
    return guard([&] {
        SMG_REQUIRE(P, "null handle");
        Field<double> f = up(fine, P->P->fine_n()), c(P->P->coarse_n());
        launch_restrict<double>(P->P->A, P->P->bv64.p(), f.p(), c.p(), S());
        downc(coarse, c.p(), c.n);
    });
}

// ------------------------------------------------------------------ coarse operator
int smg_coarse_create_galerkin(smg_coarse** out, smg_dirac* D, smg_prolongator* P) {
    return guard([&] {
        SMG_REQUIRE(out && D && P, "null argument");
        SMG_REQUIRE(P->P->A.nx == D->D->L && P->P->A.ny == D->D->L && P->P->A.nd == 2,
                    "prolongator does not act on this Dirac operator's lattice");
        auto* h = new smg_coarse_s();
        try {
            h->C.reset(new CoarseOperator(P->P->A.nbx, P->P->A.nby, P->P->A.k));
            h->C->build(*D->op64, *P->P, S());
        } catch (...) { delete h; throw; }
        *out = h;
    });
}
int smg_coarse_create_from_coarse(smg_coarse** out, smg_coarse* A, smg_prolongator* P) {
    return guard([&] {
        SMG_REQUIRE(out && A && P, "null argument");
        SMG_REQUIRE(P->P->A.nx == A->C->nx && P->P->A.ny == A->C->ny && P->P->A.nd == A->C->k,
                    "prolongator does not act on this coarse operator's lattice");
        auto* h = new smg_coarse_s();
        try {
            CoarseLinOp<double, c64> op(*A->C);
            h->C.reset(new CoarseOperator(P->P->A.nbx, P->P->A.nby, P->P->A.k));
            h->C->build(op, *P->P, S());
        } catch (...) { delete h; throw; }
        *out = h;
    });
}
int smg_coarse_create_links(smg_coarse** out, int nx, int ny, int k, const double* links) {
    return guard([&] {
        SMG_REQUIRE(out && nx > 0 && ny > 0 && k > 0, "bad arguments");
        auto* h = new smg_coarse_s();
        try {
            h->C.reset(new CoarseOperator(nx, ny, k));
            Field<double> l = up(links, h->C->nlinks());
            h->C->set_links(l.p(), S());
            Context::get().sync();
        } catch (...) { delete h; throw; }
        *out = h;
    });
}
int smg_coarse_destroy(smg_coarse* C) { return guard([&] { delete C; }); }
int smg_coarse_get_links(smg_coarse* C, double* links) {
    return guard([&] { SMG_REQUIRE(C, "null handle"); downc(links, C->C->l64.p(), C->C->nlinks()); });
}
int smg_coarse_apply(smg_coarse* C, int prec, const double* src, double* dst) {
    return guard([&] {
        SMG_REQUIRE(C, "null handle");
        const size_t n = C->C->n();
        Field<double> s = up(src, n), d(n);
        if (prec == SMG_F64) {
            CoarseLinOp<double, c64> op(*C->C);
            op.apply(d.p(), s.p(), S());
        } else {
            Field<float> s32(n), d32(n);
            launch_convert<float, double>(s32.p(), s.p(), n, S());
            if (prec == SMG_F32) {
                CoarseLinOp<float, c32> op(*C->C);
                op.apply(d32.p(), s32.p(), S());
            } else {
                SMG_REQUIRE(prec == SMG_F16, "bad precision");
                CoarseLinOp<float, ch16> op(*C->C);
                op.apply(d32.p(), s32.p(), S());
            }
            launch_convert<double, float>(d.p(), d32.p(), n, S());
        }
        downc(dst, d.p(), n);
    });
}
int smg_coarse_solve(smg_coarse* C, const double* b, double* x, int maxit, double tol, int* iters) {
    return guard([&] {
        SMG_REQUIRE(C, "null handle");
        const size_t n = C->C->n();
        if (!C->defl.vals.empty()) {   // deflated initial guess, then the same CG
            Field<double> bb = up(b, n), xx(n);
            CoarseLinOp<double, c64> op(*C->C);
            const int it = deflated_coarse_cg(op, C->defl, bb.p(), xx.p(), maxit, tol, S());
            if (iters) *iters = it;
            downc(x, xx.p(), n);
            return;
        }
        Field<double> bb = up(b, n), xx(n), r(n), p(n), Ap(n);
        DevMem sc(64 * sizeof(double));
        SMG_HIP(hipMemsetAsync(sc.p, 0, 64 * sizeof(double), S()));
        CoarseLinOp<double, c64> op(*C->C);
        int it = coarse_cg<double>(op, bb.p(), xx.p(), r.p(), p.p(), Ap.p(), maxit, tol, sc.as<double>(), true, S());
        if (iters) *iters = it;
        downc(x, xx.p(), n);
    });
}

// Dense A_c (CoarseOp, src/coarse_op.h:18-59): the Galerkin stencil expanded to
// the reference's column layout mat[j][i] (M[j*n + i] = A_c[i][j])
int smg_coarse_dense(smg_coarse* C, double* M) {
    return guard([&] {
        SMG_REQUIRE(C && M, "bad arguments");
        const CoarseOperator& c = *C->C;
        const size_t n = c.n();
        SMG_REQUIRE(n <= 16384, "dense coarse operator: at most 16384 unknowns");
        Field<double> d(n * n);
        launch_zero<double>(d.p(), n * n, S());
        launch_coarse_to_dense(CoarseView<c64>{c.l64.p(), c.nx, c.ny, c.k}, d.p(), n, S());
        std::vector<c64> h(n * n);
        down(reinterpret_cast<double*>(h.data()), d.p(), n * n * sizeof(c64));
        c64* out = reinterpret_cast<c64*>(M);
        for (size_t i = 0; i < n; i++)
            for (size_t j = 0; j < n; j++) out[j * n + i] = h[i * n + j];
    });
}
// CoarseOp::solve (src/coarse_op.h:101-137): x = A_c^-1 b through the dense inverse
// (Gauss-Jordan without pivoting, valid for the Hermitian positive-definite Galerkin
// projection of D^dag D)
int smg_coarse_dense_solve(smg_coarse* C, const double* b, double* x) {
    return guard([&] {
        SMG_REQUIRE(C && b && x, "bad arguments");
        const CoarseOperator& c = *C->C;
        const size_t n = c.n();
        SMG_REQUIRE(n <= 8192, "dense coarse solve: at most 8192 unknowns");
        Field<double> inv(n * n), prow(n), pcol(n);
        launch_zero<double>(inv.p(), n * n, S());
        launch_coarse_to_dense(CoarseView<c64>{c.l64.p(), c.nx, c.ny, c.k}, inv.p(), n, S());
        launch_dense_invert(inv.p(), n, prow.p(), pcol.p(), S());
        Field<double> bb = up(b, n), xx(n);
        launch_dense_matvec<double>(inv.p(), bb.p(), xx.p(), n, S());
        downc(x, xx.p(), n);
    });
}

// ------------------------------------------------------------------ smoothers
template <typename Op>
static void smooth_op(Op& op, int kind, double* x, const double* b, int niter, double omega) {
    const size_t n = op.n();
    Field<double> xx = up(x, n), bb = up(b, n), r(n), Ar(n);
    DevMem sc(64 * sizeof(double));
    if (kind == SMG_SMOOTH_MR)
        mr_smooth<double>(op, xx.p(), bb.p(), r.p(), Ar.p(), niter, false, sc.as<double>(), S());
    else {
        SMG_REQUIRE(kind == SMG_SMOOTH_RICHARDSON, "unknown smoother");
        richardson_smooth<double>(op, xx.p(), bb.p(), r.p(), niter, omega, false, S());
    }
    downc(x, xx.p(), n);
}
int smg_smooth_dirac(smg_dirac* D, int kind, double* x, const double* b, int niter, double omega) {
    return guard([&] { SMG_REQUIRE(D, "null handle"); smooth_op(*D->op64, kind, x, b, niter, omega); });
}
int smg_smooth_coarse(smg_coarse* C, int kind, double* x, const double* b, int niter, double omega) {
    return guard([&] {
        SMG_REQUIRE(C, "null handle");
        CoarseLinOp<double, c64> op(*C->C);
        smooth_op(op, kind, x, b, niter, omega);
    });
}
int smg_lambda_max_dirac(smg_dirac* D, const double* v0, int niter, double* out) {
    return guard([&] {
        SMG_REQUIRE(D && out, "null argument");
        const size_t n = D->op64->n();
        Field<double> v = up(v0, n), Av(n);
        *out = lambda_max<double>(*D->op64, v.p(), Av.p(), niter, S());
    });
}

// ------------------------------------------------------------------ multigrid
namespace {
MGConfig cfg_of(const smg_mg_params* prm) {
    MGConfig c;
    c.n_levels = prm->n_levels;
    for (int i = 0; i < 4; i++) {
        c.bx[i] = prm->bx[i];
        c.by[i] = prm->by[i];
        c.k[i] = prm->k[i];
    }
    c.null_iters = prm->null_iters;
    c.pre = prm->pre_smooth;
    c.post = prm->post_smooth;
    c.w_cycle = prm->w_cycle;
    c.coarse_maxit = prm->coarse_maxit;
    c.coarse_tol = prm->coarse_tol;
    c.coarse_half = prm->coarse_half;
    c.coarse_solver = prm->coarse_solver;
    c.use_graph = prm->use_graph;
    c.chiral = prm->chiral;
    c.null_prec = prm->null_prec;
    SMG_REQUIRE(c.null_prec == 0 || c.null_prec == 1, "null_prec must be 0 or 1");
    SMG_REQUIRE(c.coarse_solver >= 0 && c.coarse_solver <= 2, "coarse_solver must be 0, 1 or 2");
    SMG_REQUIRE(c.n_levels >= 2 && c.n_levels <= 5, "n_levels must be 2..5");
    for (int l = 0; l + 1 < c.n_levels; l++)
        SMG_REQUIRE(c.k[l] >= 1 && c.k[l] <= 64 && c.bx[l] >= 1 && c.by[l] >= 1, "bad level parameters");
    return c;
}
}  // namespace
int smg_mg_setup(smg_mg** out, smg_dirac* D, const smg_mg_params* prm, const double* starts) {
    return guard([&] {
        SMG_REQUIRE(out && D && prm && starts, "null argument");
        const MGConfig c = cfg_of(prm);
        auto* h = new smg_mg_s();
        h->D = D;
        try { h->mg.reset(new MG(*D->D, c, reinterpret_cast<const c64*>(starts), false)); } catch (...) { delete h; throw; }
        *out = h;
    });
}
int smg_mg_destroy(smg_mg* mg) {
    return guard([&] {
        if (mg) release_batched_cycles(mg->mg.get());
        delete mg;
    });
}
int smg_mg_set_symmetric(smg_mg* mg, double damping, const double* starts) {
    return guard([&] {
        SMG_REQUIRE(mg && starts, "null argument");
        mg->mg->set_symmetric(damping, reinterpret_cast<const c64*>(starts), false);
    });
}
static void mg_prec_dev(smg_mg* mg, int prec, const void* b, void* x) {
    check_prec(prec);
    if (prec == SMG_F64) mg->mg->precondition((const c64*)b, (c64*)x, S());
    else mg->mg->precondition((const c32*)b, (c32*)x, S());
}
int smg_mg_precondition(smg_mg* mg, int prec, const double* b, double* x) {
    return guard([&] {
        SMG_REQUIRE(mg, "null handle");
        check_prec(prec);
        const size_t n = mg->mg->dims[0];
        Field<double> bb = up(b, n), xx(n);
        if (prec == SMG_F64) mg_prec_dev(mg, prec, bb.p(), xx.p());
        else {
            Field<float> b32(n), x32(n);
            launch_convert<float, double>(b32.p(), bb.p(), n, S());
            mg_prec_dev(mg, prec, b32.p(), x32.p());
            launch_convert<double, float>(xx.p(), x32.p(), n, S());
        }
        downc(x, xx.p(), n);
    });
}
int smg_mg_precondition_dev(smg_mg* mg, int prec, const void* b, void* x, void* stream) {
    return guard([&] {
        SMG_REQUIRE(mg && b && x, "bad arguments");
        StreamJoin j(stream);
        mg_prec_dev(mg, prec, b, x);
    });
}
int smg_mg_level_dim(smg_mg* mg, int level, size_t* n) {
    return guard([&] {
        SMG_REQUIRE(mg && n && level >= 0 && level < mg->mg->cfg.n_levels, "bad arguments");
        *n = mg->mg->dims[level];
    });
}
int smg_mg_null_vecs(smg_mg* mg, int level, double* out) {
    return guard([&] {
        SMG_REQUIRE(mg && level >= 0 && level + 1 < mg->mg->cfg.n_levels, "bad level");
        downc(out, mg->mg->nullv[level]->p(), mg->mg->nullv[level]->n);
    });
}
int smg_mg_block_vecs(smg_mg* mg, int level, double* out) {
    return guard([&] {
        SMG_REQUIRE(mg && level >= 0 && level + 1 < mg->mg->cfg.n_levels, "bad level");
        downc(out, mg->mg->P[level]->bv64.p(), mg->mg->P[level]->bv64.n);
    });
}
int smg_mg_coarse_links(smg_mg* mg, int level, double* out) {
    return guard([&] {
        SMG_REQUIRE(mg && level >= 0 && level + 1 < mg->mg->cfg.n_levels, "bad level");
        downc(out, mg->mg->C[level]->l64.p(), mg->mg->C[level]->nlinks());
    });
}
int smg_mg_omega(smg_mg* mg, int level, double* out) {
    return guard([&] {
        SMG_REQUIRE(mg && out && level >= 0 && level < mg->mg->cfg.n_levels, "bad level");
        *out = mg->mg->omega[level];
    });
}

int smg_mg_set_cycle(smg_mg* mg, int pre, int post, int coarse_maxit, double coarse_tol, int use_graph) {
    return guard([&] {
        SMG_REQUIRE(mg && pre >= 0 && post >= 0 && coarse_maxit >= 1, "bad arguments");
        mg->mg->cfg.pre = pre;
        mg->mg->cfg.post = post;
        mg->mg->cfg.coarse_maxit = coarse_maxit;
        mg->mg->cfg.coarse_tol = coarse_tol;
        mg->mg->cfg.use_graph = use_graph;
        mg->mg->reset_graphs();
    });
}

// ------------------------------------------------------------------ solvers
static void fill_info(smg_solve_info* info, const SolveResult& r) {
    if (!info) return;
    info->iterations = r.iterations;
    info->final_residual = r.final_residual;
    info->true_residual = r.true_residual;
    info->restarts = r.restarts;
}
int smg_cg_solve(smg_dirac* D, int op, smg_mg* mg, const double* b, const double* x0, double* x, int maxit,
                 double tol, smg_solve_info* info) {
    return guard([&] {
        SMG_REQUIRE(D, "null handle");
        LinOp<double>* A;
        if (op == SMG_OP_DDAGD) A = D->op64.get();
        else {
            SMG_REQUIRE(op == SMG_OP_SCHUR && D->schur64, "unknown operator (SMG_OP_SCHUR needs an even L)");
            SMG_REQUIRE(mg == nullptr, "the multigrid preconditioner acts on D^dag D only");
            A = D->schur64.get();
        }
        const size_t n = A->n();
        Field<double> bb = up(b, n), xx(n);
        Field<double> xi;
        if (x0) xi = up(x0, n);
        SolveResult r = cg_solve(*A, bb.p(), xx.p(), x0 ? xi.p() : nullptr, mg ? mg->mg.get() : nullptr, maxit, tol,
                                 S());
        r.true_residual = true_residual(*A, bb.p(), xx.p(), S());
        fill_info(info, r);
        downc(x, xx.p(), n);
    });
}
static SolveResult gcr_dev(smg_dirac* D, smg_mg* mg, int prec, const c64* b, c64* x, int maxit, double tol,
                           int trunc, int refp, double inner_tol) {
    MG* m = mg ? mg->mg.get() : nullptr;
    SolveResult r;
    if (prec == SMG_F64) {
        r = gcr_solve<double>(*D->op64, b, x, m, maxit, tol, trunc, refp != 0, S());
        r.true_residual = true_residual(*D->op64, b, x, S());
    } else {
        check_prec(prec);
        r = gcr_solve_mixed(*D->op64, *D->op32, b, x, m, maxit, tol, trunc, inner_tol, S());
    }
    return r;
}
int smg_gcr_solve(smg_dirac* D, smg_mg* mg, int prec, const double* b, double* x, int maxit, double tol, int trunc,
                  int ref_projection, double inner_tol, smg_solve_info* info) {
    return guard([&] {
        SMG_REQUIRE(D, "null handle");
        const size_t n = D->op64->n();
        Field<double> bb = up(b, n), xx(n);
        SolveResult r = gcr_dev(D, mg, prec, bb.p(), xx.p(), maxit, tol, trunc, ref_projection, inner_tol);
        fill_info(info, r);
        downc(x, xx.p(), n);
    });
}
int smg_gcr_solve_dev(smg_dirac* D, smg_mg* mg, int prec, const void* b, void* x, int maxit, double tol, int trunc,
                      double inner_tol, smg_solve_info* info, void* stream) {
    return guard([&] {
        SMG_REQUIRE(D && b && x, "bad arguments");
        StreamJoin j(stream);
        SolveResult r = gcr_dev(D, mg, prec, (const c64*)b, (c64*)x, maxit, tol, trunc, 0, inner_tol);
        fill_info(info, r);
    });
}
int smg_fgmres_solve(smg_dirac* D, smg_mg* mg, const double* b, double* x, int restart, int maxit, double tol,
                     smg_solve_info* info) {
    return guard([&] {
        SMG_REQUIRE(D && restart >= 1, "bad arguments");
        const size_t n = D->op64->n();
        Field<double> bb = up(b, n), xx(n);
        SolveResult r = fgmres_solve(*D->op64, bb.p(), xx.p(), mg ? mg->mg.get() : nullptr, restart, maxit, tol, S());
        fill_info(info, r);
        downc(x, xx.p(), n);
    });
}

int smg_fgmres_solve_ritz(smg_dirac* D, smg_mg* mg, const double* b, double* x, int restart, int maxit, double tol,
                          int n_ritz, smg_solve_info* info, int* n_pairs, double* ritz_vals, double* ritz_vecs) {
    return guard([&] {
        SMG_REQUIRE(D && b && x && restart >= 1 && n_ritz >= 0, "bad arguments");
        SMG_REQUIRE(n_ritz == 0 || (n_pairs && ritz_vals && ritz_vecs), "ritz output buffers required");
        const size_t n = D->op64->n();
        Field<double> bb = up(b, n), xx(n);
        std::unique_ptr<RitzHarvest> rh;
        if (n_ritz > 0) rh.reset(new RitzHarvest(n_ritz, restart, n));
        SolveResult r = fgmres_solve(*D->op64, bb.p(), xx.p(), mg ? mg->mg.get() : nullptr, restart, maxit, tol, S(),
                                     rh.get());
        fill_info(info, r);
        downc(x, xx.p(), n);
        if (n_pairs) *n_pairs = rh ? rh->count() : 0;
        if (rh && rh->count() > 0) {
            std::copy(rh->vals.begin(), rh->vals.end(), ritz_vals);
            downc(ritz_vecs, rh->vecs.p(), (size_t)rh->count() * n);
        }
    });
}

// ------------------------------------------------------------------ many right-hand sides
static void check_nrhs(int N, int maxN) {
    SMG_REQUIRE(N >= 1 && N <= maxN && (N & (N - 1)) == 0, "nrhs must be a power of two <= " + std::to_string(maxN));
}
template <typename T>
static void dirac_kind_b(const Dirac& d, int kind, const cplx<T>* src, cplx<T>* dst, int N, hipStream_t st) {
    const DiracView<T> v = d.view<T>();
    const T mu = (T)d.mu_t;
    switch (kind) {
    case SMG_D: launch_b_dslash<T>(v, dst, src, mu, 0, 0, EPI_NONE, nullptr, nullptr, FinB(), N, st); break;
    case SMG_DDAG: launch_b_dslash<T>(v, dst, src, -mu, 1, 1, EPI_NONE, nullptr, nullptr, FinB(), N, st); break;
    case SMG_G5D: launch_b_dslash<T>(v, dst, src, mu, 0, 1, EPI_NONE, nullptr, nullptr, FinB(), N, st); break;
    case SMG_DDAGD: {
        DdagDOpB<T> op(d, N);   // its workspace returns to the stream-ordered cache
        op.apply(dst, src, st);
        break;
    }
    default: SMG_REQUIRE(false, "unknown Dirac kind");
    }
}
int smg_dirac_apply_batched_dev(smg_dirac* D, int kind, int prec, const void* src, void* dst, int nrhs,
                                void* stream) {
    return guard([&] {
        SMG_REQUIRE(D && src && dst, "bad arguments");
        check_nrhs(nrhs, 64);
        check_prec(prec);
        StreamJoin j(stream);
        if (prec == SMG_F64) dirac_kind_b<double>(*D->D, kind, (const c64*)src, (c64*)dst, nrhs, S());
        else dirac_kind_b<float>(*D->D, kind, (const c32*)src, (c32*)dst, nrhs, S());
    });
}
int smg_coarse_apply_batched(smg_coarse* C, int prec, int nrhs, const double* src, double* dst) {
    return guard([&] {
        SMG_REQUIRE(C && src && dst, "bad arguments");
        check_nrhs(nrhs, 64);
        SMG_REQUIRE(prec == SMG_F64 || prec == SMG_F32 || prec == SMG_F16 || prec == SMG_BF16V,
                    "prec must be SMG_F64, SMG_F32, SMG_F16 or SMG_BF16V");
        const size_t n = C->C->n(), nN = n * nrhs;
        Field<double> s = up(src, nN), d(nN);
        if (prec == SMG_BF16V) {
            Field<float> sb(nN), db(nN);
            DevMem sh(nN * sizeof(cb16)), dh(nN * sizeof(cb16));
            launch_b_pack<float, double>(sb.p(), s.p(), n, nrhs, S());
            launch_h_from_f32(sh.as<cb16>(), sb.p(), n, nrhs, S());
            launch_b_coarse_apply_h(C->C->view16(),
                                    dh.as<cb16>(), sh.as<cb16>(), EPI_NONE, nullptr, FinB(), nrhs, S());
            launch_h_to_f32(db.p(), dh.as<cb16>(), n, nrhs, S());
            launch_b_unpack<double, float>(d.p(), db.p(), n, nrhs, S());
        } else if (prec == SMG_F64) {
            Field<double> sb(nN), db(nN);
            launch_b_pack<double, double>(sb.p(), s.p(), n, nrhs, S());
            CoarseLinOpB<double, c64> op(*C->C, nrhs);
            op.apply(db.p(), sb.p(), S());
            launch_b_unpack<double, double>(d.p(), db.p(), n, nrhs, S());
        } else {
            Field<float> sb(nN), db(nN);
            launch_b_pack<float, double>(sb.p(), s.p(), n, nrhs, S());
            if (prec == SMG_F32) {
                CoarseLinOpB<float, c32> op(*C->C, nrhs);
                op.apply(db.p(), sb.p(), S());
            } else {
                CoarseLinOpB<float, ch16> op(*C->C, nrhs);
                op.apply(db.p(), sb.p(), S());
            }
            launch_b_unpack<double, float>(d.p(), db.p(), n, nrhs, S());
        }
        downc(dst, d.p(), nN);
    });
}
int smg_coarse_apply_batched_dev(smg_coarse* C, int prec, int nrhs, const void* src, void* dst, void* stream) {
    return guard([&] {
        SMG_REQUIRE(C && src && dst, "bad arguments");
        check_nrhs(nrhs, 64);
        SMG_REQUIRE(prec == SMG_F64 || prec == SMG_F32 || prec == SMG_F16, "prec must be SMG_F64, SMG_F32 or SMG_F16");
        StreamJoin j(stream);
        if (prec == SMG_F64) {
            CoarseLinOpB<double, c64> op(*C->C, nrhs);
            op.apply((c64*)dst, (const c64*)src, S());
        } else if (prec == SMG_F32) {
            CoarseLinOpB<float, c32> op(*C->C, nrhs);
            op.apply((c32*)dst, (const c32*)src, S());
        } else {
            CoarseLinOpB<float, ch16> op(*C->C, nrhs);
            op.apply((c32*)dst, (const c32*)src, S());
        }
    });
}
int smg_mg_coarse_apply_dev(smg_mg* mg, int level, int prec, int nrhs, const void* src, void* dst, int dots,
                            void* stream) {
    return guard([&] {
        SMG_REQUIRE(mg && src && dst, "bad arguments");
        MG& m = *mg->mg;
        SMG_REQUIRE(level >= 1 && level < (int)m.dims.size(), "level must name a coarse level of the hierarchy");
        check_nrhs(nrhs, 64);
        SMG_REQUIRE(prec == SMG_F64 || prec == SMG_F32 || prec == SMG_F16 || prec == SMG_BF16V,
                    "prec must be SMG_F64, SMG_F32, SMG_F16 or SMG_BF16V");
        const CoarseOperator& C = *m.C[level - 1];
        StreamJoin j(stream);
        // the dot products {<dst, src>, |dst|^2} per right-hand side go to a scratch block
        static thread_local std::unique_ptr<Field<double>> dots_out;
        if (!dots_out) dots_out.reset(new Field<double>(3 * 64));
        FinB fin;
        if (dots) {
            const Fin f = Context::get().fin(FIN_STORE, reinterpret_cast<double*>(dots_out->p()));
            fin.op = FIN_STORE;
            fin.out = f.out;
            fin.ostride = 3;
            fin.part = f.part;
            fin.counter = f.counter;
        }
        SMG_REQUIRE(dots >= 0 && dots <= 2 && (dots < 2 || prec == SMG_BF16V),
                    "dots must be 0, 1 or (SMG_BF16V only) 2");
        const Epi epi = dots ? EPI_DOTS : EPI_NONE;
        if (prec == SMG_F64) {
            CoarseLinOpB<double, c64> op(C, nrhs);
            op.apply((c64*)dst, (const c64*)src, epi, nullptr, (const c64*)src, fin, S());
        } else if (prec == SMG_F32) {
            CoarseLinOpB<float, c32> op(C, nrhs);
            op.apply((c32*)dst, (const c32*)src, epi, nullptr, (const c32*)src, fin, S());
        } else if (prec == SMG_F16) {
            CoarseLinOpB<float, ch16> op(C, nrhs);
            op.apply((c32*)dst, (const c32*)src, epi, nullptr, (const c32*)src, fin, S());
        } else if (dots == 2) {
            // one Chebyshev step: r and x are scratch here (updated in place, discarded)
            static thread_local std::unique_ptr<DevMem> rx;
            const size_t bytes = C.n() * nrhs * sizeof(cb16);
            if (!rx || rx->bytes < 2 * bytes) rx.reset(new DevMem(2 * bytes));
            FinB fc;
            fc.cheb_c1 = 0.5;
            fc.cheb_c2 = 0.25;
            cb16* r = rx->as<cb16>();
            launch_b_coarse_apply_h(C.view16(), (cb16*)dst, (const cb16*)src, EPI_CHEB, r + C.n() * nrhs, fc, nrhs,
                                    S(), r);
        } else {
            launch_b_coarse_apply_h(C.view16(), (cb16*)dst, (const cb16*)src, epi, (const cb16*)src, fin, nrhs, S());
        }
    });
}
int smg_mg_precondition_batched(smg_mg* mg, int prec, int nrhs, const double* B, double* X) {
    return guard([&] {
        SMG_REQUIRE(mg, "null handle");
        check_nrhs(nrhs, 64);
        check_prec(prec);
        const size_t n = mg->mg->dims[0], nN = n * nrhs;
        Field<double> b = up(B, nN), x(nN), bb(nN), xb(nN);
        launch_b_pack<double, double>(bb.p(), b.p(), n, nrhs, S());
        if (prec == SMG_F64) {
            mg_precondition_b<double>(*mg->mg, bb.p(), xb.p(), nrhs, S());
        } else {
            Field<float> b32(nN), x32(nN);
            launch_convert<float, double>(b32.p(), bb.p(), nN, S());
            mg_precondition_b<float>(*mg->mg, b32.p(), x32.p(), nrhs, S());
            launch_convert<double, float>(xb.p(), x32.p(), nN, S());
        }
        launch_b_unpack<double, double>(x.p(), xb.p(), n, nrhs, S());
        downc(X, x.p(), nN);
    });
}
int smg_mg_set_cheb_smoother(smg_mg* mg, double ratio, int fused, double* hi) {
    return guard([&] {
        SMG_REQUIRE(mg, "null handle");
        SMG_REQUIRE(fused == 0 || fused == 1, "fused must be 0 or 1");
        MG& m = *mg->mg;
        m.set_cheb_smoother(ratio, fused == 1, S());
        if (hi)
            for (int l = 0; l + 1 < m.cfg.n_levels; l++) hi[l] = m.cheb_smooth_hi[l];
    });
}
int smg_mg_set_cheb_interval(smg_mg* mg, double lo, double hi) {
    return guard([&] {
        SMG_REQUIRE(mg, "null handle");
        SMG_REQUIRE(mg->mg->cfg.coarse_solver == 2, "the hierarchy has no Chebyshev coarsest solve (coarse_solver 2)");
        SMG_REQUIRE(lo > 0.0 && hi > lo, "the interval needs 0 < lo < hi");
        mg->mg->cheb_lo = lo;
        mg->mg->cheb_hi = hi;
        mg->mg->reset_graphs();   // captured cycles hold the old coefficients
    });
}
int smg_mg_cheb_interval(smg_mg* mg, double* lo, double* hi) {
    return guard([&] {
        SMG_REQUIRE(mg && lo && hi, "bad arguments");
        SMG_REQUIRE(mg->mg->cfg.coarse_solver == 2, "the hierarchy has no Chebyshev coarsest solve (coarse_solver 2)");
        *lo = mg->mg->cheb_lo;
        *hi = mg->mg->cheb_hi;
    });
}
int smg_mg_coarsest_batched(smg_mg* mg, int nrhs, const double* B, double* X) {
    return guard([&] {
        SMG_REQUIRE(mg && B && X, "bad arguments");
        check_nrhs(nrhs, 64);
        MG& m = *mg->mg;
        const size_t n = m.dims.back(), nN = n * nrhs;
        Field<double> b = up(B, nN), x(nN), bb(nN), xb(nN);
        Field<float> b32(nN), x32(nN);
        launch_b_pack<double, double>(bb.p(), b.p(), n, nrhs, S());
        launch_convert<float, double>(b32.p(), bb.p(), nN, S());
        coarsest_solve_b32(m, b32.p(), x32.p(), nrhs, S());
        launch_convert<double, float>(xb.p(), x32.p(), nN, S());
        launch_b_unpack<double, double>(x.p(), xb.p(), n, nrhs, S());
        downc(X, x.p(), nN);
    });
}
static void fill_batched(const BatchedSolveResult& r, double* true_res, smg_solve_info* info) {
    double worst = 0.0;
    for (double v : r.true_residual) worst = std::max(worst, v);
    if (true_res) std::copy(r.true_residual.begin(), r.true_residual.end(), true_res);
    if (info) {
        info->iterations = r.iterations;
        info->final_residual = worst;
        info->true_residual = worst;
        info->restarts = r.restarts;
    }
}
int smg_gcr_solve_batched(smg_dirac* D, smg_mg* mg, int prec, int nrhs, const double* B, double* X, int maxit,
                          double tol, int trunc, double inner_tol, double* true_res, smg_solve_info* info) {
    return guard([&] {
        SMG_REQUIRE(D, "null handle");
        check_nrhs(nrhs, 32);
        check_prec(prec);
        const size_t nN = 2 * (size_t)D->D->V * nrhs;
        Field<double> b = up(B, nN), x(nN);
        BatchedSolveResult r = gcr_solve_batched(*D->D, mg ? mg->mg.get() : nullptr, prec, b.p(), x.p(), nrhs, maxit,
                                                 tol, trunc, inner_tol, S());
        fill_batched(r, true_res, info);
        downc(X, x.p(), nN);
    });
}
int smg_gcr_solve_batched_dev(smg_dirac* D, smg_mg* mg, int prec, int nrhs, const void* B, void* X, int maxit,
                              double tol, int trunc, double inner_tol, double* true_res, smg_solve_info* info,
                              void* stream) {
    return guard([&] {
        SMG_REQUIRE(D && B && X, "bad arguments");
        check_nrhs(nrhs, 32);
        check_prec(prec);
        StreamJoin j(stream);
        BatchedSolveResult r = gcr_solve_batched(*D->D, mg ? mg->mg.get() : nullptr, prec, (const c64*)B, (c64*)X,
                                                 nrhs, maxit, tol, trunc, inner_tol, S());
        fill_batched(r, true_res, info);
    });
}

// ------------------------------------------------------------------ forces
int smg_force_forward(smg_dirac* D, const double* X, double* chi) { return smg_dirac_apply(D, SMG_D, X, chi); }
int smg_force_backward(smg_dirac* D, const double* chi, const double* X, double* F) {
    return guard([&] {
        SMG_REQUIRE(D, "null handle");
        const size_t n = 2 * (size_t)D->D->V;
        Field<double> c = up(chi, n), xx = up(X, n);
        DevMem f(n * sizeof(double));
        launch_force_backward(D->D->v64(), D->D->csw, c.p(), xx.p(), f.as<double>(), S());
        down(F, f.p, n * sizeof(double));
    });
}
// fermion_force (src/hmc_forces.cpp:134-139) in one kernel: chi = D X is recomputed
// per tile inside the force kernel, never written to memory
static void fermion_force_dev(smg_dirac* D, const c64* X, double* F, hipStream_t st) {
    const Dirac& d = *D->D;
    launch_force_fused(d.v64(), d.csw, d.mu_t, X, F, st);
}
int smg_fermion_force(smg_dirac* D, const double* X, double* F) {
    return guard([&] {
        SMG_REQUIRE(D, "null handle");
        const size_t n = 2 * (size_t)D->D->V;
        Field<double> xx = up(X, n);
        DevMem f(n * sizeof(double));
        fermion_force_dev(D, xx.p(), f.as<double>(), S());
        down(F, f.p, n * sizeof(double));
    });
}
int smg_fermion_force_dev(smg_dirac* D, const void* X, void* F, void* stream) {
    return guard([&] {
        SMG_REQUIRE(D && X && F, "bad arguments");
        StreamJoin j(stream);
        fermion_force_dev(D, (const c64*)X, (double*)F, S());
    });
}
int smg_hopping_force(smg_dirac* D, const double* chi, const double* X, double* F) {
    return guard([&] {
        SMG_REQUIRE(D, "null handle");
        const size_t n = 2 * (size_t)D->D->V;
        Field<double> c = up(chi, n), xx = up(X, n);
        DevMem f(n * sizeof(double));
        launch_hopping_force(D->D->v64(), c.p(), xx.p(), 1.0, 0, f.as<double>(), S());
        down(F, f.p, n * sizeof(double));
    });
}
static void logdet_force_dev(const Dirac& d, double* F, int accumulate, hipStream_t st) {
    const size_t V = (size_t)d.V;
    if (!accumulate) SMG_HIP(hipMemsetAsync(F, 0, 2 * V * sizeof(double), st));
    if (d.csw == 0.0) return;
    DevMem w(V * sizeof(double));
    launch_logdet_weight(d.v64(), d.csw, d.mu_t, w.as<double>(), st);
    launch_clover_insert(d.v64(), w.as<double>(), 0.5, F, st);
}
int smg_logdet_force(smg_dirac* D, double* F) {
    return guard([&] {
        SMG_REQUIRE(D, "null handle");
        SMG_REQUIRE(D->D->L % 2 == 0, "even-odd operators need an even L");
        DevMem f(2 * (size_t)D->D->V * sizeof(double));
        logdet_force_dev(*D->D, f.as<double>(), 0, S());
        down(F, f.p, 2 * (size_t)D->D->V * sizeof(double));
    });
}
// schur_deriv_plus, src/hmc_forces.cpp:763-828, accumulated into F (device)
static void schur_deriv_plus(const Dirac& d, const c64* chi_o, const c64* psi_o, double* F, hipStream_t st) {
    const size_t nh = (size_t)d.V, n = 2 * nh;
    const DiracView<double> v = d.v64();
    Field<double> tmp2(nh), tmp3(nh), chif(n), psif(n);
    launch_eo_hop<double>(v, tmp2.p(), psi_o, 0, 1, d.mu_t, 0, 0, nullptr, EPI_NONE, nullptr, nullptr, Fin(), st);
    // tmp3 = D_ee(-mu)^-1 g5 H (g5 chi): the diagonal commutes with g5
    launch_eo_hop<double>(v, tmp3.p(), chi_o, 0, 1, -d.mu_t, 1, 1, nullptr, EPI_NONE, nullptr, nullptr, Fin(), st);
    if (d.csw != 0.0) {
        DevMem w(nh * sizeof(double));
        launch_eo_force_weights(v, chi_o, psi_o, tmp3.p(), tmp2.p(), w.as<double>(), st);
        launch_clover_insert(v, w.as<double>(), d.csw / 8.0, F, st);
    }
    launch_eo_scatter<double>(d.L, chif.p(), tmp3.p(), 0, st);
    launch_eo_scatter<double>(d.L, chif.p(), chi_o, 1, st);
    launch_eo_scatter<double>(d.L, psif.p(), tmp2.p(), 0, st);
    launch_eo_scatter<double>(d.L, psif.p(), psi_o, 1, st);
    launch_hopping_force(v, chif.p(), psif.p(), -0.5, 1, F, st);
}
// eo_fermion_force, src/hmc_forces.cpp:830-867, on device arrays: F = 2 *
// schur_deriv_plus(y_o, x_o) + the log-det force (a real array of 2V doubles is
// scaled as V complex numbers)
static void eo_force_dev(const Dirac& d, const c64* y_o, const c64* x_o, double* F, hipStream_t st) {
    const size_t V = (size_t)d.V;
    SMG_HIP(hipMemsetAsync(F, 0, 2 * V * sizeof(double), st));
    schur_deriv_plus(d, y_o, x_o, F, st);
    launch_axpby<double>(reinterpret_cast<c64*>(F), reinterpret_cast<const c64*>(F), c64{0, 0}, c64{2.0, 0}, V, st);
    if (d.csw != 0.0) {
        DevMem fl(2 * V * sizeof(double));
        logdet_force_dev(d, fl.as<double>(), 0, st);
        launch_axpby<double>(reinterpret_cast<c64*>(F), fl.as<c64>(), c64{1, 0}, c64{1, 0}, V, st);
    }
}
int smg_eo_fermion_force(smg_dirac* D, const double* x_o, const double* y_o, double* F) {
    return guard([&] {
        SMG_REQUIRE(D, "null handle");
        const Dirac& d = *D->D;
        SMG_REQUIRE(d.L % 2 == 0, "even-odd operators need an even L");
        const size_t V = (size_t)d.V;
        Field<double> xo = up(x_o, V), yo = up(y_o, V);
        DevMem f(2 * V * sizeof(double));
        eo_force_dev(d, yo.p(), xo.p(), f.as<double>(), S());
        down(F, f.p, 2 * V * sizeof(double));
    });
}
int smg_fermion_action(smg_dirac* D, const double* phi, int maxit, double tol, double* S_out, double* X,
                       int* iters) {
    return guard([&] {
        SMG_REQUIRE(D && S_out, "null argument");
        const size_t n = 2 * (size_t)D->D->V;
        Field<double> ph = up(phi, n), xx(n);
        const auto r = cg_solve(*D->op64, ph.p(), xx.p(), nullptr, nullptr, maxit, tol, S());
        if (iters) *iters = r.iterations;
        *S_out = dev_dot<double>(ph.p(), xx.p(), n, S()).re;
        if (X) downc(X, xx.p(), n);
    });
}

// ------------------------------------------------------------------ HMC molecular dynamics
namespace {
// Device state of one trajectory: momenta, the two forces, the pseudofermion and
// the CG solution, all resident for the whole evolution.  Full lattice: phi and X
// have 2V complex entries, the action is Re phi^dag (D^dag D)^-1 phi.  Even-odd
// (src/hmc_trajectory.cpp:53-135): phi_o and x_o have V complex entries (2 x V/2),
// the action is Re phi_o^dag (M^dag M)^-1 phi_o - 2 log det D_ee.
struct MDState {
    smg_dirac* D;
    bool eo;
    size_t V, nphi;
    DevMem pi, gf, ff;
    Field<double> phi, X, Y;
    double max_tres = 0.0;   // the largest FP64 true residual |phi - A X|/|phi| of the solves
    MDState(smg_dirac* d, bool eo_, const double* pi_h, const double* phi_h)
        : D(d), eo(eo_), V((size_t)d->D->V), nphi(eo_ ? (size_t)d->D->V : 2 * (size_t)d->D->V),
          pi(2 * V * sizeof(double)), gf(2 * V * sizeof(double)), ff(2 * V * sizeof(double)), phi(nphi), X(nphi),
          Y(nphi) {
        SMG_REQUIRE(!eo || (d->D->L % 2 == 0 && d->schur64), "the even-odd trajectory needs an even L");
        SMG_HIP(hipMemcpyAsync(pi.p, pi_h, 2 * V * sizeof(double), hipMemcpyHostToDevice, S()));
        if (phi_h) SMG_HIP(hipMemcpyAsync(phi.p(), phi_h, nphi * sizeof(c64), hipMemcpyHostToDevice, S()));
        SMG_HIP(hipStreamSynchronize(S()));
    }
    LinOp<double>& op() { return eo ? static_cast<LinOp<double>&>(*D->schur64) : *D->op64; }
    // gauge force and fermion force, both -dS/dtheta: full lattice X = (D^dag D)^-1 phi
    // and fermion_force (src/hmc_trajectory.cpp:254-343); even-odd x_o = (M^dag M)^-1
    // phi_o, y_o = M x_o and eo_fermion_force (src/hmc_trajectory.cpp:60-110)
    int force(double beta, int maxit, double tol) {
        const Dirac& d = *D->D;
        launch_gauge_force(d.L, d.U64.p(), beta, gf.as<double>(), S());
        const SolveResult r = cg_solve(op(), phi.p(), X.p(), nullptr, nullptr, maxit, tol, S());
        max_tres = std::max(max_tres, true_residual(op(), phi.p(), X.p(), S()));
        if (!eo) {
            fermion_force_dev(D, X.p(), ff.as<double>(), S());
        } else {
            D->schur64->apply_m(Y.p(), X.p(), 0, S());
            eo_force_dev(d, Y.p(), X.p(), ff.as<double>(), S());
        }
        return r.iterations;
    }
    void mom(double scale) { launch_mom_update(pi.as<double>(), gf.as<double>(), ff.as<double>(), scale, 2 * V, S()); }
    void gauge(double scale) { D->D->update_links(pi.as<double>(), scale, S()); }
    // {KE, S_g, S_f, LD} on the current gauge field (src/hmc_trajectory.cpp:87-103, 115-127)
    int hamiltonian(double beta, int maxit, double tol, double* out) {
        const Dirac& d = *D->D;
        out[0] = 0.5 * dev_dot<double>(pi.as<c64>(), pi.as<c64>(), V, S()).re;   // 2V reals as V complex
        Scalar1 sg;
        launch_plaq_sum(d.L, d.U64.p(), sg.fin(), S());
        out[1] = -beta * sg.read();
        const SolveResult r = cg_solve(op(), phi.p(), X.p(), nullptr, nullptr, maxit, tol, S());
        max_tres = std::max(max_tres, true_residual(op(), phi.p(), X.p(), S()));
        out[2] = dev_dot<double>(phi.p(), X.p(), nphi, S()).re;
        out[3] = 0.0;
        if (eo && d.csw != 0.0) {
            Scalar1 ld;
            launch_logdet(d.v64(), d.mu_t, ld.fin(), S());
            out[3] = -2.0 * ld.read();
        }
        return r.iterations;
    }
    // plain_leapfrog_evolve, src/hmc_trajectory.cpp:243-409 (the even-odd trajectory
    // always takes the leapfrog, src/hmc_trajectory.cpp:135-170)
    int evolve(const smg_hmc_params& p) {
        SMG_REQUIRE(p.n_steps >= 1, "n_steps must be >= 1");
        const double dt = p.tau / p.n_steps;
        int cg = 0;
        if (p.omelyan && !eo) {   // P(l e) [Q(e/2) P((1-2l) e) Q(e/2) P(2l e)]^(N-1) ... P(l e), l = 0.1932
            const double lambda = 0.1932;
            cg += force(p.beta, p.cg_maxiter, p.cg_tol);
            mom(lambda * dt);
            for (int step = 0; step < p.n_steps; step++) {
                gauge(0.5 * dt);
                cg += force(p.beta, p.cg_maxiter, p.cg_tol);
                mom((1.0 - 2.0 * lambda) * dt);
                gauge(0.5 * dt);
                cg += force(p.beta, p.cg_maxiter, p.cg_tol);
                mom((step < p.n_steps - 1 ? 2.0 : 1.0) * lambda * dt);
            }
        } else {   // P(dt/2) [Q(dt) P(dt)]^(N-1) Q(dt) P(dt/2)
            cg += force(p.beta, p.cg_maxiter, p.cg_tol);
            mom(0.5 * dt);
            for (int step = 0; step < p.n_steps - 1; step++) {
                gauge(dt);
                cg += force(p.beta, p.cg_maxiter, p.cg_tol);
                mom(dt);
            }
            gauge(dt);
            cg += force(p.beta, p.cg_maxiter, p.cg_tol);
            mom(0.5 * dt);
        }
        return cg;
    }
    void pi_out(double* h) { down(h, pi.p, 2 * V * sizeof(double)); }
};
void check_hmc(const smg_hmc_params* p) {
    SMG_REQUIRE(p, "null parameters");
    SMG_REQUIRE(p->n_steps >= 1 && p->tau > 0.0, "tau > 0 and n_steps >= 1 required");
    SMG_REQUIRE(p->cg_maxiter >= 1 && p->cg_tol > 0.0, "bad CG budget");
}
}  // namespace

int smg_dirac_get_gauge(smg_dirac* D, double* U) {
    return guard([&] {
        SMG_REQUIRE(D && U, "null argument");
        downc(U, D->D->U64.p(), 2 * (size_t)D->D->V);
    });
}
int smg_hmc_evolve(smg_dirac* D, const smg_hmc_params* p, double* pi, const double* phi, int* cg_iters) {
    return guard([&] {
        SMG_REQUIRE(D && pi && phi, "null argument");
        check_hmc(p);
        MDState md(D, p->use_eo != 0, pi, phi);
        const int it = md.evolve(*p);
        md.pi_out(pi);
        if (cg_iters) *cg_iters = it;
    });
}
int smg_hmc_hamiltonian(smg_dirac* D, int use_eo, double beta, const double* pi, const double* phi, int cg_maxiter,
                        double cg_tol, double* out4, int* cg_iters) {
    return guard([&] {
        SMG_REQUIRE(D && pi && phi && out4, "null argument");
        MDState md(D, use_eo != 0, pi, phi);
        const int it = md.hamiltonian(beta, cg_maxiter, cg_tol, out4);
        if (cg_iters) *cg_iters = it;
    });
}
int smg_hmc_trajectory(smg_dirac* D, const smg_hmc_params* p, const double* pi0, const double* eta, double u,
                       smg_hmc_result* res) {
    return guard([&] {
        SMG_REQUIRE(D && pi0 && eta && res, "null argument");
        check_hmc(p);
        Dirac& d = *D->D;
        const size_t V = (size_t)d.V;
        Field<double> saved(2 * V);   // gauge_old, src/hmc_trajectory.cpp:48-50
        launch_copy<double>(saved.p(), d.U64.p(), 2 * V, S());
        MDState md(D, p->use_eo != 0, pi0, nullptr);
        Field<double> e = up(eta, md.nphi);
        if (!md.eo)   // phi = D^dag eta (generate_pseudofermion, src/hmc_trajectory.cpp:32-36)
            launch_dslash<double>(d.v64(), md.phi.p(), e.p(), -d.mu_t, 1, 1, EPI_NONE, nullptr, nullptr, Fin(), 1,
                                  2 * V, S());
        else          // phi_o = M^dag eta_o (generate_pseudofermion_eo, src/dirac.h:504-510)
            D->schur64->apply_m(md.phi.p(), e.p(), 1, S());
        double h0[4], h1[4];
        int cg = md.hamiltonian(p->beta, p->cg_maxiter, p->cg_tol, h0);
        cg += md.evolve(*p);
        cg += md.hamiltonian(p->beta, p->cg_maxiter, p->cg_tol, h1);
        const double dH = (h1[0] + h1[1] + h1[2] + h1[3]) - (h0[0] + h0[1] + h0[2] + h0[3]);
        const bool accept = p->force_accept || dH < 0 || u < std::exp(-dH);
        if (!accept) {
            launch_copy<double>(d.U64.p(), saved.p(), 2 * V, S());
            d.refresh_derived(S());
        }
        SMG_HIP(hipStreamSynchronize(S()));
        res->accepted = accept ? 1 : 0;
        res->dH = dH;
        res->dKE = h1[0] - h0[0];
        res->dSG = h1[1] - h0[1];
        res->dSF = h1[2] - h0[2];
        res->dLD = h1[3] - h0[3];
        res->total_cg_iters = cg;
        res->max_true_residual = md.max_tres;
    });
}

// ------------------------------------------------------------------ multi-timescale MG HMC
namespace {
// coarse_lowmode_force, src/hmc_multiscale.cpp:237-276, on the device; the MR
// smoother carries its residual recursively, the residual handed to the low-mode
// solution is recomputed
void lowmode_force_dev(smg_dirac* D, const Aggregation& P, const Eigenpairs& E, const c64* phi, int smooth, double* F,
                       hipStream_t st) {
    LinOp<double>& A = *D->op64;
    const size_t n = A.n();
    Field<double> x(n), r(n), Ar(n), e(n);
    if (smooth <= 0) {
        lowmode_solution(P, E, phi, x.p(), st);
    } else {
        DevMem sc(64 * sizeof(double));
        mr_smooth<double>(A, x.p(), phi, r.p(), Ar.p(), smooth, true, sc.as<double>(), st);
        A.apply(Ar.p(), x.p(), st);
        launch_copy<double>(r.p(), phi, n, st);
        launch_axpby<double>(r.p(), Ar.p(), c64{-1.0, 0.0}, c64{1.0, 0.0}, n, st);
        lowmode_solution(P, E, r.p(), e.p(), st);
        launch_axpby<double>(x.p(), e.p(), c64{1.0, 0.0}, c64{1.0, 0.0}, n, st);
        mr_smooth<double>(A, x.p(), phi, r.p(), Ar.p(), smooth, false, sc.as<double>(), st);
    }
    fermion_force_dev(D, x.p(), F, st);
}

// The molecular dynamics of hmc_trajectory_mg_multiscale (src/hmc_multiscale.cpp:340-888)
// without tracking, forecasting or prolongator perturbation.  The full-lattice MD
// state (momenta, forces, phi) of the plain trajectory, plus the low-mode force and
// the deflation space (held by the smg_coarse handle, updated in place).
struct MSState : MDState {
    MG* mg;
    const Aggregation& P;
    Eigenpairs& E;
    DevMem fl, zero;
    int lowmode_evals = 0, counter = 0;
    MSState(smg_dirac* d, MG* m, const Aggregation& P_, Eigenpairs& E_, const double* pi_h, const double* phi_h)
        : MDState(d, false, pi_h, phi_h), mg(m), P(P_), E(E_), fl(2 * V * sizeof(double)),
          zero(2 * V * sizeof(double)) {
        SMG_HIP(hipMemsetAsync(zero.p, 0, 2 * V * sizeof(double), S()));
    }
    void low_force(int smooth) { lowmode_force_dev(D, P, E, phi.p(), smooth, fl.as<double>(), S()); }
    // compute_outer_force (:446-512): F_g + F_f - F_low, the solve MG-preconditioned when
    // a hierarchy is given
    int outer_force(const smg_ms_params& p) {
        const Dirac& d = *D->D;
        launch_gauge_force(d.L, d.U64.p(), p.beta, gf.as<double>(), S());
        const SolveResult r = cg_solve(*D->op64, phi.p(), X.p(), nullptr, mg, p.cg_maxiter, p.cg_tol, S());
        fermion_force_dev(D, X.p(), ff.as<double>(), S());
        low_force(p.inner_smooth);
        launch_mom_update(ff.as<double>(), fl.as<double>(), zero.as<double>(), -1.0, 2 * V, S());   // ff -= fl
        return r.iterations;
    }
    void kick_low(double scale) {
        launch_mom_update(pi.as<double>(), fl.as<double>(), zero.as<double>(), scale, 2 * V, S());
    }
    // inner_integrator (:667-685) with maybe_refresh_deflation (:601-614)
    void inner(const smg_ms_params& p, double dt_total) {
        const int ni = p.n_inner;
        const double dti = dt_total / ni;
        low_force(p.inner_smooth);
        lowmode_evals++;
        kick_low(0.5 * dti);
        for (int i = 0; i < ni; i++) {
            gauge(dti);
            counter++;
            if (p.defl_refresh > 0 && counter % p.defl_refresh == 0) {   // evolve_coarse_deflation
                CoarseOperator Cn(P.A.nbx, P.A.nby, P.A.k);
                Cn.build(*D->op64, P, S());
                CoarseLinOp<double, c64> op(Cn);
                rr_evolve(op, E, S());
            }
            low_force(p.inner_smooth);
            lowmode_evals++;
            kick_low(i < ni - 1 ? dti : 0.5 * dti);
        }
    }
    int evolve(const smg_ms_params& p) {
        const double h = p.tau / p.n_outer;
        int cg = 0;
        if (p.omelyan) {
            const double lam = 0.1932;
            cg += outer_force(p);
            mom(lam * h);
            for (int o = 0; o < p.n_outer; o++) {
                inner(p, 0.5 * h);
                cg += outer_force(p);
                mom((1.0 - 2.0 * lam) * h);
                inner(p, 0.5 * h);
                cg += outer_force(p);
                mom((o < p.n_outer - 1 ? 2.0 : 1.0) * lam * h);
            }
        } else {
            cg += outer_force(p);
            mom(0.5 * h);
            for (int o = 0; o < p.n_outer; o++) {
                inner(p, h);
                cg += outer_force(p);
                mom(o < p.n_outer - 1 ? h : 0.5 * h);
            }
        }
        return cg;
    }
    // KE + S_g + S_f with the (preconditioned) CG of the trajectory
    int ham(const smg_ms_params& p, double* h3) {
        double h4[4];
        const Dirac& d = *D->D;
        h4[0] = 0.5 * dev_dot<double>(pi.as<c64>(), pi.as<c64>(), V, S()).re;
        Scalar1 sg;
        launch_plaq_sum(d.L, d.U64.p(), sg.fin(), S());
        h4[1] = -p.beta * sg.read();
        const SolveResult r = cg_solve(*D->op64, phi.p(), X.p(), nullptr, mg, p.cg_maxiter, p.cg_tol, S());
        h4[2] = dev_dot<double>(phi.p(), X.p(), nphi, S()).re;
        for (int i = 0; i < 3; i++) h3[i] = h4[i];
        return r.iterations;
    }
};
void check_ms(smg_dirac* D, smg_prolongator* P, smg_coarse* defl, const smg_ms_params* p) {
    SMG_REQUIRE(D && P && defl && p, "null argument");
    SMG_REQUIRE(p->n_outer >= 1 && p->n_inner >= 1 && p->tau > 0.0, "tau > 0, n_outer >= 1 and n_inner >= 1");
    SMG_REQUIRE(p->cg_maxiter >= 1 && p->cg_tol > 0.0, "bad CG budget");
    SMG_REQUIRE(P->P->fine_n() == 2 * (size_t)D->D->V && P->P->A.nd == 2, "the prolongator is not of this lattice");
    SMG_REQUIRE(defl->defl.vals.empty() || defl->C->n() == P->P->coarse_n(),
                "the deflation space is not of the prolongator's coarse space");
}
}  // namespace

int smg_coarse_lowmode_force(smg_dirac* D, smg_prolongator* P, smg_coarse* defl, const double* phi, int smooth_iters,
                             double* F) {
    return guard([&] {
        SMG_REQUIRE(D && P && defl && phi && F, "null argument");
        SMG_REQUIRE(P->P->fine_n() == 2 * (size_t)D->D->V && defl->C->n() == P->P->coarse_n(), "shape mismatch");
        const size_t V = (size_t)D->D->V;
        Field<double> ph = up(phi, 2 * V);
        DevMem f(2 * V * sizeof(double));
        lowmode_force_dev(D, *P->P, defl->defl, ph.p(), smooth_iters, f.as<double>(), S());
        down(F, f.p, 2 * V * sizeof(double));
    });
}
int smg_coarse_evolve_deflation(smg_coarse* defl, smg_coarse* A_new) {
    return guard([&] {
        SMG_REQUIRE(defl && A_new && defl->C->n() == A_new->C->n(), "bad arguments");
        CoarseLinOp<double, c64> op(*A_new->C);
        rr_evolve(op, defl->defl, S());
    });
}
int smg_hmc_ms_evolve(smg_dirac* D, smg_mg* mg, smg_prolongator* P, smg_coarse* defl, const smg_ms_params* p,
                      double* pi, const double* phi, int* cg_iters, int* lowmode_evals) {
    return guard([&] {
        check_ms(D, P, defl, p);
        SMG_REQUIRE(pi && phi, "null argument");
        MSState ms(D, mg ? mg->mg.get() : nullptr, *P->P, defl->defl, pi, phi);
        const int it = ms.evolve(*p);
        ms.pi_out(pi);
        if (cg_iters) *cg_iters = it;
        if (lowmode_evals) *lowmode_evals = ms.lowmode_evals;
    });
}
int smg_hmc_ms_trajectory(smg_dirac* D, smg_mg* mg, smg_prolongator* P, smg_coarse* defl, const smg_ms_params* p,
                          const double* pi0, const double* eta, double u, smg_ms_result* res) {
    return guard([&] {
        check_ms(D, P, defl, p);
        SMG_REQUIRE(pi0 && eta && res, "null argument");
        Dirac& d = *D->D;
        const size_t V = (size_t)d.V;
        Field<double> saved(2 * V);
        launch_copy<double>(saved.p(), d.U64.p(), 2 * V, S());
        MSState ms(D, mg ? mg->mg.get() : nullptr, *P->P, defl->defl, pi0, nullptr);
        Field<double> e = up(eta, 2 * V);   // phi = D^dag eta (generate_pseudofermion)
        launch_dslash<double>(d.v64(), ms.phi.p(), e.p(), -d.mu_t, 1, 1, EPI_NONE, nullptr, nullptr, Fin(), 1, 2 * V,
                              S());
        double h0[3], h1[3];
        int cg = ms.ham(*p, h0);
        cg += ms.evolve(*p);
        cg += ms.ham(*p, h1);
        const double dH = (h1[0] + h1[1] + h1[2]) - (h0[0] + h0[1] + h0[2]);
        const bool accept = p->force_accept || dH < 0 || u < std::exp(-dH);
        if (!accept) {   // the gauge field is restored; the deflation space keeps its evolution
            launch_copy<double>(d.U64.p(), saved.p(), 2 * V, S());
            d.refresh_derived(S());
        }
        SMG_HIP(hipStreamSynchronize(S()));
        res->accepted = accept ? 1 : 0;
        res->dH = dH;
        res->dKE = h1[0] - h0[0];
        res->dSG = h1[1] - h0[1];
        res->dSF = h1[2] - h0[2];
        res->highmode_cg_iters = cg;
        res->lowmode_force_evals = ms.lowmode_evals;
    });
}

// ------------------------------------------------------------------ nested integrator, gauge force on the fine scale
namespace {
// The fermion force on the outer scale, the gauge force on the fine scale: each outer
// position update Q(h) is n_gauge inner steps of h / n_gauge kicking with the gauge
// force alone; the outer kicks use the fermion force alone.  The solves of D^dag D X =
// phi are FP64 CG (preconditioned by the FP64 cycle when a hierarchy is given) or the
// mixed-precision GCR(10) inside FP64 refinement.
struct NestedState : MDState {
    smg_mg* mgh;
    const smg_nested_params& p;
    DevMem zero;
    int iters = 0, solves = 0;
    bool have_x = false;      // X holds a solution on an earlier field (the chronological guess)
    double max_tres = 0.0;    // the largest FP64 true residual of the trajectory's solves
    NestedState(smg_dirac* d, smg_mg* m, const smg_nested_params& p_, const double* pi_h, const double* phi_h)
        : MDState(d, false, pi_h, phi_h), mgh(m), p(p_), zero(2 * V * sizeof(double)) {
        SMG_HIP(hipMemsetAsync(zero.p, 0, 2 * V * sizeof(double), S()));
    }
    void solve() {
        SolveResult r;
        const bool guess = p.chrono_guess && have_x;
        if (p.solver == SMG_F32) {
            r = gcr_solve_mixed(*D->op64, *D->op32, phi.p(), X.p(), mgh->mg.get(), p.cg_maxiter, p.cg_tol, 10,
                                p.inner_tol, S(), guess);
        } else {
            r = cg_solve(*D->op64, phi.p(), X.p(), guess ? X.p() : nullptr, mgh ? mgh->mg.get() : nullptr,
                         p.cg_maxiter, p.cg_tol, S());
            r.true_residual = true_residual(*D->op64, phi.p(), X.p(), S());
        }
        have_x = true;
        max_tres = std::max(max_tres, r.true_residual);
        iters += r.iterations;
        solves++;
    }
    void fermion_kick(double scale) {
        solve();
        fermion_force_dev(D, X.p(), ff.as<double>(), S());
        launch_mom_update(pi.as<double>(), ff.as<double>(), zero.as<double>(), scale, 2 * V, S());
    }
    void gauge_kick(double scale) {
        const Dirac& d = *D->D;
        launch_gauge_force(d.L, d.U64.p(), p.beta, gf.as<double>(), S());
        launch_mom_update(pi.as<double>(), gf.as<double>(), zero.as<double>(), scale, 2 * V, S());
    }
    void inner(double h) {
        const int ng = p.n_gauge;
        const double dt = h / ng;
        if (p.omelyan_inner) {
            const double lam = 0.1932;
            gauge_kick(lam * dt);
            for (int s = 0; s < ng; s++) {
                gauge(0.5 * dt);
                gauge_kick((1.0 - 2.0 * lam) * dt);
                gauge(0.5 * dt);
                gauge_kick((s < ng - 1 ? 2.0 : 1.0) * lam * dt);
            }
        } else {
            gauge_kick(0.5 * dt);
            for (int s = 0; s < ng - 1; s++) {
                gauge(dt);
                gauge_kick(dt);
            }
            gauge(dt);
            gauge_kick(0.5 * dt);
        }
    }
    int moved = 0;
    // after every mg_refresh-th outer position update the hierarchy follows the field
    void follow() {
        if (p.mg_refresh > 0 && mgh && ++moved % p.mg_refresh == 0) mgh->mg->refresh(p.mg_refresh_iters, S());
    }
    void evolve() {
        const double h = p.tau / p.n_outer;
        if (p.omelyan_outer) {
            const double lam = 0.1932;
            fermion_kick(lam * h);
            for (int o = 0; o < p.n_outer; o++) {
                inner(0.5 * h);
                follow();
                fermion_kick((1.0 - 2.0 * lam) * h);
                inner(0.5 * h);
                follow();
                fermion_kick((o < p.n_outer - 1 ? 2.0 : 1.0) * lam * h);
            }
        } else {
            fermion_kick(0.5 * h);
            for (int o = 0; o < p.n_outer; o++) {
                inner(h);
                follow();
                fermion_kick(o < p.n_outer - 1 ? h : 0.5 * h);
            }
        }
    }
    // KE, S_g, S_f with the trajectory's solver
    void ham(double* h3) {
        const Dirac& d = *D->D;
        h3[0] = 0.5 * dev_dot<double>(pi.as<c64>(), pi.as<c64>(), V, S()).re;
        Scalar1 sg;
        launch_plaq_sum(d.L, d.U64.p(), sg.fin(), S());
        h3[1] = -p.beta * sg.read();
        solve();
        h3[2] = dev_dot<double>(phi.p(), X.p(), nphi, S()).re;
    }
};
void check_nested(smg_dirac* D, smg_mg* mg, const smg_nested_params* p) {
    SMG_REQUIRE(D && p, "null argument");
    SMG_REQUIRE(p->n_outer >= 1 && p->n_gauge >= 1 && p->tau > 0.0, "tau > 0, n_outer >= 1 and n_gauge >= 1");
    SMG_REQUIRE(p->cg_maxiter >= 1 && p->cg_tol > 0.0, "bad solver budget");
    check_prec(p->solver);
    SMG_REQUIRE(p->solver == SMG_F64 || mg, "the mixed-precision solver needs a hierarchy");
    SMG_REQUIRE(p->mg_refresh >= 0 && p->mg_refresh_iters >= 0, "bad refresh parameters");
}
}  // namespace

int smg_hmc_nested_evolve(smg_dirac* D, smg_mg* mg, const smg_nested_params* p, double* pi, const double* phi,
                          int* solver_iters) {
    return guard([&] {
        check_nested(D, mg, p);
        SMG_REQUIRE(pi && phi, "null argument");
        NestedState ns(D, mg, *p, pi, phi);
        ns.evolve();
        ns.pi_out(pi);
        if (solver_iters) *solver_iters = ns.iters;
    });
}
int smg_hmc_nested_trajectory(smg_dirac* D, smg_mg* mg, const smg_nested_params* p, const double* pi0,
                              const double* eta, double u, smg_hmc_result* res) {
    return guard([&] {
        check_nested(D, mg, p);
        SMG_REQUIRE(pi0 && eta && res, "null argument");
        Dirac& d = *D->D;
        const size_t V = (size_t)d.V;
        Field<double> saved(2 * V);
        launch_copy<double>(saved.p(), d.U64.p(), 2 * V, S());
        // a hierarchy that follows the proposal keeps its near-null vectors for a rejection
        std::vector<Field<double>> saved_null;
        const bool follows = p->mg_refresh > 0 && mg;
        if (follows)
            for (auto& v : mg->mg->nullv) {
                saved_null.emplace_back(v->n);
                launch_copy<double>(saved_null.back().p(), v->p(), v->n, S());
            }
        NestedState ns(D, mg, *p, pi0, nullptr);
        Field<double> e = up(eta, 2 * V);   // phi = D^dag eta
        launch_dslash<double>(d.v64(), ns.phi.p(), e.p(), -d.mu_t, 1, 1, EPI_NONE, nullptr, nullptr, Fin(), 1, 2 * V,
                              S());
        double h0[3], h1[3];
        ns.ham(h0);
        ns.evolve();
        ns.ham(h1);
        const double dH = (h1[0] + h1[1] + h1[2]) - (h0[0] + h0[1] + h0[2]);
        const bool accept = p->force_accept || dH < 0 || u < std::exp(-dH);
        if (!accept) {
            launch_copy<double>(d.U64.p(), saved.p(), 2 * V, S());
            d.refresh_derived(S());
            // the hierarchy followed the rejected proposal: its near-null vectors come back and
            // the aggregations and Galerkin operators are rebuilt on the kept field, which
            // restores the hierarchy the trajectory started from
            if (follows) {
                for (size_t l = 0; l < saved_null.size(); l++)
                    launch_copy<double>(mg->mg->nullv[l]->p(), saved_null[l].p(), saved_null[l].n, S());
                mg->mg->refresh(0, S());
            }
        }
        SMG_HIP(hipStreamSynchronize(S()));
        res->accepted = accept ? 1 : 0;
        res->dH = dH;
        res->dKE = h1[0] - h0[0];
        res->dSG = h1[1] - h0[1];
        res->dSF = h1[2] - h0[2];
        res->dLD = 0.0;
        res->total_cg_iters = ns.iters;
        res->max_true_residual = ns.max_tres;
    });
}

// ------------------------------------------------------------------ U(1) heatbath (no reference counterpart)
int smg_dirac_heatbath(smg_dirac* D, double beta, int n_sweeps, unsigned long long seed, unsigned first_sweep) {
    return guard([&] {
        SMG_REQUIRE(D && n_sweeps >= 0, "bad arguments");
        Dirac& d = *D->D;
        launch_heatbath(d.U64.p(), d.L, beta, n_sweeps, seed, first_sweep, S());
        d.refresh_derived(S());
        SMG_HIP(hipStreamSynchronize(S()));
    });
}
int smg_heatbath(int L, double* U, double beta, int n_sweeps, unsigned long long seed, unsigned first_sweep) {
    return guard([&] {
        SMG_REQUIRE(U && L >= 2 && L % 2 == 0 && n_sweeps >= 0, "bad arguments (L even)");
        const size_t n = 2 * (size_t)L * L;
        Field<double> u = up(U, n);
        launch_heatbath(u.p(), L, beta, n_sweeps, seed, first_sweep, S());
        downc(U, u.p(), n);
    });
}

// ------------------------------------------------------------------ many lattices in one launch
int smg_lattices_create(smg_lattices** out, int G, const int* L, const double* U, const double* mass, double r,
                        double csw, double mu_t) {
    return guard([&] {
        // one completion counter per lattice (of the context's 1024) in every reducing launch
        SMG_REQUIRE(out && L && U && mass && G >= 1 && G <= 1023, "bad arguments (1 <= G <= 1023)");
        auto* h = new smg_lattices_s();
        try {
            size_t off = 0;
            std::vector<long long> sites(G);
            for (int g = 0; g < G; g++) sites[g] = (long long)L[g] * L[g];
            const std::vector<int> nwg = lattice_workgroups(sites);
            for (int g = 0; g < G; g++) {
                SMG_REQUIRE(L[g] >= 2, "every L must be >= 2");
                h->D.emplace_back(new Dirac(L[g], reinterpret_cast<const c64*>(U) + off, mass[g], r, csw, mu_t));
                off += 2 * (size_t)L[g] * L[g];
                const DiracView<double> v = h->D.back()->v64();
                LatDesc e;
                e.U = v.U;
                e.C = v.C;
                e.diag = v.diag;
                e.r = v.r;
                e.mu = mu_t;
                e.L = L[g];
                e.wg0 = h->nwg;
                e.nwg = nwg[g];
                e.site0 = (long long)h->sites;
                h->h.push_back(e);
                h->sites += (size_t)L[g] * L[g];
                h->nwg += nwg[g];
            }
            SMG_REQUIRE((size_t)h->nwg * 3 <= (size_t(1) << 21), "too many sites for one launch's partial sums");
            h->d.alloc(h->h.size() * sizeof(LatDesc));
            SMG_HIP(hipMemcpyAsync(h->d.p, h->h.data(), h->h.size() * sizeof(LatDesc), hipMemcpyHostToDevice, S()));
            Context::get().sync();
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}
int smg_lattices_destroy(smg_lattices* H) { return guard([&] { delete H; }); }
int smg_lattices_size(smg_lattices* H, size_t* n) {
    return guard([&] {
        SMG_REQUIRE(H && n, "null argument");
        *n = 2 * H->sites;
    });
}
namespace {
LatFin latfin(int op, double* out, double* aux) {
    const Fin f = Context::get().fin(op, out, aux);
    LatFin l;
    l.op = op;
    l.out = out;
    l.aux = aux;
    l.part = f.part;
    l.counter = f.counter;
    return l;
}
void ml_apply(smg_lattices* H, int kind, c64* dst, const c64* src, c64* tmp, const LatFin& skip_fin,
              const LatFin* dots, const c64* aux, hipStream_t st) {
    const LatDesc* d = H->d.as<LatDesc>();
    const int G = (int)H->h.size();
    if (kind == SMG_DDAGD) {
        launch_ml_dslash(d, G, H->nwg, tmp, src, 1.0, 0, 0, EPI_NONE, nullptr, skip_fin, st);
        launch_ml_dslash(d, G, H->nwg, dst, tmp, -1.0, 1, 1, dots ? EPI_DOTS : EPI_NONE, aux, dots ? *dots : skip_fin,
                         st);
        return;
    }
    SMG_REQUIRE(kind == SMG_D || kind == SMG_DDAG || kind == SMG_G5D, "unknown operator kind");
    const double ms = kind == SMG_DDAG ? -1.0 : 1.0;
    const int g5in = kind == SMG_DDAG ? 1 : 0, g5out = kind == SMG_D ? 0 : 1;
    launch_ml_dslash(d, G, H->nwg, dst, src, ms, g5in, g5out, EPI_NONE, nullptr, skip_fin, st);
}
}  // namespace
int smg_lattices_apply(smg_lattices* H, int kind, const double* src, double* dst) {
    return guard([&] {
        SMG_REQUIRE(H && src && dst, "null argument");
        const size_t n = 2 * H->sites;
        Field<double> s_ = up(src, n), d_(n), t(n);
        ml_apply(H, kind, d_.p(), s_.p(), t.p(), LatFin(), nullptr, nullptr, S());
        downc(dst, d_.p(), n);
    });
}
int smg_lattices_apply_dev(smg_lattices* H, int kind, const void* src, void* dst, void* stream) {
    return guard([&] {
        SMG_REQUIRE(H && src && dst, "null argument");
        StreamJoin j(stream);
        Field<double> t(kind == SMG_DDAGD ? 2 * H->sites : 1);
        ml_apply(H, kind, static_cast<c64*>(dst), static_cast<const c64*>(src), t.p(), LatFin(), nullptr, nullptr, S());
    });
}
int smg_lattices_cg(smg_lattices* H, const double* b, double* x, int maxit, double tol, int* iters,
                    double* true_res) {
    return guard([&] {
        SMG_REQUIRE(H && b && x && maxit >= 0 && tol > 0.0, "bad arguments");
        const int G = (int)H->h.size();
        const size_t n = 2 * H->sites;
        const LatDesc* d = H->d.as<LatDesc>();
        Field<double> bb = up(b, n), xx(n), r(n), p(n), Ap(n), t(n);
        DevMem sc((4 * (size_t)G + 16 * (size_t)G + 4 * (size_t)G) * sizeof(double) + 64);
        double* alpha = sc.as<double>();
        double* beta = alpha + 2 * G;
        double* aux = beta + 2 * G;
        double* nrm = aux + 16 * G;   // [G] |b - A x|^2, then [G] |b|^2
        int* done = reinterpret_cast<int*>(nrm + 4 * G);
        int* late = done + 1;
        launch_ml_cg_state(aux, G, tol * tol, done, late, S());
        launch_ml_cg_init(d, G, H->nwg, bb.p(), xx.p(), r.p(), p.p(), latfin(FIN_CG_INIT, nullptr, aux), S());
        LatFin skip;
        skip.skip = done;
        LatFin fa = latfin(FIN_CG_ALPHA, alpha, aux), fbt = latfin(FIN_CG_BETA, beta, aux);
        fa.skip = done;
        fbt.skip = done;
        fbt.done = done;
        for (int it = 0; it < maxit; it++) {
            ml_apply(H, SMG_DDAGD, Ap.p(), p.p(), t.p(), skip, &fa, p.p(), S());
            launch_ml_cg_r(d, G, H->nwg, r.p(), Ap.p(), alpha, fbt, late, S());
            launch_ml_cg_xp(d, G, H->nwg, xx.p(), p.p(), r.p(), alpha, beta, late, S());
            if (it % 32 == 31) {   // stop enqueueing once every lattice has converged
                int h = 0;
                SMG_HIP(hipMemcpyAsync(&h, done, sizeof(int), hipMemcpyDeviceToHost, S()));
                SMG_HIP(hipStreamSynchronize(S()));
                if (h) break;
            }
        }
        // iterations and FP64 true residuals per lattice
        std::vector<double> ah(16 * (size_t)G);
        SMG_HIP(hipMemcpyAsync(ah.data(), aux, ah.size() * sizeof(double), hipMemcpyDeviceToHost, S()));
        ml_apply(H, SMG_DDAGD, Ap.p(), xx.p(), t.p(), LatFin(), nullptr, nullptr, S());
        launch_axpby<double>(Ap.p(), bb.p(), c64{1.0, 0.0}, c64{-1.0, 0.0}, n, S());   // b - A x
        launch_ml_norm2(d, G, H->nwg, Ap.p(), latfin(FIN_STORE, nrm, nullptr), S());
        // FIN_STORE writes one value per lattice at out + t*ostride: use stride 1
        std::vector<double> hn(2 * (size_t)G);
        SMG_HIP(hipStreamSynchronize(S()));
        {
            LatFin fs = latfin(FIN_STORE, nrm + 2 * G, nullptr);
            fs.ostride = 1;
            launch_ml_norm2(d, G, H->nwg, bb.p(), fs, S());
        }
        SMG_HIP(hipMemcpyAsync(hn.data(), nrm, 2 * G * sizeof(double), hipMemcpyDeviceToHost, S()));
        std::vector<double> hb(G);
        SMG_HIP(hipMemcpyAsync(hb.data(), nrm + 2 * G, G * sizeof(double), hipMemcpyDeviceToHost, S()));
        SMG_HIP(hipStreamSynchronize(S()));
        for (int g = 0; g < G; g++) {
            if (iters) iters[g] = (int)ah[(size_t)g * 16 + 3];
            if (true_res) true_res[g] = hb[g] > 0.0 ? std::sqrt(hn[2 * g] / hb[g]) : 0.0;
        }
        downc(x, xx.p(), n);
    });
}

// ------------------------------------------------------------------ multigrid over many lattices
int smg_lattices_mg_setup(smg_lattices_mg** out, smg_lattices* H, const smg_mg_params* prm, const double* starts) {
    return guard([&] {
        SMG_REQUIRE(out && H && prm && starts, "null argument");
        const MGConfig c = cfg_of(prm);
        std::vector<const Dirac*> dirac;
        for (const auto& d : H->D) dirac.push_back(d.get());
        auto* h = new smg_lattices_mg_s();
        h->H = H;
        try {
            h->mg.reset(new LatticeMG(dirac, H->d.as<LatDesc>(), H->nwg, c, reinterpret_cast<const c64*>(starts)));
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}
int smg_lattices_mg_destroy(smg_lattices_mg* M) { return guard([&] { delete M; }); }
int smg_lattices_mg_precondition(smg_lattices_mg* M, const double* b, double* x) {
    return guard([&] {
        SMG_REQUIRE(M && b && x, "null argument");
        const size_t n = 2 * M->H->sites;
        Field<double> bb = up(b, n), xx(n);
        M->mg->precondition(bb.p(), xx.p(), S());
        downc(x, xx.p(), n);
    });
}
int smg_lattices_mg_precondition_dev(smg_lattices_mg* M, const void* b, void* x, void* stream) {
    return guard([&] {
        SMG_REQUIRE(M && b && x, "null argument");
        StreamJoin j(stream);
        M->mg->precondition(static_cast<const c64*>(b), static_cast<c64*>(x), S());
    });
}
int smg_lattices_gcr(smg_lattices_mg* M, const double* b, double* x, int maxit, double tol, int trunc, int* iters,
                     double* true_res, int* total) {
    return guard([&] {
        SMG_REQUIRE(M && b && x && maxit >= 0 && tol > 0.0, "bad arguments");
        const size_t n = 2 * M->H->sites;
        Field<double> bb = up(b, n), xx(n);
        const int t = M->mg->gcr(bb.p(), xx.p(), maxit, tol, trunc, iters, true_res, S());
        if (total) *total = t;
        downc(x, xx.p(), n);
    });
}
int smg_lattices_gcr_dev(smg_lattices_mg* M, const void* b, void* x, int maxit, double tol, int trunc, int* iters,
                         double* true_res, int* total, void* stream) {
    return guard([&] {
        SMG_REQUIRE(M && b && x && maxit >= 0 && tol > 0.0, "bad arguments");
        StreamJoin j(stream);
        const int t = M->mg->gcr(static_cast<const c64*>(b), static_cast<c64*>(x), maxit, tol, trunc, iters,
                                 true_res, S());
        if (total) *total = t;
    });
}

// ------------------------------------------------------------------ gauge field files (host I/O)
namespace {
// GaugeField::avg_plaq, src/gauge.h:26-36, summed in site order (the reference's
// single-thread order)
double host_avg_plaq(int L, const c64* U) {
    const int V = L * L;
    double sum = 0.0;
    for (int s = 0; s < V; s++) {
        const int x = s % L, y = s / L;
        const int sxp = (x + 1) % L + L * y, syp = x + L * ((y + 1) % L);
        const std::complex<double> u0(U[s].re, U[s].im), u1p(U[V + sxp].re, U[V + sxp].im);
        const std::complex<double> u0p(U[syp].re, U[syp].im), u1(U[V + s].re, U[V + s].im);
        sum += std::real(u0 * u1p * std::conj(u0p) * std::conj(u1));
    }
    return sum / V;
}
struct File {
    FILE* f;
    File(const char* path, const char* mode) : f(std::fopen(path, mode)) {
        SMG_REQUIRE(f != nullptr, std::string("cannot open ") + path);
    }
    ~File() {
        if (f) std::fclose(f);
    }
    void write(const void* p, size_t bytes) {
        SMG_REQUIRE(std::fwrite(p, 1, bytes, f) == bytes, "short write");
    }
    void read(void* p, size_t bytes) {
        SMG_REQUIRE(std::fread(p, 1, bytes, f) == bytes, "short read");
    }
};
}  // namespace

int smg_gauge_write_raw(const char* path, int L, const double* U) {
    return guard([&] {
        SMG_REQUIRE(path && U && L >= 2, "bad arguments");
        const int V = L * L;
        File f(path, "wb");
        f.write(&V, sizeof(V));
        f.write(U, 2 * (size_t)V * sizeof(c64));
    });
}
int smg_gauge_read_raw(const char* path, int L, double* U) {
    return guard([&] {
        SMG_REQUIRE(path && U && L >= 2, "bad arguments");
        File f(path, "rb");
        int V = 0;
        f.read(&V, sizeof(V));
        SMG_REQUIRE(V == L * L, "volume in the file does not match L");
        f.read(U, 2 * (size_t)V * sizeof(c64));
    });
}
int smg_gauge_write(const char* path, int L, const double* U, double beta, double mass) {
    return guard([&] {
        SMG_REQUIRE(path && U && L >= 2, "bad arguments");
        smg_gauge_header h{};
        h.magic = 0x53574E47u;
        h.L = (uint32_t)L;
        h.beta = beta;
        h.mass = mass;
        h.avg_plaq = host_avg_plaq(L, reinterpret_cast<const c64*>(U));
        File f(path, "wb");
        f.write(&h, sizeof(h));
        f.write(U, 2 * (size_t)L * L * sizeof(c64));
    });
}
int smg_gauge_read(const char* path, int L, double* U, smg_gauge_header* hdr) {
    return guard([&] {
        SMG_REQUIRE(path && U && L >= 2, "bad arguments");
        File f(path, "rb");
        smg_gauge_header h{};
        f.read(&h, sizeof(h));
        SMG_REQUIRE(h.magic == 0x53574E47u, std::string("bad magic number in ") + path);
        SMG_REQUIRE((int)h.L == L, "lattice size in the file does not match L");
        const size_t n = 2 * (size_t)L * L;
        f.read(U, n * sizeof(c64));
        c64* u = reinterpret_cast<c64*>(U);
        for (size_t i = 0; i < n; i++) {   // links renormalised to |U| = 1 (|U| as cabs: hypot)
            const double a = std::hypot(u[i].re, u[i].im);
            u[i].re /= a;
            u[i].im /= a;
        }
        if (hdr) *hdr = h;
    });
}
int smg_therm_cfg_path(int L, double beta, int n_therm, char* out, size_t cap) {
    return guard([&] {
        SMG_REQUIRE(out && cap > 0, "bad arguments");
        const std::string p = "gauge_L" + std::to_string(L) + "_b" + std::to_string(beta).substr(0, 4) + "_t" +
                              std::to_string(n_therm) + ".bin";
        SMG_REQUIRE(p.size() + 1 <= cap, "output buffer too small");
        std::memcpy(out, p.c_str(), p.size() + 1);
    });
}

// ---------------------------------------------------------------- eigensolvers and deflation
int smg_random_vec(int n, unsigned seed, double* out) {
    return guard([&] {
        SMG_REQUIRE(n >= 0 && (n == 0 || out), "bad arguments");
        std::mt19937 rng(seed);
        std::normal_distribution<double> dist(0.0, 1.0);
        for (int i = 0; i < n; i++) {
            // the reference's cx(dist(rng), dist(rng)) as g++ evaluates it: the
            // imaginary part's draw first
            const double im = dist(rng);
            const double re = dist(rng);
            out[2 * i] = re;
            out[2 * i + 1] = im;
        }
    });
}
int smg_dense_eigh(int n, const double* A, double* evals, double* evecs) {
    return guard([&] {
        SMG_REQUIRE(n >= 0 && (n == 0 || (A && evals && evecs)), "bad arguments");
        dense_eigh(n, reinterpret_cast<const c64*>(A), evals, reinterpret_cast<c64*>(evecs));
    });
}
static void trlm_out(const Eigenpairs& E, size_t n, double* evals, double* evecs, smg_trlm_info* info) {
    const int ne = (int)E.vals.size();
    if (evals) std::copy(E.vals.begin(), E.vals.end(), evals);
    if (evecs && ne) downc(evecs, E.vecs.p(), (size_t)ne * n);
    if (info) {
        info->iterations = E.iterations;
        info->restarts = E.restarts;
        info->converged = E.converged ? 1 : 0;
    }
}
int smg_dirac_trlm(smg_dirac* D, int op, int n_ev, int n_kr, int max_restarts, double tol, const double* start,
                   double* evals, double* evecs, smg_trlm_info* info) {
    return guard([&] {
        SMG_REQUIRE(D && start, "bad arguments");
        LinOp<double>* A;
        if (op == SMG_OP_DDAGD) A = D->op64.get();
        else {
            SMG_REQUIRE(op == SMG_OP_SCHUR && D->schur64, "unknown operator (SMG_OP_SCHUR needs an even L)");
            A = D->schur64.get();
        }
        const size_t n = A->n();
        SMG_REQUIRE(n_ev >= 1 && (size_t)n_ev < n, "need 1 <= n_ev < dimension");
        Field<double> s0 = up(start, n);
        Eigenpairs E = trlm(*A, n_ev, n_kr, max_restarts, tol, s0.p(), S());
        trlm_out(E, n, evals, evecs, info);
    });
}
int smg_coarse_trlm(smg_coarse* C, int n_ev, int n_kr, int max_restarts, double tol, const double* start,
                    double* evals, double* evecs, smg_trlm_info* info) {
    return guard([&] {
        SMG_REQUIRE(C && start, "bad arguments");
        const size_t n = C->C->n();
        SMG_REQUIRE(n_ev >= 1 && (size_t)n_ev < n, "need 1 <= n_ev < dimension");
        Field<double> s0 = up(start, n);
        CoarseLinOp<double, c64> op(*C->C);
        Eigenpairs E = trlm(op, n_ev, n_kr, max_restarts, tol, s0.p(), S());
        trlm_out(E, n, evals, evecs, info);
    });
}
int smg_coarse_setup_deflation(smg_coarse* C, int n_ev, int n_kr, int max_restarts, double tol, const double* start,
                               smg_trlm_info* info) {
    return guard([&] {
        SMG_REQUIRE(C && start, "bad arguments");
        const int dim = (int)C->C->n();
        if (dim <= 0 || n_ev <= 0) return;
        if (n_kr <= 0) n_kr = std::min(2 * n_ev + 10, dim);
        if (n_kr > dim) n_kr = dim;
        if (n_kr < n_ev + 2) n_kr = std::min(n_ev + 6, dim);
        SMG_REQUIRE(n_ev < dim, "need n_ev < coarse dimension");
        Field<double> s0 = up(start, (size_t)dim);
        CoarseLinOp<double, c64> op(*C->C);
        C->defl = trlm(op, n_ev, n_kr, max_restarts, tol, s0.p(), S());
        trlm_out(C->defl, (size_t)dim, nullptr, nullptr, info);
    });
}
int smg_coarse_deflation(smg_coarse* C, int* ndefl, double* evals, double* evecs) {
    return guard([&] {
        SMG_REQUIRE(C, "null handle");
        if (ndefl) *ndefl = (int)C->defl.vals.size();
        trlm_out(C->defl, C->C->n(), evals, evecs, nullptr);
    });
}
int smg_coarse_set_deflation(smg_coarse* C, int ndefl, const double* evals, const double* evecs) {
    return guard([&] {
        SMG_REQUIRE(C && ndefl >= 0 && (ndefl == 0 || (evals && evecs)), "bad arguments");
        const size_t n = C->C->n();
        Eigenpairs E;
        if (ndefl > 0) {
            E.vals.assign(evals, evals + ndefl);
            E.vecs = up(evecs, (size_t)ndefl * n);
            Context::get().sync();
        }
        E.converged = true;
        C->defl = std::move(E);
    });
}
int smg_cg_solve_deflated(smg_dirac* D, int op, smg_mg* mg, const double* b, int ndefl, const double* defl_vecs,
                          double* x, int maxit, double tol, smg_solve_info* info) {
    return guard([&] {
        SMG_REQUIRE(D && ndefl >= 0 && (ndefl == 0 || defl_vecs), "bad arguments");
        LinOp<double>* A;
        if (op == SMG_OP_DDAGD) A = D->op64.get();
        else {
            SMG_REQUIRE(op == SMG_OP_SCHUR && D->schur64, "unknown operator (SMG_OP_SCHUR needs an even L)");
            SMG_REQUIRE(mg == nullptr, "the multigrid preconditioner acts on D^dag D only");
            A = D->schur64.get();
        }
        const size_t n = A->n();
        Field<double> bb = up(b, n), xx(n), uu;
        if (ndefl) uu = up(defl_vecs, (size_t)ndefl * n);
        SolveResult r = cg_deflated(*A, mg ? mg->mg.get() : nullptr, ndefl ? uu.p() : nullptr, ndefl, bb.p(), xx.p(),
                                    maxit, tol, S());
        r.true_residual = true_residual(*A, bb.p(), xx.p(), S());
        fill_info(info, r);
        downc(x, xx.p(), n);
    });
}

int smg_mg_rebuild(smg_mg* mg) {
    return guard([&] {
        SMG_REQUIRE(mg, "null handle");
        mg->mg->rebuild(S());
    });
}
int smg_mg_refresh(smg_mg* mg, int null_iters) {
    return guard([&] {
        SMG_REQUIRE(mg && null_iters >= 0, "bad arguments");
        mg->mg->refresh(null_iters, S());
    });
}
int smg_mg_setup_sparse_coarse(smg_mg* mg, int n_defl, double cg_tol, int max_cg_iter, const double* start) {
    return guard([&] {
        SMG_REQUIRE(mg && start && n_defl >= 0 && max_cg_iter >= 0 && cg_tol > 0.0, "bad arguments");
        MG& m = *mg->mg;
        SMG_REQUIRE(m.cfg.n_levels == 2,
                    "setup_sparse_coarse wires the level-1 operator as the coarsest solve: two-level hierarchies");
        const int dim = (int)m.C[0]->n();
        // SparseCoarseOp::setup_deflation(n_defl, 0, 100, 1e-10) on the level-1 operator
        int n_kr = std::min(2 * n_defl + 10, dim);
        if (n_kr < n_defl + 2) n_kr = std::min(n_defl + 6, dim);
        std::unique_ptr<Eigenpairs> E(new Eigenpairs());
        if (n_defl > 0) {
            SMG_REQUIRE(n_defl < dim, "n_defl must be below the coarse dimension");
            Field<double> s0 = up(start, (size_t)dim);
            CoarseLinOp<double, c64> op(*m.C[0]);
            *E = trlm(op, n_defl, n_kr, 100, 1e-10, s0.p(), S());
        }
        m.sparse_defl = std::move(E);
        m.sparse_maxit = max_cg_iter;
        m.sparse_tol = cg_tol;
    });
}
static int next_pow2(int k) {
    int N = 1;
    while (N < k) N <<= 1;
    return N;
}
static void lobpcg_out(const LobpcgResult& r, size_t n, double* evals, double* X, int* iterations) {
    if (evals) std::copy(r.vals.begin(), r.vals.end(), evals);
    if (X) downc(X, r.vecs.p(), r.vals.size() * n);
    if (iterations) *iterations = r.iterations;
}
int smg_dirac_lobpcg(smg_dirac* D, smg_mg* mg, int k, const double* X0, int max_iter, double tol, double* evals,
                     double* X, int* iterations) {
    return guard([&] {
        SMG_REQUIRE(D && X0 && k >= 1 && k <= 32, "bad arguments (1 <= k <= 32)");
        const size_t n = 2 * (size_t)D->D->V;
        DdagDOpB<double> A(*D->D, next_pow2(k));
        Field<double> x0 = up(X0, (size_t)k * n);
        LobpcgResult r = lobpcg(A, mg ? mg->mg.get() : nullptr, k, x0.p(), max_iter, tol, S());
        lobpcg_out(r, n, evals, X, iterations);
    });
}
int smg_dirac_lobpcg_ex(smg_dirac* D, smg_mg* mg, int k, const double* X0, int max_iter, double tol, int exact_ax,
                        double* evals, double* X, int* iterations) {
    return guard([&] {
        SMG_REQUIRE(D && X0 && k >= 1 && k <= 32, "bad arguments (1 <= k <= 32)");
        const size_t n = 2 * (size_t)D->D->V;
        DdagDOpB<double> A(*D->D, next_pow2(k));
        Field<double> x0 = up(X0, (size_t)k * n);
        LobpcgResult r = lobpcg(A, mg ? mg->mg.get() : nullptr, k, x0.p(), max_iter, tol, S(), exact_ax != 0);
        lobpcg_out(r, n, evals, X, iterations);
    });
}
int smg_coarse_lobpcg(smg_coarse* C, int k, const double* X0, int max_iter, double tol, double* evals, double* X,
                      int* iterations) {
    return guard([&] {
        SMG_REQUIRE(C && X0 && k >= 1 && k <= 32, "bad arguments (1 <= k <= 32)");
        const size_t n = C->C->n();
        CoarseLinOpB<double, c64> A(*C->C, next_pow2(k));
        Field<double> x0 = up(X0, (size_t)k * n);
        LobpcgResult r = lobpcg(A, nullptr, k, x0.p(), max_iter, tol, S());
        lobpcg_out(r, n, evals, X, iterations);
    });
}
int smg_mg_update_coarsest_deflation(smg_mg* mg, int n_defl, int lobpcg_iters) {
    return guard([&] {
        SMG_REQUIRE(mg, "null handle");
        MG& m = *mg->mg;
        SMG_REQUIRE(!m.C.empty(), "the hierarchy has no coarse level");
        const size_t dim = m.dims.back();
        if (dim == 0 || n_defl <= 0) return;
        SMG_REQUIRE(n_defl <= 32 && (size_t)n_defl <= dim, "n_defl: 1..32 and at most the coarsest dimension");
        Field<double> x0((size_t)n_defl * dim);
        if ((int)m.defl_vals.size() >= n_defl) {   // warm start
            launch_copy<double>(x0.p(), m.defl_vecs.p(), (size_t)n_defl * dim, S());
        } else {   // cold start: random_vec(dim, rng) n_defl times from one mt19937(42)
            std::vector<double> h(2 * (size_t)n_defl * dim);
            SMG_REQUIRE(smg_random_vec((int)(n_defl * dim), 42, h.data()) == 0, "random_vec failed");
            SMG_HIP(hipMemcpyAsync(x0.p(), h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, S()));
            SMG_HIP(hipStreamSynchronize(S()));
        }
        CoarseLinOpB<double, c64> A(*m.C.back(), next_pow2(n_defl));
        LobpcgResult r = lobpcg(A, nullptr, n_defl, x0.p(), lobpcg_iters, 1e-10, S());
        m.defl_vals = r.vals;
        m.defl_vecs = std::move(r.vecs);
    });
}
int smg_mg_coarsest_deflation(smg_mg* mg, int* ndefl, double* evals, double* evecs) {
    return guard([&] {
        SMG_REQUIRE(mg, "null handle");
        const MG& m = *mg->mg;
        const int nd = (int)m.defl_vals.size();
        if (ndefl) *ndefl = nd;
        if (evals) std::copy(m.defl_vals.begin(), m.defl_vals.end(), evals);
        if (evecs && nd) downc(evecs, m.defl_vecs.p(), (size_t)nd * m.dims.back());
    });
}

// ------------------------------------------------------------------ caller-supplied operators
namespace {
// The operator of an smg_op: D^dag D of a Dirac handle, a coarse handle's operator, or
// the caller's host callback.  A callback is called with host arrays of n complex
// numbers; its input is copied from the device before the call and its output back
// after it, and the Krylov vectors, reductions and updates stay on the device.
struct CallbackOp : LinOp<double> {
    size_t nn;
    smg_op_fn fn;
    void* ctx;
    std::vector<c64> hin, hout;
    CallbackOp(size_t n_, smg_op_fn f, void* c) : nn(n_), fn(f), ctx(c), hin(n_), hout(n_) {}
    size_t n() const override { return nn; }
    void apply(c64* dst, const c64* src, Epi epi, const c64* b, const c64* aux, const Fin& fin,
               hipStream_t st) override {
        SMG_HIP(hipMemcpyAsync(hin.data(), src, nn * sizeof(c64), hipMemcpyDeviceToHost, st));
        SMG_HIP(hipStreamSynchronize(st));
        SMG_REQUIRE(fn(ctx, reinterpret_cast<const double*>(hin.data()), reinterpret_cast<double*>(hout.data())) == 0,
                    "the operator callback reported a failure");
        SMG_HIP(hipMemcpyAsync(dst, hout.data(), nn * sizeof(c64), hipMemcpyHostToDevice, st));
        SMG_HIP(hipStreamSynchronize(st));   // hout may be overwritten by the next call
        if (epi == EPI_RESID) launch_axpby<double>(dst, b, c64{1.0, 0.0}, c64{-1.0, 0.0}, nn, st);
        else if (epi == EPI_DOTS) launch_dots<double>(dst, aux, nn, fin, st);
    }
    using LinOp<double>::apply;
};
struct CallbackPrec : Preconditioner {
    size_t nn;
    smg_op_fn fn;
    void* ctx;
    std::vector<c64> hin, hout;
    CallbackPrec(size_t n_, smg_op_fn f, void* c) : nn(n_), fn(f), ctx(c), hin(n_), hout(n_) {}
    void precondition(const c64* b, c64* x, hipStream_t st) override {
        SMG_HIP(hipMemcpyAsync(hin.data(), b, nn * sizeof(c64), hipMemcpyDeviceToHost, st));
        SMG_HIP(hipStreamSynchronize(st));
        SMG_REQUIRE(fn(ctx, reinterpret_cast<const double*>(hin.data()), reinterpret_cast<double*>(hout.data())) == 0,
                    "the preconditioner callback reported a failure");
        SMG_HIP(hipMemcpyAsync(x, hout.data(), nn * sizeof(c64), hipMemcpyHostToDevice, st));
        SMG_HIP(hipStreamSynchronize(st));
    }
    using Preconditioner::precondition;
};
// the operator an smg_op describes (own holds what had to be built)
LinOp<double>& op_of(const smg_op* A, std::unique_ptr<LinOp<double>>& own) {
    SMG_REQUIRE(A, "null operator");
    const int set = (A->dirac != nullptr) + (A->coarse != nullptr) + (A->fn != nullptr);
    SMG_REQUIRE(set == 1, "an smg_op names exactly one of dirac, coarse, fn");
    if (A->dirac) {
        SMG_REQUIRE(A->n == 0 || A->n == A->dirac->op64->n(), "smg_op: n does not match the Dirac operator");
        return *A->dirac->op64;
    }
    if (A->coarse) {
        SMG_REQUIRE(A->n == 0 || A->n == A->coarse->C->n(), "smg_op: n does not match the coarse operator");
        own.reset(new CoarseLinOp<double, c64>(*A->coarse->C));
        return *own;
    }
    SMG_REQUIRE(A->n > 0, "smg_op: a callback operator needs its dimension n");
    own.reset(new CallbackOp(A->n, A->fn, A->ctx));
    return *own;
}
Preconditioner* prec_of(const smg_precond* M, size_t n, std::unique_ptr<Preconditioner>& own) {
    if (!M || (!M->mg && !M->fn)) return nullptr;   // the identity
    SMG_REQUIRE(!(M->mg && M->fn), "an smg_precond names at most one of mg, fn");
    if (M->mg) {
        SMG_REQUIRE(M->mg->mg->dims[0] == n, "smg_precond: the hierarchy's fine dimension does not match");
        return M->mg->mg.get();
    }
    own.reset(new CallbackPrec(n, M->fn, M->ctx));
    return own.get();
}
}  // namespace

int smg_cg_solve_op(const smg_op* A, const smg_precond* M, const double* b, const double* x0, double* x, int maxit,
                    double tol, smg_solve_info* info) {
    return guard([&] {
        SMG_REQUIRE(b && x, "null argument");
        std::unique_ptr<LinOp<double>> oa;
        std::unique_ptr<Preconditioner> om;
        LinOp<double>& op = op_of(A, oa);
        const size_t n = op.n();
        Preconditioner* pm = prec_of(M, n, om);
        Field<double> bb = up(b, n), xx(n), xi;
        if (x0) xi = up(x0, n);
        SolveResult r = cg_solve(op, bb.p(), xx.p(), x0 ? xi.p() : nullptr, pm, maxit, tol, S());
        r.true_residual = true_residual(op, bb.p(), xx.p(), S());
        fill_info(info, r);
        downc(x, xx.p(), n);
    });
}
int smg_gcr_solve_op(const smg_op* A, const smg_precond* M, const double* b, double* x, int maxit, double tol,
                     int trunc, int ref_projection, smg_solve_info* info) {
    return guard([&] {
        SMG_REQUIRE(b && x, "null argument");
        std::unique_ptr<LinOp<double>> oa;
        std::unique_ptr<Preconditioner> om;
        LinOp<double>& op = op_of(A, oa);
        const size_t n = op.n();
        Preconditioner* pm = prec_of(M, n, om);
        Field<double> bb = up(b, n), xx(n);
        SolveResult r = gcr_solve<double>(op, bb.p(), xx.p(), pm, maxit, tol, trunc, ref_projection != 0, S());
        r.true_residual = true_residual(op, bb.p(), xx.p(), S());
        fill_info(info, r);
        downc(x, xx.p(), n);
    });
}
int smg_fgmres_solve_op(const smg_op* A, const smg_precond* M, const double* b, double* x, int restart, int maxit,
                        double tol, int n_ritz, smg_solve_info* info, int* n_pairs, double* ritz_vals,
                        double* ritz_vecs) {
    return guard([&] {
        SMG_REQUIRE(b && x && restart >= 1 && n_ritz >= 0, "bad arguments");
        SMG_REQUIRE(n_ritz == 0 || (n_pairs && ritz_vals && ritz_vecs), "ritz output buffers required");
        std::unique_ptr<LinOp<double>> oa;
        std::unique_ptr<Preconditioner> om;
        LinOp<double>& op = op_of(A, oa);
        const size_t n = op.n();
        Preconditioner* pm = prec_of(M, n, om);
        Field<double> bb = up(b, n), xx(n);
        std::unique_ptr<RitzHarvest> rh;
        if (n_ritz > 0) rh.reset(new RitzHarvest(n_ritz, restart, n));
        SolveResult r = fgmres_solve(op, bb.p(), xx.p(), pm, restart, maxit, tol, S(), rh.get());
        fill_info(info, r);
        downc(x, xx.p(), n);
        if (n_pairs) *n_pairs = rh ? rh->count() : 0;
        if (rh && rh->count() > 0) {
            std::copy(rh->vals.begin(), rh->vals.end(), ritz_vals);
            downc(ritz_vecs, rh->vecs.p(), (size_t)rh->count() * n);
        }
    });
}
int smg_smooth_op(const smg_op* A, int kind, double* x, const double* b, int niter, double omega) {
    return guard([&] {
        SMG_REQUIRE(x && b, "null argument");
        std::unique_ptr<LinOp<double>> oa;
        smooth_op(op_of(A, oa), kind, x, b, niter, omega);
    });
}
int smg_lambda_max_op(const smg_op* A, const double* v0, int niter, double* out) {
    return guard([&] {
        SMG_REQUIRE(v0 && out, "null argument");
        std::unique_ptr<LinOp<double>> oa;
        LinOp<double>& op = op_of(A, oa);
        Field<double> v = up(v0, op.n()), Av(op.n());
        *out = lambda_max<double>(op, v.p(), Av.p(), niter, S());
    });
}
int smg_coarse_create_galerkin_op(smg_coarse** out, const smg_op* A, smg_prolongator* P) {
    return guard([&] {
        SMG_REQUIRE(out && P, "null argument");
        std::unique_ptr<LinOp<double>> oa;
        LinOp<double>& op = op_of(A, oa);
        SMG_REQUIRE(op.n() == P->P->fine_n(), "the operator's dimension does not match the prolongator's fine side");
        auto* h = new smg_coarse_s();
        try {
            h->C.reset(new CoarseOperator(P->P->A.nbx, P->P->A.nby, P->P->A.k));
            h->C->build(op, *P->P, S());
        } catch (...) { delete h; throw; }
        *out = h;
    });
}

// ------------------------------------------------------------------ launch-shape autotuner
int smg_tune_mg(smg_mg* mg, int nrhs, int* entries) {
    return guard([&] {
        SMG_REQUIRE(mg, "null handle");
        const int n = tune_mg(*mg->mg, nrhs, S());
        if (entries) *entries = n;
    });
}
int smg_tune_save(const char* path) {
    return guard([&] {
        SMG_REQUIRE(path, "null path");
        tune_save(path);
    });
}
int smg_tune_load(const char* path) {
    return guard([&] {
        SMG_REQUIRE(path, "null path");
        tune_load(path);
    });
}
int smg_tune_clear(void) {
    return guard([&] { tune_clear(); });
}
int smg_tune_get(int family, const int* shape, int* choice) {
    return guard([&] {
        SMG_REQUIRE(shape && choice, "null argument");
        *choice = tune_lookup(TuneKey{family, shape[0], shape[1], shape[2], shape[3], shape[4]});
    });
}
int smg_tune_set(int family, const int* shape, int choice) {
    return guard([&] {
        SMG_REQUIRE(shape, "null argument");
        tune_store(TuneKey{family, shape[0], shape[1], shape[2], shape[3], shape[4]}, choice);
    });
}

// ------------------------------------------------------------------ even-odd multigrid
int smg_eo_mg_setup(smg_eo_mg** out, smg_dirac* D, int bx, int by, int k, int pre_smooth, int post_smooth,
                    int null_iters, const double* starts, const double* lambda_start) {
    return guard([&] {
        SMG_REQUIRE(out && D && starts && lambda_start, "null argument");
        auto* h = new smg_eo_mg_s();
        try {
            h->mg.reset(new EOMG(*D->D, bx, by, k, pre_smooth, post_smooth, null_iters,
                                 reinterpret_cast<const c64*>(starts), reinterpret_cast<const c64*>(lambda_start)));
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}
int smg_eo_mg_destroy(smg_eo_mg* M) { return guard([&] { delete M; }); }
int smg_eo_mg_info(smg_eo_mg* M, double* omega, double* lambda_max, int* coarse_dim) {
    return guard([&] {
        SMG_REQUIRE(M, "null handle");
        if (omega) *omega = M->mg->omega;
        if (lambda_max) *lambda_max = M->mg->lmax;
        if (coarse_dim) *coarse_dim = (int)M->mg->dim;
    });
}
int smg_eo_mg_coarse_matrix(smg_eo_mg* M, double* out) {
    return guard([&] {
        SMG_REQUIRE(M && out, "null argument");
        downc(out, M->mg->mat.p(), M->mg->dim * M->mg->dim);
    });
}
int smg_eo_mg_precondition(smg_eo_mg* M, const smg_op* op, const double* b, double* x) {
    return guard([&] {
        SMG_REQUIRE(M && b && x, "null argument");
        std::unique_ptr<LinOp<double>> own;
        LinOp<double>& A = op ? op_of(op, own) : M->mg->A;
        const size_t n = M->mg->A.n();
        Field<double> bb = up(b, n), xx(n);
        M->mg->precondition(A, bb.p(), xx.p(), S());
        downc(x, xx.p(), n);
    });
}
int smg_eo_mg_precondition_dev(smg_eo_mg* M, const void* b, void* x, void* stream) {
    return guard([&] {
        SMG_REQUIRE(M && b && x, "null argument");
        StreamJoin j(stream);
        M->mg->precondition(M->mg->A, static_cast<const c64*>(b), static_cast<c64*>(x), S());
    });
}
