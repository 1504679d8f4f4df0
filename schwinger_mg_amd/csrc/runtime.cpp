// runtime.cpp -- host-side sequencing of the gfx950 kernels.
#include "runtime.h"

#include "eigen.h"
#include "runtime_b.h"

#include <cmath>
#include <cstring>
#include <algorithm>
#include <deque>
#include <map>
#include <mutex>
#include <random>

namespace smg {

// ------------------------------------------------------------------ memory
// Caching allocator: freed blocks are kept per rounded size and handed out again,
// so solver temporaries cost no hipMalloc/hipFree (and no device synchronisation)
// after their first use.  All library work runs on one stream, so a block
// returned to the cache is only reused by work ordered after its last use.
namespace {
std::mutex g_pool_mu;
std::multimap<size_t, void*>& pool() {
    static auto* m = new std::multimap<size_t, void*>();
    return *m;
}
size_t round_bytes(size_t b) {
    if (b <= 256) return 256;
    size_t r = 256;
    while (r < b) r <<= 1;
    if (r - b > (b >> 2) && b > (size_t(1) << 20)) r = (b + (size_t(1) << 20) - 1) & ~((size_t(1) << 20) - 1);
    return r;
}
}  // namespace
void DevMem::alloc(size_t b) {
    release();
    const size_t r = round_bytes(b);
    {
        std::lock_guard<std::mutex> lk(g_pool_mu);
        auto it = pool().find(r);
        if (it != pool().end()) {
            p = it->second;
            pool().erase(it);
            bytes = r;
            return;
        }
    }
    hipError_t e = hipMalloc(&p, r);
    if (e != hipSuccess) {
        // give the cache back to the runtime and retry once
        (void)hipGetLastError();
        {
            std::lock_guard<std::mutex> lk(g_pool_mu);
            for (auto& kv : pool()) (void)hipFree(kv.second);
            pool().clear();
        }
        SMG_HIP(hipMalloc(&p, r));
    }
    bytes = r;
}
void DevMem::release() {
    if (p) {
        std::lock_guard<std::mutex> lk(g_pool_mu);
        pool().emplace(bytes, p);
    }
    p = nullptr;
    bytes = 0;
}

Context& Context::get() {
    static Context* ctx = nullptr;
    if (!ctx) {
        ctx = new Context();
        SMG_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        ctx->part.alloc(size_t(1) << 24);  // 2M doubles of partial sums
        ctx->counter.alloc(4096);   // 1024 completion counters
        SMG_HIP(hipMemset(ctx->counter.p, 0, 4096));
        void* h = nullptr;
        SMG_HIP(hipHostMalloc(&h, 4096, hipHostMallocDefault));
        ctx->hostscal.p = h;
        ctx->hostscal.bytes = 4096;
    }
    return *ctx;
}
Context::~Context() {}

void Context::read(const double* dev, double* host, size_t count) {
    double* hs = hostscal.as<double>();
    SMG_REQUIRE(count <= 512, "read: too many scalars");
    SMG_HIP(hipMemcpyAsync(hs, dev, count * sizeof(double), hipMemcpyDeviceToHost, stream));
    SMG_HIP(hipStreamSynchronize(stream));
    std::memcpy(host, hs, count * sizeof(double));
}
void Context::sync() { SMG_HIP(hipStreamSynchronize(stream)); }

// A small device scalar block owned by one object.
struct Scalars {
    DevMem m;
    explicit Scalars(size_t n = 64) : m(n * sizeof(double)) {
        SMG_HIP(hipMemsetAsync(m.p, 0, n * sizeof(double), Context::get().stream));
    }
    double* p(size_t i = 0) const { return m.as<double>() + i; }
};

double dev_norm(const c64* v, size_t n, hipStream_t st) {
    Context& cx = Context::get();
    Scalars s(4);
    launch_norm2<double>(v, n, cx.fin(FIN_STORE, s.p()), st);
    double h;
    cx.read(s.p(), &h, 1);
    return std::sqrt(h);
}
template <typename T>
c64 dev_dot(const cplx<T>* a, const cplx<T>* b, size_t n, hipStream_t st) {
    Context& cx = Context::get();
    Scalars s(4);
    launch_dot<T>(a, b, n, cx.fin(FIN_STORE, s.p()), st);
    double h[2];
    cx.read(s.p(), h, 2);
    return c64{h[0], h[1]};
}
template c64 dev_dot<double>(const c64*, const c64*, size_t, hipStream_t);
template c64 dev_dot<float>(const c32*, const c32*, size_t, hipStream_t);

template <typename T>
static double dev_norm_t(const cplx<T>* v, size_t n, hipStream_t st) {
    Context& cx = Context::get();
    Scalars s(4);
    launch_norm2<T>(v, n, cx.fin(FIN_STORE, s.p()), st);
    double h;
    cx.read(s.p(), &h, 1);
    return std::sqrt(h);
}

// device copy as a kernel (graph-capturable)
template <typename T>
static void dcopy(cplx<T>* d, const cplx<T>* s, size_t n, hipStream_t st) {
    launch_copy<T>(d, s, n, st);
}

// ------------------------------------------------------------------ Dirac
Dirac::Dirac(int L_, const c64* U_host, double m, double r_, double csw_, double mu_)
    : L(L_), V(L_ * L_), mass(m), r(r_), csw(csw_), mu_t(mu_), U64(2 * (size_t)L_ * L_), U32(2 * (size_t)L_ * L_) {
    SMG_REQUIRE(L >= 2, "L must be >= 2");
    F01.alloc(V * sizeof(double));
    if (csw != 0.0) {
        C64.alloc(V * sizeof(double));
        C32.alloc(V * sizeof(float));
    }
    set_gauge(U_host, false);
}

void Dirac::set_gauge(const c64* U, bool device_ptr) {
    hipStream_t st = Context::get().stream;
    SMG_HIP(hipMemcpyAsync(U64.p(), U, 2 * (size_t)V * sizeof(c64),
                           device_ptr ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st));
    refresh_derived(st);
    if (!device_ptr) SMG_HIP(hipStreamSynchronize(st));   // the host array may go away
}
void Dirac::refresh_derived(hipStream_t st) {
    launch_convert<float, double>(U32.p(), U64.p(), 2 * (size_t)V, st);
    launch_clover(L, U64.p(), F01.as<double>(), st);
    // C = 0.5 * csw * F01 (src/dirac.h:102), kept in FP64 and FP32
    if (csw != 0.0) launch_clover_scale(F01.as<double>(), csw, C64.as<double>(), C32.as<float>(), V, st);
}
void Dirac::update_links(const double* pi, double scale, hipStream_t st) {
    launch_link_update(U64.p(), U32.p(), pi, scale, 2 * (size_t)V, st);
    launch_clover(L, U64.p(), F01.as<double>(), st);
    if (csw != 0.0) launch_clover_scale(F01.as<double>(), csw, C64.as<double>(), C32.as<float>(), V, st);
}

DiracView<double> Dirac::v64() const {
    return DiracView<double>{U64.p(), csw != 0.0 ? C64.as<double>() : nullptr, 2.0 * r + mass, r, L};
}
DiracView<float> Dirac::v32() const {
    return DiracView<float>{U32.p(), csw != 0.0 ? C32.as<float>() : nullptr, (float)(2.0 * r + mass), (float)r, L};
}
template <>
DiracView<double> Dirac::view<double>() const { return v64(); }
template <>
DiracView<float> Dirac::view<float>() const { return v32(); }

template <typename T>
void DdagDOp<T>::apply(cplx<T>* dst, const cplx<T>* src, Epi epi, const cplx<T>* b, const cplx<T>* aux,
                       const Fin& fin, hipStream_t st) {
    const DiracView<T> v = D.view<T>();
    const T mu = (T)D.mu_t;
    launch_dslash<T>(v, tmp.p(), src, mu, 0, 0, EPI_NONE, nullptr, nullptr, Fin(), 1, 0, st);
    launch_dslash<T>(v, dst, tmp.p(), -mu, 1, 1, epi, b, aux, fin, 1, 0, st);
}

template <typename T>
void SchurOp<T>::apply_m(cplx<T>* dst, const cplx<T>* src, int dag, hipStream_t st) {
    const DiracView<T> v = D.view<T>();
    const T mu = dag ? (T)-D.mu_t : (T)D.mu_t;
    launch_eo_hop<T>(v, te.p(), src, 0, 1, mu, dag, 0, nullptr, EPI_NONE, nullptr, nullptr, Fin(), st);
    launch_eo_hop<T>(v, dst, te.p(), 1, 2, mu, dag, dag, src, EPI_NONE, nullptr, nullptr, Fin(), st);
}
template <typename T>
void SchurOp<T>::apply(cplx<T>* dst, const cplx<T>* src, Epi epi, const cplx<T>* b, const cplx<T>* aux,
                       const Fin& fin, hipStream_t st) {
    const DiracView<T> v = D.view<T>();
    apply_m(to.p(), src, 0, st);
    const T mu = (T)-D.mu_t;
    launch_eo_hop<T>(v, te.p(), to.p(), 0, 1, mu, 1, 0, nullptr, EPI_NONE, nullptr, nullptr, Fin(), st);
    launch_eo_hop<T>(v, dst, te.p(), 1, 2, mu, 1, 1, to.p(), epi, b, aux, fin, st);
}

template struct DdagDOp<double>;
template struct DdagDOp<float>;
template struct SchurOp<double>;
template struct SchurOp<float>;

// ------------------------------------------------------------------ aggregation
Aggregation::Aggregation(int nx, int ny, int nd, int bx, int by, int k, int chiral) {
    SMG_REQUIRE(bx >= 1 && by >= 1 && nx % bx == 0 && ny % by == 0, "aggregate shape must tile the lattice");
    SMG_REQUIRE(k >= 1, "k must be >= 1");
    SMG_REQUIRE(!chiral || (k % 2 == 0 && nd % 2 == 0), "chiral aggregation needs even k and an even dof count");
    A = AggView{nx, ny, nd, bx, by, k, nx / bx, ny / by, chiral ? 1 : 0};
    const size_t nbv = (size_t)A.nbx * A.nby * k * bx * by * nd;
    bv64 = Field<double>(nbv);
    bv32 = Field<float>(nbv);
}
void Aggregation::build(const c64* null_vecs, hipStream_t st) {
    const int nv = nvec();
    for (int kv = 0; kv < nv; kv++) {
        const c64* v = null_vecs + (size_t)kv * fine_n();
        if (A.chiral) {
            launch_block_gather(A, bv64.p(), v, kv, 0, st);        // coarse dof kv: gamma5 = +1
            launch_block_gather(A, bv64.p(), v, nv + kv, 1, st);   // coarse dof nv + kv: gamma5 = -1
        } else {
            launch_block_gather(A, bv64.p(), v, kv, -1, st);
        }
    }
    launch_block_orthonormalize(A, bv64.p(), st);
    launch_convert<float, double>(bv32.p(), bv64.p(), bv64.n, st);
}
void Aggregation::set_block_vecs(const c64* bv_dev, hipStream_t st) {
    dcopy(bv64.p(), bv_dev, bv64.n, st);
    launch_convert<float, double>(bv32.p(), bv64.p(), bv64.n, st);
}
template <>
const c64* Aggregation::bv<double>() const { return bv64.p(); }
template <>
const c32* Aggregation::bv<float>() const { return bv32.p(); }

// ------------------------------------------------------------------ coarse operator
int color_period(int nb) {
    if (nb < 3) return nb;
    for (int c = 3; c <= nb; c++)
        if (nb % c == 0) return c;
    return nb;
}

CoarseOperator::CoarseOperator(int nx_, int ny_, int k_)
    : nx(nx_), ny(ny_), k(k_), l64((size_t)nx_ * ny_ * 9 * k_ * k_), l32((size_t)nx_ * ny_ * 9 * k_ * k_) {
    l16.alloc(nlinks() * sizeof(ch16));
}

void CoarseOperator::build(LinOp<double>& A, const Aggregation& P, hipStream_t st) {
    SMG_REQUIRE(P.A.nbx == nx && P.A.nby == ny && P.A.k == k, "coarse operator / aggregation mismatch");
    hermitian = true;
    launch_zero<double>(l64.p(), l64.n, st);
    Field<double> e(n()), f(P.fine_n()), Af(P.fine_n()), col(n());
    const int cx = color_period(nx), cy = color_period(ny);
    for (int cls = 0; cls < cx * cy; cls++)
        for (int kj = 0; kj < k; kj++) {
            launch_probe_vector(nx, ny, k, cx, cy, cls, kj, e.p(), st);
            launch_prolong<double>(P.A, P.bv64.p(), e.p(), f.p(), 0, st);
            A.apply(Af.p(), f.p(), st);
            launch_restrict<double>(P.A, P.bv64.p(), Af.p(), col.p(), st);
            launch_probe_extract(nx, ny, k, cx, cy, cls, kj, col.p(), l64.p(), st);
        }
    refresh_low_precision(st);
    SMG_HIP(hipStreamSynchronize(st));
}
void CoarseOperator::set_links(const c64* dev, hipStream_t st) {
    hermitian = false;   // links from the caller: no symmetry is assumed
    dcopy(l64.p(), dev, nlinks(), st);
    refresh_low_precision(st);
}
void CoarseOperator::refresh_low_precision(hipStream_t st) {
    launch_convert<float, double>(l32.p(), l64.p(), nlinks(), st);
    launch_convert_half(l16.as<ch16>(), l64.p(), nlinks(), st);
    // P^dag A P of a Hermitian A is Hermitian up to roundoff; the FP16 copy is made so
    // exactly, and its matrix-core applies then read 5 of the 9 blocks of a site
    herm16 = (hermitian && nx >= 3 && ny >= 3) ? 1 : 0;
    if (herm16) launch_half_hermitian(l16.as<ch16>(), nx, ny, k, st);
}

template <typename T, typename TL>
void CoarseLinOp<T, TL>::apply(cplx<T>* dst, const cplx<T>* src, Epi epi, const cplx<T>* b, const cplx<T>* aux,
                               const Fin& fin, hipStream_t st) {
    const TL* link;
    if constexpr (std::is_same<TL, c64>::value) link = C.l64.p();
    else if constexpr (std::is_same<TL, c32>::value) link = C.l32.p();
    else link = C.l16.template as<ch16>();
    CoarseView<TL> v{link, C.nx, C.ny, C.k};
    launch_coarse_apply<T, TL>(v, dst, src, epi, b, aux, fin, st);
}
template struct CoarseLinOp<double, c64>;
template struct CoarseLinOp<float, c32>;
template struct CoarseLinOp<float, ch16>;

// ------------------------------------------------------------------ smoothers
template <typename T>
void mr_smooth(LinOp<T>& A, cplx<T>* x, const cplx<T>* b, cplx<T>* r, cplx<T>* Ar, int niter, bool zero_init,
               double* sc, hipStream_t st) {
    const size_t n = A.n();
    const Fin fin = Context::get().fin(FIN_MR, sc);
    if (zero_init) {
        launch_zero<T>(x, n, st);
        dcopy(r, b, n, st);
    } else {
        A.apply(r, x, EPI_RESID, b, nullptr, Fin(), st);
    }
    for (int it = 0; it < niter; it++) {
        // Ar = A r with {<Ar,r>, |Ar|^2} -> omega = <Ar,r>/|Ar|^2; x += omega r; r -= omega Ar
        A.apply(Ar, r, EPI_DOTS, nullptr, r, fin, st);
        launch_mr_update<T>(x, r, Ar, sc, n, st);
    }
}

template <typename T>
void richardson_smooth(LinOp<T>& A, cplx<T>* x, const cplx<T>* b, cplx<T>* r, int niter, double omega,
                       bool zero_init, hipStream_t st) {
    const size_t n = A.n();
    if (zero_init) launch_zero<T>(x, n, st);
    for (int it = 0; it < niter; it++) {
        if (it == 0 && zero_init) dcopy(r, b, n, st);
        else A.apply(r, x, EPI_RESID, b, nullptr, Fin(), st);
        launch_axpby<T>(x, r, c64{omega, 0.0}, c64{1.0, 0.0}, n, st);
    }
    if (niter == 0 && zero_init) dcopy(r, b, n, st);
    else A.apply(r, x, EPI_RESID, b, nullptr, Fin(), st);
}

template <typename T>
double lambda_max(LinOp<T>& A, cplx<T>* v, cplx<T>* Av, int niter, hipStream_t st) {
    const size_t n = A.n();
    double vn = dev_norm_t(v, n, st);
    launch_axpby<T>(v, v, c64{0.0, 0.0}, c64{1.0 / vn, 0.0}, n, st);
    double lambda = 0.0;
    for (int i = 0; i < niter; i++) {
        A.apply(Av, v, st);
        lambda = dev_norm_t(Av, n, st);
        if (lambda < 1e-30) break;
        launch_axpby<T>(v, Av, c64{1.0 / lambda, 0.0}, c64{0.0, 0.0}, n, st);
    }
    return lambda;
}

template <typename T>
int coarse_cg(LinOp<T>& A, const cplx<T>* b, cplx<T>* x, cplx<T>* r, cplx<T>* p, cplx<T>* Ap, int maxit, double tol,
              double* sc, bool check, hipStream_t st) {
    const size_t n = A.n();
    Context& cx = Context::get();
    double* alpha = sc;
    double* beta = sc + 2;
    double* aux = sc + 16;
    int* flag = reinterpret_cast<int*>(aux + 8);
    launch_zero<T>(x, n, st);
    dcopy(r, b, n, st);
    dcopy(p, b, n, st);
    launch_set_scalars(aux + 5, 1, tol * tol, 0, 0, 0, flag, st);
    launch_norm2<T>(r, n, cx.fin(FIN_CG_INIT, nullptr, aux), st);
    const Fin fa = cx.fin(FIN_CG_ALPHA, alpha, aux), fb = cx.fin(FIN_CG_BETA, beta, aux);
    const int chunk = 8;
    for (int it = 0; it < maxit; it++) {
        A.apply(Ap, p, EPI_DOTS, nullptr, p, fa, st);
        launch_cg_update<T>(x, r, p, Ap, alpha, flag, n, fb, st);
        launch_xpay_dev<T>(p, r, beta, flag, n, st);
        if (check && (it + 1) % chunk == 0 && it + 1 < maxit) {
            double h[9];
            cx.read(aux, h, 9);
            int f;
            std::memcpy(&f, &h[8], sizeof(int));
            if (f) break;
        }
    }
    if (!check) return -1;
    double h[9];
    cx.read(aux, h, 9);
    return (int)h[3];
}

template void mr_smooth<double>(LinOp<double>&, c64*, const c64*, c64*, c64*, int, bool, double*, hipStream_t);
template void mr_smooth<float>(LinOp<float>&, c32*, const c32*, c32*, c32*, int, bool, double*, hipStream_t);
template void richardson_smooth<double>(LinOp<double>&, c64*, const c64*, c64*, int, double, bool, hipStream_t);
template void richardson_smooth<float>(LinOp<float>&, c32*, const c32*, c32*, int, double, bool, hipStream_t);
template double lambda_max<double>(LinOp<double>&, c64*, c64*, int, hipStream_t);
template double lambda_max<float>(LinOp<float>&, c32*, c32*, int, hipStream_t);
template int coarse_cg<double>(LinOp<double>&, const c64*, c64*, c64*, c64*, c64*, int, double, double*, bool,
                               hipStream_t);
template int coarse_cg<float>(LinOp<float>&, const c32*, c32*, c32*, c32*, c32*, int, double, double*, bool,
                              hipStream_t);

// ------------------------------------------------------------------ multigrid cycle
template <typename T>
struct Cycle {
    MG& mg;
    std::vector<LinOp<T>*> ops;
    std::vector<std::unique_ptr<LinOp<T>>> own;
    std::vector<Field<T>> xs, bs, rs, Ars, ts, e2s, ps;
    std::vector<std::unique_ptr<Scalars>> sc;
    int nlev;
    Cycle(MG& m, LinOp<T>* fine) : mg(m), nlev(m.cfg.n_levels) {
        ops.resize(nlev);
        ops[0] = fine;
        for (int l = 1; l < nlev; l++) {
            if constexpr (std::is_same<T, double>::value)
                own.emplace_back(new CoarseLinOp<double, c64>(*m.C[l - 1]));
            else if (m.cfg.coarse_half)
                own.emplace_back(new CoarseLinOp<float, ch16>(*m.C[l - 1]));
            else
                own.emplace_back(new CoarseLinOp<float, c32>(*m.C[l - 1]));
            ops[l] = own.back().get();
        }
        for (int l = 0; l < nlev; l++) {
            const size_t n = m.dims[l];
            xs.emplace_back(n);
            bs.emplace_back(n);
            rs.emplace_back(n);
            Ars.emplace_back(n);
            ts.emplace_back(n);
            e2s.emplace_back(n);
            ps.emplace_back(l == nlev - 1 ? n : 1);
            sc.emplace_back(new Scalars(64));
        }
    }
    // src/multigrid.cpp:43-102
    void run(int lev, const cplx<T>* b, cplx<T>* x, hipStream_t st) {
        LinOp<T>& A = *ops[lev];
        const int last = nlev - 1;
        if (lev == last) {
            if constexpr (std::is_same<T, double>::value) {
                if (mg.sparse_defl) {   // SparseCoarseOp::solve with its deflation space
                    deflated_coarse_cg(A, *mg.sparse_defl, b, x, mg.sparse_maxit, mg.sparse_tol, st);
                    return;
                }
            }
            if (mg.cfg.coarse_solver == 1) {
                const cplx<T>* inv;
                if constexpr (std::is_same<T, double>::value) inv = mg.inv64.p();
                else inv = mg.inv32.p();
                launch_dense_matvec<T>(inv, b, x, mg.dims[lev], st);
            } else {
                coarse_cg<T>(A, b, x, rs[lev].p(), ps[lev].p(), Ars[lev].p(), mg.cfg.coarse_maxit,
                             mg.cfg.coarse_tol, sc[lev]->p(), false, st);
            }
            return;
        }
        const double om = mg.omega[lev];
        if (om > 0) richardson_smooth<T>(A, x, b, rs[lev].p(), mg.cfg.pre, om, true, st);
        else mr_smooth<T>(A, x, b, rs[lev].p(), Ars[lev].p(), mg.cfg.pre, true, sc[lev]->p(), st);
        // r = b - A x is carried by the smoother
        const Aggregation& P = *mg.P[lev];
        launch_restrict<T>(P.A, P.bv<T>(), rs[lev].p(), bs[lev + 1].p(), st);
        run(lev + 1, bs[lev + 1].p(), xs[lev + 1].p(), st);
        if (mg.cfg.w_cycle && lev + 1 < last) {
            ops[lev + 1]->apply(ts[lev + 1].p(), xs[lev + 1].p(), EPI_RESID, bs[lev + 1].p(), nullptr, Fin(), st);
            run(lev + 1, ts[lev + 1].p(), e2s[lev + 1].p(), st);
            launch_axpby<T>(xs[lev + 1].p(), e2s[lev + 1].p(), c64{1.0, 0.0}, c64{1.0, 0.0}, mg.dims[lev + 1], st);
        }
        launch_prolong<T>(P.A, P.bv<T>(), xs[lev + 1].p(), x, 1, st);
        if (om > 0) richardson_smooth<T>(A, x, b, rs[lev].p(), mg.cfg.post, om, false, st);
        else mr_smooth<T>(A, x, b, rs[lev].p(), Ars[lev].p(), mg.cfg.post, false, sc[lev]->p(), st);
    }
};

template <>
Cycle<double>& MG::cycle<double>() { return *cyc64; }
template <>
Cycle<float>& MG::cycle<float>() { return *cyc32; }

LinOp<double>& MG::level_op64(int l) { return *cyc64->ops[l]; }

// src/multigrid.cpp:545-587: inverse iteration with 5 MR sweeps per outer step.
// With T = float the MR sweeps run in FP32 on the FP32 operator (the vectors, their
// orthogonalisation and normalisation stay FP64): the right-hand side is rounded to
// FP32 before the sweeps and the iterate widened after them.
template <typename T>
static void null_space(LinOp<T>& A, int k, int outer, int inner, c64* vecs, hipStream_t st) {
    const size_t n = A.n();
    Field<double> x(n), rhs(n);
    const size_t nlow = std::is_same<T, double>::value ? 0 : n;   // FP32 copies of x and the rhs
    Field<T> xs(nlow), r(n), Ar(n), rhss(nlow);
    Scalars sc(64);
    for (int i = 0; i < k; i++) {
        c64* v = vecs + (size_t)i * n;
        double nv = dev_norm(v, n, st);
        launch_axpby<double>(v, v, c64{0, 0}, c64{1.0 / nv, 0}, n, st);
        for (int it = 0; it < outer; it++) {
            if constexpr (std::is_same<T, double>::value) {
                dcopy(rhs.p(), v, n, st);
                mr_smooth<double>(A, x.p(), rhs.p(), r.p(), Ar.p(), inner, true, sc.p(), st);
            } else {
                launch_convert<T, double>(rhss.p(), v, n, st);
                mr_smooth<T>(A, xs.p(), rhss.p(), r.p(), Ar.p(), inner, true, sc.p(), st);
                launch_convert<double, T>(x.p(), xs.p(), n, st);
            }
            for (int j = 0; j < i; j++) {
                c64 pr = dev_dot<double>(vecs + (size_t)j * n, x.p(), n, st);
                launch_axpby<double>(x.p(), vecs + (size_t)j * n, c64{-pr.re, -pr.im}, c64{1, 0}, n, st);
            }
            nv = dev_norm(x.p(), n, st);
            if (nv > 1e-14) launch_axpby<double>(v, x.p(), c64{1.0 / nv, 0}, c64{0, 0}, n, st);
            else break;
        }
    }
    for (int i = 0; i < k; i++) {
        c64* v = vecs + (size_t)i * n;
        for (int j = 0; j < i; j++) {
            c64 pr = dev_dot<double>(vecs + (size_t)j * n, v, n, st);
            launch_axpby<double>(v, vecs + (size_t)j * n, c64{-pr.re, -pr.im}, c64{1, 0}, n, st);
        }
        double nv = dev_norm(v, n, st);
        if (nv > 1e-14) launch_axpby<double>(v, v, c64{0, 0}, c64{1.0 / nv, 0}, n, st);
    }
}

// level l's near-null vectors by inverse iteration on A (level l's FP64 operator),
// the MR sweeps in FP64 or, with null_prec = 1, in FP32 on the same operator's FP32
// form (the fine FP32 D^dag D, or the FP32 links of the coarse operator of level l - 1)
void MG::near_null(size_t l, LinOp<double>& A, int nv, int iters, c64* vecs, hipStream_t st) {
    if (cfg.null_prec == 0) {
        null_space<double>(A, nv, iters, cfg.null_inner, vecs, st);
    } else if (l == 0) {
        null_space<float>(*fine32, nv, iters, cfg.null_inner, vecs, st);
    } else {
        CoarseLinOp<float, c32> A32(*C[l - 1]);
        null_space<float>(A32, nv, iters, cfg.null_inner, vecs, st);
    }
}

// the half-lattice's row length, after checking the block shape
static int eo_half_extent(const Dirac& d, int bx, int by) {
    SMG_REQUIRE(d.L % 2 == 0 && bx >= 2 && bx % 2 == 0 && by >= 1 && d.L % bx == 0 && d.L % by == 0,
                "even-odd multigrid: even bx, and bx, by dividing an even L");
    return d.L / 2;
}

// src/eo_multigrid.h:229-254 (EvenOddMG::build)
EOMG::EOMG(const Dirac& d, int bx_, int by_, int k_, int pre_, int post_, int null_iters, const c64* starts,
           const c64* lam_start)
    : D(d), bx(bx_), by(by_), k(k_), pre(pre_), post(post_), A(d),
      P(eo_half_extent(d, bx_, by_), d.L, 2, bx_ / 2, by_, k_) {
    SMG_REQUIRE(k >= 1 && pre >= 0 && post >= 0 && null_iters >= 0, "even-odd multigrid: bad parameters");
    hipStream_t st = Context::get().stream;
    const size_t n = A.n();
    dim = P.coarse_n();
    SMG_REQUIRE(dim <= 8192, "even-odd multigrid: the dense coarse operator needs <= 8192 unknowns");
    Field<double> vecs((size_t)k * n);
    SMG_HIP(hipMemcpyAsync(vecs.p(), starts, (size_t)k * n * sizeof(c64), hipMemcpyHostToDevice, st));
    null_space<double>(A, k, null_iters, 5, vecs.p(), st);   // compute_near_null_space_generic, 5 MR steps
    P.build(vecs.p(), st);
    // EvenOddCoarseOp::build (src/eo_multigrid.h:164-176): column j = P^dag A P e_j
    mat = Field<double>(dim * dim);
    Field<double> e(dim), f(n), Af(n), col(dim);
    static const c64 one{1.0, 0.0};
    for (size_t j = 0; j < dim; j++) {
        launch_zero<double>(e.p(), dim, st);
        SMG_HIP(hipMemcpyAsync(e.p(j), &one, sizeof(c64), hipMemcpyHostToDevice, st));
        launch_prolong<double>(P.A, P.bv64.p(), e.p(), f.p(), 0, st);
        A.apply(Af.p(), f.p(), st);
        launch_restrict<double>(P.A, P.bv64.p(), Af.p(), col.p(), st);
        SMG_HIP(hipMemcpy2DAsync(mat.p(j), dim * sizeof(c64), col.p(), sizeof(c64), sizeof(c64), dim,
                                 hipMemcpyDeviceToDevice, st));
    }
    // the dense solve of src/eo_multigrid.h:188-220 (LU without pivoting) as the
    // inverse, formed once by Gauss-Jordan without pivoting
    inv = Field<double>(dim * dim);
    launch_copy<double>(inv.p(), mat.p(), dim * dim, st);
    Field<double> prow(dim), pcol(dim);
    launch_dense_invert(inv.p(), dim, prow.p(), pcol.p(), st);
    // estimate_lambda_max(A, n, 30) from the caller's start vector; omega = 0.8 / lambda_max
    Field<double> v(n), Av(n);
    SMG_HIP(hipMemcpyAsync(v.p(), lam_start, n * sizeof(c64), hipMemcpyHostToDevice, st));
    lmax = lambda_max<double>(A, v.p(), Av.p(), 30, st);
    omega = 0.8 / lmax;
    r = Field<double>(n);
    rc = Field<double>(dim);
    ec = Field<double>(dim);
    SMG_HIP(hipStreamSynchronize(st));
}

// src/eo_multigrid.h:257-286: Richardson pre-smoothing from x = 0, the residual
// restricted, the dense coarse solve prolonged and added, Richardson post-smoothing
void EOMG::precondition(LinOp<double>& Aop, const c64* b, c64* x, hipStream_t st) {
    SMG_REQUIRE(Aop.n() == A.n(), "even-odd multigrid: the operator's size differs from the half-lattice's");
    richardson_smooth<double>(Aop, x, b, r.p(), pre, omega, true, st);   // leaves r = b - A x
    launch_restrict<double>(P.A, P.bv64.p(), r.p(), rc.p(), st);
    launch_dense_matvec<double>(inv.p(), rc.p(), ec.p(), dim, st);
    launch_prolong<double>(P.A, P.bv64.p(), ec.p(), x, 1, st);
    richardson_smooth<double>(Aop, x, b, r.p(), post, omega, false, st);
}

// src/multigrid.cpp:595-687 with geometric aggregation at every level and the
// 9-point Galerkin operator at every coarse level.
MG::MG(const Dirac& d, const MGConfig& c, const c64* starts, bool on_device, bool schur) : cfg(c), D(&d) {
    SMG_REQUIRE(c.n_levels >= 2 && c.n_levels <= 5, "n_levels must be 2..5");
    hipStream_t st = Context::get().stream;
    if (schur) {
        fine64.reset(new SchurOp<double>(d));
        fine32.reset(new SchurOp<float>(d));
    } else {
        fine64.reset(new DdagDOp<double>(d));
        fine32.reset(new DdagDOp<float>(d));
    }
    SMG_REQUIRE(!schur, "multigrid on the Schur operator is not supported");
    dims.push_back(fine64->n());
    lnx.push_back(d.L);
    lny.push_back(d.L);
    lnd.push_back(2);
    omega.assign(c.n_levels, 0.0);
    size_t soff = 0;
    std::vector<std::unique_ptr<LinOp<double>>> ops64;
    LinOp<double>* cur = fine64.get();
    for (int l = 0; l + 1 < c.n_levels; l++) {
        const int k = c.k[l];
        SMG_REQUIRE(!c.chiral || k % 2 == 0, "chiral aggregation needs an even k at every level");
        const int nv = c.chiral ? k / 2 : k;   // null vectors of this level
        const size_t n = dims[l];
        nullv.emplace_back(new Field<double>((size_t)nv * n));
        SMG_HIP(hipMemcpyAsync(nullv.back()->p(), starts + soff, (size_t)nv * n * sizeof(c64),
                               on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st));
        soff += (size_t)nv * n;
        near_null(l, *cur, nv, c.null_iters, nullv.back()->p(), st);
        P.emplace_back(new Aggregation(lnx[l], lny[l], lnd[l], c.bx[l], c.by[l], k, c.chiral));
        P.back()->build(nullv.back()->p(), st);
        C.emplace_back(new CoarseOperator(P.back()->A.nbx, P.back()->A.nby, k));
        galerkin(l, st);
        ops64.emplace_back(new CoarseLinOp<double, c64>(*C.back()));
        cur = ops64.back().get();
        dims.push_back(C.back()->n());
        lnx.push_back(P.back()->A.nbx);
        lny.push_back(P.back()->A.nby);
        lnd.push_back(k);
    }
    if (cfg.coarse_solver == 1) build_dense_inverse(st);
    if (cfg.coarse_solver == 2) cheb_bounds(st);
    if (cheb_smooth_ratio > 0.0) set_cheb_smoother(cheb_smooth_ratio, cheb_smooth_fused, st);
    cyc64.reset(new Cycle<double>(*this, fine64.get()));
    cyc32.reset(new Cycle<float>(*this, fine32.get()));
    SMG_HIP(hipStreamSynchronize(st));
}
MG::~MG() { reset_graphs(); }

// level l's Galerkin operator from level l's operator (the fine D^dag D or the
// coarse operator of level l - 1), probed 32 columns at a time
void MG::galerkin(size_t l, hipStream_t st) {
    constexpr int kProbeBatch = 32;
    if (l == 0) {
        DdagDOpB<double> A(*D, kProbeBatch);
        C[0]->build_b(A, *P[0], st);
    } else {
        CoarseLinOpB<double, c64> A(*C[l - 1], kProbeBatch);
        C[l]->build_b(A, *P[l], st);
    }
}

void MG::rebuild(hipStream_t st) {
    for (size_t l = 0; l < C.size(); l++) galerkin(l, st);
    if (cfg.coarse_solver == 1) build_dense_inverse(st);
    if (cfg.coarse_solver == 2) cheb_bounds(st);
    if (cheb_smooth_ratio > 0.0) set_cheb_smoother(cheb_smooth_ratio, cheb_smooth_fused, st);
    reset_graphs();   // captured cycles are recaptured on their next use
    SMG_HIP(hipStreamSynchronize(st));
}

void MG::refresh(int iters, hipStream_t st) {
    LinOp<double>* cur = fine64.get();
    std::vector<std::unique_ptr<LinOp<double>>> ops64;
    for (size_t l = 0; l < C.size(); l++) {
        const int nv = cfg.chiral ? cfg.k[l] / 2 : cfg.k[l];
        if (iters > 0) near_null(l, *cur, nv, iters, nullv[l]->p(), st);
        P[l]->build(nullv[l]->p(), st);
        galerkin(l, st);
        ops64.emplace_back(new CoarseLinOp<double, c64>(*C[l]));
        cur = ops64.back().get();
    }
    if (cfg.coarse_solver == 1) build_dense_inverse(st);
    if (cfg.coarse_solver == 2) cheb_bounds(st);
    if (cheb_smooth_ratio > 0.0) set_cheb_smoother(cheb_smooth_ratio, cheb_smooth_fused, st);
    reset_graphs();
    SMG_HIP(hipStreamSynchronize(st));
}

void MG::reset_graphs() {
    release_batched_cycles(this);   // the batched cycles' graphs too
    for (auto& g : graph) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        if (g.g) (void)hipGraphDestroy(g.g);
        g.exec = nullptr;
        g.g = nullptr;
    }
}

// The coarsest operator is Hermitian positive definite (a Galerkin projection of
// D^dag D), so its inverse is formed once by Gauss-Jordan without pivoting, and a
// coarsest solve becomes one dense matrix-vector product (the exact solve of
// src/coarse_op.h:101-137 without refactorising on every call).
void MG::build_dense_inverse(hipStream_t st) {
    const size_t n = dims.back();
    SMG_REQUIRE(n <= 8192, "dense coarsest inverse: the coarsest level must have <= 8192 unknowns");
    const CoarseOperator& Cc = *C.back();
    inv64 = Field<double>(n * n);
    inv32 = Field<float>(n * n);
    launch_zero<double>(inv64.p(), n * n, st);
    launch_coarse_to_dense(CoarseView<c64>{Cc.l64.p(), Cc.nx, Cc.ny, Cc.k}, inv64.p(), n, st);
    Field<double> prow(n), pcol(n);
    launch_dense_invert(inv64.p(), n, prow.p(), pcol.p(), st);
    launch_convert<float, double>(inv32.p(), inv64.p(), n * n, st);
}

// The coarsest operator's spectral interval for the Chebyshev coarsest solve: Lanczos on
// the FP64 Galerkin operator from a fixed pseudo-random start (so the interval, and the
// Chebyshev coefficients, depend only on the operator)
void MG::cheb_bounds(hipStream_t st) {
    const CoarseOperator& Cc = *C.back();
    CoarseLinOp<double, c64> A(Cc);
    const size_t n = A.n();
    std::vector<c64> h(n);
    std::mt19937_64 rng(0x5eed);
    std::normal_distribution<double> g;
    for (auto& v : h) v = c64{g(rng), g(rng)};
    Field<double> start(n);
    SMG_HIP(hipMemcpyAsync(start.p(), h.data(), n * sizeof(c64), hipMemcpyHostToDevice, st));
    double lo = 0.0, hi = 0.0;
    lanczos_extremes(A, kChebLanczos, start.p(), lo, hi, st);
    SMG_REQUIRE(lo > 0.0 && hi > lo, "Chebyshev coarsest solve: the coarsest operator is not positive definite");
    cheb_lo = lo;
    cheb_hi = kChebMargin * hi;
}

void MG::set_cheb_smoother(double ratio, bool fused, hipStream_t st) {
    cheb_smooth_ratio = ratio > 0.0 ? ratio : 0.0;
    cheb_smooth_fused = fused;
    cheb_smooth_hi.assign(cfg.n_levels, 0.0);
    if (cheb_smooth_ratio > 0.0) {
        for (int l = 0; l + 1 < cfg.n_levels; l++) {
            const size_t n = dims[l];
            std::vector<c64> h(n);
            std::mt19937_64 rng(0x5eed + l);
            std::normal_distribution<double> g;
            for (auto& v : h) v = c64{g(rng), g(rng)};
            Field<double> v(n), Av(n);
            SMG_HIP(hipMemcpyAsync(v.p(), h.data(), n * sizeof(c64), hipMemcpyHostToDevice, st));
            cheb_smooth_hi[l] = kChebMargin * lambda_max<double>(*cyc64->ops[l], v.p(), Av.p(), 40, st);
        }
    }
    reset_graphs();
    SMG_HIP(hipStreamSynchronize(st));
}

// src/multigrid.cpp:146-152 with explicit power-iteration start vectors
void MG::set_symmetric(double damping, const c64* starts, bool on_device) {
    hipStream_t st = Context::get().stream;
    size_t off = 0;
    for (int l = 0; l + 1 < cfg.n_levels; l++) {
        const size_t n = dims[l];
        Field<double> v(n), Av(n);
        SMG_HIP(hipMemcpyAsync(v.p(), starts + off, n * sizeof(c64),
                               on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st));
        off += n;
        const double lmax = lambda_max<double>(*cyc64->ops[l], v.p(), Av.p(), 30, st);
        omega[l] = damping / lmax;
    }
    reset_graphs();
}

// Replays a hipGraph of the whole cycle (captured on first use with fixed
// input/output buffers); the cycle makes no host synchronisation and no
// allocation, so it captures as is.
template <typename T>
static void run_graphed(MG& mg, Cycle<T>& cyc, MG::Graph& G, Field<T>& gin, Field<T>& gout, const cplx<T>* b,
                        cplx<T>* x, hipStream_t st) {
    const size_t n = mg.dims[0];
    if (!G.exec) {
        gin = Field<T>(n);
        gout = Field<T>(n);
        SMG_HIP(hipStreamSynchronize(st));
        SMG_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        try {
            cyc.run(0, gin.p(), gout.p(), st);
        } catch (...) {
            hipGraph_t tmp;
            (void)hipStreamEndCapture(st, &tmp);
            if (tmp) (void)hipGraphDestroy(tmp);
            throw;
        }
        SMG_HIP(hipStreamEndCapture(st, &G.g));
        SMG_HIP(hipGraphInstantiate(&G.exec, G.g, nullptr, nullptr, 0));
    }
    launch_copy<T>(gin.p(), b, n, st);
    SMG_HIP(hipGraphLaunch(G.exec, st));
    launch_copy<T>(x, gout.p(), n, st);
}

void MG::precondition(const c64* b, c64* x, hipStream_t st) {
    // the deflated coarsest CG reads its scalars back on the host: not capturable
    if (cfg.use_graph && !sparse_defl) run_graphed<double>(*this, *cyc64, graph[0], gin64, gout64, b, x, st);
    else cyc64->run(0, b, x, st);
}
void MG::precondition(const c32* b, c32* x, hipStream_t st) {
    if (cfg.use_graph) run_graphed<float>(*this, *cyc32, graph[1], gin32, gout32, b, x, st);
    else cyc32->run(0, b, x, st);
}

// ------------------------------------------------------------------ Krylov
double true_residual(LinOp<double>& A, const c64* b, const c64* x, hipStream_t st) {
    const size_t n = A.n();
    Field<double> r(n);
    A.apply(r.p(), x, EPI_RESID, b, nullptr, Fin(), st);
    return dev_norm(r.p(), n, st) / dev_norm(b, n, st);
}

template <typename T>
static void precond_apply(Preconditioner* mg, const cplx<T>* r, cplx<T>* z, size_t n, hipStream_t st) {
    if (mg) mg->precondition(r, z, st);
    else dcopy(z, r, n, st);
}

// src/solvers.cpp:16-72
SolveResult cg_solve(LinOp<double>& A, const c64* b, c64* x, const c64* x0, Preconditioner* mg, int maxit, double tol,
                     hipStream_t st) {
    SolveResult res;
    const size_t n = A.n();
    Context& cx = Context::get();
    const double bnorm = dev_norm(b, n, st);
    if (bnorm < 1e-30) {
        launch_zero<double>(x, n, st);
        return res;
    }
    Field<double> r(n), z(n), p(n), Ap(n);
    Scalars sc(64);
    double* alpha = sc.p(0);
    double* beta = sc.p(2);
    double* rr = sc.p(4);
    double* aux = sc.p(16);
    int* flag = reinterpret_cast<int*>(aux + 8);
    launch_set_scalars(aux + 5, 1, 0.0, 0, 0, 0, flag, st);
    if (x0) {
        dcopy(x, x0, n, st);
        A.apply(r.p(), x, EPI_RESID, b, nullptr, Fin(), st);
    } else {
        launch_zero<double>(x, n, st);
        dcopy(r.p(), b, n, st);
    }
    c64* zp = mg ? z.p() : r.p();
    if (mg) mg->precondition(r.p(), z.p(), st);
    dcopy(p.p(), zp, n, st);
    launch_dot<double>(r.p(), zp, n, cx.fin(FIN_STORE, aux), st);
    int iter = 0;
    double rn = bnorm;
    while (iter < maxit) {
        A.apply(Ap.p(), p.p(), EPI_DOTS, nullptr, p.p(), cx.fin(FIN_CG_ALPHA, alpha, aux), st);
        launch_cg_update<double>(x, r.p(), p.p(), Ap.p(), alpha, flag, n, cx.fin(FIN_STORE, rr), st);
        iter++;
        double h;
        cx.read(rr, &h, 1);
        rn = std::sqrt(h);
        if (rn / bnorm < tol) break;
        if (mg) mg->precondition(r.p(), z.p(), st);
        launch_dot<double>(r.p(), zp, n, cx.fin(FIN_PCG_BETA, beta, aux), st);
        launch_xpay_dev<double>(p.p(), zp, beta, flag, n, st);
    }
    res.iterations = iter;
    res.final_residual = rn / bnorm;
    return res;
}

// src/solvers.cpp:223-300
template <typename T>
SolveResult gcr_solve(LinOp<T>& A, const cplx<T>* b, cplx<T>* x, Preconditioner* mg, int maxit, double tol, int trunc,
                      bool ref_projection, hipStream_t st) {
    SolveResult res;
    const size_t n = A.n();
    SMG_REQUIRE(trunc >= 1 && trunc <= 32, "GCR truncation window must be 1..32");
    Context& cx = Context::get();
    launch_zero<T>(x, n, st);
    const double bnorm = dev_norm_t(b, n, st);
    if (bnorm < 1e-30) return res;
    Field<T> r(n);
    dcopy(r.p(), b, n, st);
    std::vector<Field<T>> Pd, APd;
    for (int i = 0; i <= trunc; i++) {
        Pd.emplace_back(n);
        APd.emplace_back(n);
    }
    Scalars sc(128);   // [0..1] alpha, [2] rr, [8..] dots (64), [80..] ApAp ring (33)
    double* alpha = sc.p(0);
    double* rr = sc.p(2);
    double* dots = sc.p(8);
    double* apap = sc.p(80);
    std::deque<int> order;
    std::vector<int> freeslots;
    for (int i = trunc; i >= 0; i--) freeslots.push_back(i);
    double rn = bnorm;
    int iter = 0, collapses = 0;
    while (iter < maxit) {
        if (rn / bnorm < tol) break;
        const int slot = freeslots.back();
        freeslots.pop_back();
        cplx<T>* p = Pd[slot].p();
        cplx<T>* Ap = APd[slot].p();
        precond_apply<T>(mg, r.p(), p, n, st);
        A.apply(Ap, p, st);
        const int nd = (int)order.size();
        if (nd > 0) {
            VecPtrs<T> Pv, APv;
            const double* ap[32];
            for (int i = 0; i < nd; i++) {
                Pv.v[i] = Pd[order[i]].p();
                APv.v[i] = APd[order[i]].p();
                ap[i] = apap + order[i];
            }
            // beta_i from the unmodified A z (classical Gram-Schmidt, src/solvers.cpp:260-266)
            launch_multi_dot<T>(APv, nd, Ap, n, ref_projection ? 1 : 0, cx.fin(FIN_STORE, dots), st);
            launch_gcr_orth<T>(p, Ap, Pv, APv, nd, dots, ap, n, st);
        }
        launch_dots<T>(Ap, r.p(), n, cx.fin(FIN_GCR_ALPHA, alpha, apap + slot), st);
        launch_cg_update<T>(x, r.p(), p, Ap, alpha, reinterpret_cast<const int*>(sc.p(120)), n,
                            cx.fin(FIN_STORE, rr), st);
        double h[2];
        cx.read(rr, &h[0], 1);
        cx.read(apap + slot, &h[1], 1);
        if (h[1] < 1e-30) {
            // direction collapsed: clear the history and restart (src/solvers.cpp:272-276)
            freeslots.push_back(slot);
            for (int s : order) freeslots.push_back(s);
            order.clear();
            if (++collapses > 3) break;   // the preconditioned direction stays zero
            continue;
        }
        collapses = 0;
        rn = std::sqrt(h[0]);
        iter++;
        if ((int)order.size() >= trunc) {
            freeslots.push_back(order.front());
            order.pop_front();
        }
        order.push_back(slot);
    }
    res.iterations = iter;
    res.final_residual = rn / bnorm;
    return res;
}
template SolveResult gcr_solve<double>(LinOp<double>&, const c64*, c64*, Preconditioner*, int, double, int, bool,
                                        hipStream_t);
template SolveResult gcr_solve<float>(LinOp<float>&, const c32*, c32*, Preconditioner*, int, double, int, bool,
                                       hipStream_t);

SolveResult gcr_solve_mixed(LinOp<double>& A64, LinOp<float>& A32, const c64* b, c64* x, Preconditioner* mg, int maxit,
                            double tol, int trunc, double inner_tol, hipStream_t st, bool use_x0) {
    SolveResult res;
    const size_t n = A64.n();
    Context& cx = Context::get();
    if (!use_x0) launch_zero<double>(x, n, st);
    const double bnorm = dev_norm(b, n, st);
    if (bnorm < 1e-30) {
        launch_zero<double>(x, n, st);
        return res;
    }
    Field<double> r(n);
    Field<float> r32(n), e32(n);
    Scalars sc(8);
    double rn = bnorm;
    if (use_x0) {
        A64.apply(r.p(), x, EPI_RESID, b, nullptr, Fin(), st);
        rn = dev_norm(r.p(), n, st);
    } else {
        dcopy(r.p(), b, n, st);
    }
    int total = 0;
    while (total < maxit) {
        if (rn / bnorm < tol) break;
        launch_convert<float, double>(r32.p(), r.p(), n, st);
        // inner FP32 solve to a relative residual max(inner_tol, tol*|b|/|r|)
        const double itol = std::max(inner_tol, 0.5 * tol * bnorm / rn);
        SolveResult in = gcr_solve<float>(A32, r32.p(), e32.p(), mg, maxit - total, itol, trunc, false, st);
        total += std::max(in.iterations, 1);
        launch_accum_mixed(x, e32.p(), n, st);
        // FP64 true residual
        A64.apply(r.p(), x, EPI_RESID, b, nullptr, Fin(), st);
        launch_norm2<double>(r.p(), n, cx.fin(FIN_STORE, sc.p()), st);
        double h;
        cx.read(sc.p(), &h, 1);
        rn = std::sqrt(h);
        res.restarts++;
    }
    res.iterations = total;
    res.final_residual = rn / bnorm;
    res.true_residual = rn / bnorm;
    return res;
}

// src/solvers.cpp:555-772; the Ritz harvest and intra-solve deflation when ritz
SolveResult fgmres_solve(LinOp<double>& A, const c64* b, c64* x, Preconditioner* mg, int restart, int maxit, double tol,
                         hipStream_t st, RitzHarvest* ritz) {
    SolveResult res;
    const size_t n = A.n();
    launch_zero<double>(x, n, st);
    const double bnorm = dev_norm(b, n, st);
    if (bnorm < 1e-30) return res;
    std::vector<Field<double>> Vv;
    for (int i = 0; i <= restart; i++) Vv.emplace_back(n);
    Field<double> Z((size_t)restart * n);   // preconditioned vectors, contiguous for the harvest
    auto zk = [&](int k) { return Z.p((size_t)k * n); };
    Field<double> w(n);
    int total = 0;
    auto cdivh = [](c64 a, c64 b) { return cdiv(a, b); };
    while (total < maxit) {
        A.apply(Vv[0].p(), x, EPI_RESID, b, nullptr, Fin(), st);
        const double beta = dev_norm(Vv[0].p(), n, st);
        if (beta / bnorm < tol) break;
        const int m = std::min(restart, maxit - total);
        launch_axpby<double>(Vv[0].p(), Vv[0].p(), c64{0, 0}, c64{1.0 / beta, 0}, n, st);
        std::vector<c64> H((size_t)(restart + 1) * restart, c64{0, 0}), g(restart + 1, c64{0, 0}),
            cs(restart, c64{0, 0}), sn(restart, c64{0, 0}), y(restart, c64{0, 0});
        std::vector<c64> Hraw(ritz ? H.size() : 0, c64{0, 0});   // the Hessenberg before the rotations
        auto h = [&](int i, int j) -> c64& { return H[(size_t)i * restart + j]; };
        g[0] = c64{beta, 0};
        int k;
        for (k = 0; k < m; k++) {
            precond_apply<double>(mg, Vv[k].p(), zk(k), n, st);
            if (ritz) ritz->deflate(Vv[k].p(), zk(k), st);
            A.apply(w.p(), zk(k), st);
            for (int pass = 0; pass < 2; pass++)
                for (int j = 0; j <= k; j++) {
                    c64 c = dev_dot<double>(Vv[j].p(), w.p(), n, st);
                    h(j, k) = h(j, k) + c;
                    launch_axpby<double>(w.p(), Vv[j].p(), c64{-c.re, -c.im}, c64{1, 0}, n, st);
                }
            const double hn = dev_norm(w.p(), n, st);
            h(k + 1, k) = c64{hn, 0};
            if (ritz)
                for (int j = 0; j <= k + 1; j++) Hraw[(size_t)j * restart + k] = h(j, k);
            if (hn < 1e-14) {
                k++;
                break;
            }
            launch_axpby<double>(Vv[k + 1].p(), w.p(), c64{1.0 / hn, 0}, c64{0, 0}, n, st);
            for (int j = 0; j < k; j++) {
                const c64 hj = h(j, k), hj1 = h(j + 1, k);
                h(j, k) = conj(cs[j]) * hj + conj(sn[j]) * hj1;
                h(j + 1, k) = (-sn[j]) * hj + cs[j] * hj1;
            }
            const double a = std::sqrt(norm2(h(k, k))), bb = std::sqrt(norm2(h(k + 1, k)));
            const double den = std::sqrt(a * a + bb * bb);
            cs[k] = c64{h(k, k).re / den, h(k, k).im / den};
            sn[k] = c64{h(k + 1, k).re / den, h(k + 1, k).im / den};
            g[k + 1] = (-sn[k]) * g[k];
            g[k] = conj(cs[k]) * g[k];
            h(k, k) = c64{den, 0};
            h(k + 1, k) = c64{0, 0};
            total++;
            if (std::sqrt(norm2(g[k + 1])) / bnorm < tol) {
                k++;
                break;
            }
        }
        const int kk = k;
        for (int i = kk - 1; i >= 0; i--) {
            c64 s = g[i];
            for (int j = i + 1; j < kk; j++) s = s - h(i, j) * y[j];
            y[i] = cdivh(s, h(i, i));
        }
        for (int j = 0; j < kk; j++) launch_axpby<double>(x, zk(j), y[j], c64{1, 0}, n, st);
        if (ritz) ritz->harvest(Hraw, restart, kk, Z.p(), st);
        if (std::sqrt(norm2(g[kk])) / bnorm < tol) break;
    }
    res.iterations = total;
    res.final_residual = true_residual(A, b, x, st);
    res.true_residual = res.final_residual;
    return res;
}

}  // namespace smg
