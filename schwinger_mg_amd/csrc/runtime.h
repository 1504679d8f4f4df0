// runtime.h -- host-side C++ runtime of the gfx950 Schwinger-MG library.
//
// Objects mirror the reference's operator API (src/dirac.h, src/prolongator.h,
// src/coarse_op.h, src/multigrid.h, src/solvers.h) but keep every field in
// device memory; host code only sequences kernels on one HIP stream.
#pragma once
#include <memory>
#include <vector>
#include "smg_kernels.h"

namespace smg {

// ------------------------------------------------------------------ memory
struct DevMem {
    void* p = nullptr;
    size_t bytes = 0;
    DevMem() = default;
    explicit DevMem(size_t b) { alloc(b); }
    ~DevMem() { release(); }
    DevMem(const DevMem&) = delete;
    DevMem& operator=(const DevMem&) = delete;
    DevMem(DevMem&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    DevMem& operator=(DevMem&& o) noexcept {
        if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
        return *this;
    }
    void alloc(size_t b);
    void release();
    template <typename X>
    X* as() const { return reinterpret_cast<X*>(p); }
};

template <typename T>
struct Field {
    DevMem m;
    size_t n = 0;  // complex entries
    Field() = default;
    explicit Field(size_t n_) : m(n_ * sizeof(cplx<T>)), n(n_) {}
    cplx<T>* p() const { return m.as<cplx<T>>(); }
    cplx<T>* p(size_t i) const { return m.as<cplx<T>>() + i; }
};

// Per-process execution context: one stream, reduction partials, scalars.
struct Context {
    hipStream_t stream = nullptr;
    DevMem part;      // reduction partial sums
    DevMem counter;   // workgroup-completion counter of the fused reductions (zero between kernels)
    DevMem hostscal;  // pinned host mirror (hipHostMalloc)
    static Context& get();
    double* part_ptr() const { return part.as<double>(); }
    // a reduction request finishing with scalar program op into out (aux: program state)
    Fin fin(int op, double* out, double* aux = nullptr) const {
        Fin f;
        f.op = op;
        f.out = out;
        f.aux = aux;
        f.part = part.as<double>();
        f.counter = counter.as<unsigned>();
        return f;
    }
    // read `count` doubles from device into host (synchronises the stream)
    void read(const double* dev, double* host, size_t count);
    void sync();
    ~Context();
};

// ------------------------------------------------------------------ operators
template <typename T>
struct LinOp {
    virtual ~LinOp() = default;
    virtual size_t n() const = 0;
    // dst = A src | b - A src | A src with the reduction {<dst,aux>, |dst|^2} finished by fin
    virtual void apply(cplx<T>* dst, const cplx<T>* src, Epi epi, const cplx<T>* b, const cplx<T>* aux,
                       const Fin& fin, hipStream_t st) = 0;
    void apply(cplx<T>* dst, const cplx<T>* src, hipStream_t st) { apply(dst, src, EPI_NONE, nullptr, nullptr, Fin(), st); }
};

// Wilson-clover(-twisted) Dirac operator on an L x L torus (src/dirac.h:8-183).
struct Dirac {
    int L, V;
    double mass, r, csw, mu_t;
    Field<double> U64;
    Field<float> U32;
    DevMem F01;  // clover field F01 (double [V])
    DevMem C64;  // 0.5*csw*F01 (double [V])
    DevMem C32;  // float [V]
    Dirac(int L, const c64* U_host, double mass, double r, double csw, double mu_t);
    void set_gauge(const c64* U, bool device_ptr);
    // U <- U exp(i scale pi) on the device (pi: 2V reals), then the FP32 links and
    // the clover term are recomputed
    void update_links(const double* pi, double scale, hipStream_t st);
    // FP32 links, F01 and the clover term from U64
    void refresh_derived(hipStream_t st);
    DiracView<double> v64() const;
    DiracView<float> v32() const;
    template <typename T>
    DiracView<T> view() const;
};

// D^dag D (src/dirac.h:136-140) as a fused pair of stencil launches.
template <typename T>
struct DdagDOp : LinOp<T> {
    const Dirac& D;
    Field<T> tmp;
    explicit DdagDOp(const Dirac& d) : D(d), tmp(2 * (size_t)d.V) {}
    size_t n() const override { return 2 * (size_t)D.V; }
    void apply(cplx<T>* dst, const cplx<T>* src, Epi epi, const cplx<T>* b, const cplx<T>* aux, const Fin& fin,
               hipStream_t st) override;
    using LinOp<T>::apply;
};

// M^dag M on odd sites (src/dirac.h:399-445), four half-lattice launches.
template <typename T>
struct SchurOp : LinOp<T> {
    const Dirac& D;
    Field<T> te, to;
    explicit SchurOp(const Dirac& d) : D(d), te((size_t)d.V), to((size_t)d.V) {}
    size_t n() const override { return (size_t)D.V; }
    void apply(cplx<T>* dst, const cplx<T>* src, Epi epi, const cplx<T>* b, const cplx<T>* aux, const Fin& fin,
               hipStream_t st) override;
    using LinOp<T>::apply;
    // M (dag = 0) or M^dag (dag = 1) only
    void apply_m(cplx<T>* dst, const cplx<T>* src, int dag, hipStream_t st);
};

// Aggregation (geometric block) prolongator, src/prolongator.h:14-142.
struct Aggregation {
    AggView A;
    Field<double> bv64;
    Field<float> bv32;
    // k coarse dof per aggregate; chiral: k = 2 * (null vectors), each vector split
    // into its two chiralities
    Aggregation(int nx, int ny, int nd, int bx, int by, int k, int chiral = 0);
    size_t fine_n() const { return (size_t)A.nx * A.ny * A.nd; }
    size_t coarse_n() const { return (size_t)A.nbx * A.nby * A.k; }
    int nvec() const { return A.chiral ? A.k / 2 : A.k; }
    // null_vecs: nvec() device vectors of fine_n() entries, contiguous
    void build(const c64* null_vecs, hipStream_t st);
    void set_block_vecs(const c64* bv_dev, hipStream_t st);
    template <typename T>
    const cplx<T>* bv() const;
};

// 9-point coarse operator, src/coarse_op.h:154-297, stored in FP64, FP32, FP16.
template <typename T>
struct LinOpB;

// the colour period of the probing along a coarse direction of nb sites
int color_period(int nb);

struct CoarseOperator {
    int nx, ny, k;
    Field<double> l64;
    Field<float> l32;
    DevMem l16;
    bool hermitian = false;   // a Galerkin projection of a Hermitian operator (build, build_b)
    int herm16 = 0;           // l16 made exactly Hermitian (hermitian and both extents >= 3)
    CoarseView<ch16> view16() const { return CoarseView<ch16>{l16.as<ch16>(), nx, ny, k, herm16}; }
    CoarseOperator(int nx, int ny, int k);
    size_t n() const { return (size_t)nx * ny * k; }
    size_t nlinks() const { return (size_t)nx * ny * 9 * k * k; }
    // Galerkin P^dag A P by coloured probing in FP64
    void build(LinOp<double>& A, const Aggregation& P, hipStream_t st);
    // the same probing with A.N probes at a time, each batch prolonged, applied and
    // restricted as one block of right-hand sides
    void build_b(LinOpB<double>& A, const Aggregation& P, hipStream_t st);
    void set_links(const c64* dev, hipStream_t st);
    void refresh_low_precision(hipStream_t st);
};

template <typename T, typename TL>
struct CoarseLinOp : LinOp<T> {
    const CoarseOperator& C;
    explicit CoarseLinOp(const CoarseOperator& c) : C(c) {}
    size_t n() const override { return C.n(); }
    void apply(cplx<T>* dst, const cplx<T>* src, Epi epi, const cplx<T>* b, const cplx<T>* aux, const Fin& fin,
               hipStream_t st) override;
    using LinOp<T>::apply;
};

// ------------------------------------------------------------------ smoothers
// MR smoother (src/smoother.cpp:26-40).  x holds the initial guess unless
// zero_init; on return r holds the recursive residual b - A x.
template <typename T>
void mr_smooth(LinOp<T>& A, cplx<T>* x, const cplx<T>* b, cplx<T>* r, cplx<T>* Ar, int niter, bool zero_init,
               double* sc, hipStream_t st);
// Richardson (src/smoother.cpp:63-73); on return r = b - A x (recomputed)
template <typename T>
void richardson_smooth(LinOp<T>& A, cplx<T>* x, const cplx<T>* b, cplx<T>* r, int niter, double omega,
                       bool zero_init, hipStream_t st);
// power iteration (src/smoother.cpp:43-59) from the start vector v (overwritten)
template <typename T>
double lambda_max(LinOp<T>& A, cplx<T>* v, cplx<T>* Av, int niter, hipStream_t st);

// CG on a (coarse) operator with device-side scalars (src/coarse_op.h:492-547
// without deflation).  Returns iterations if check != 0, else -1 (not read back).
template <typename T>
int coarse_cg(LinOp<T>& A, const cplx<T>* b, cplx<T>* x, cplx<T>* r, cplx<T>* p, cplx<T>* Ap, int maxit, double tol,
              double* sc, bool check, hipStream_t st);

// ------------------------------------------------------------------ multigrid
struct MGParamsC;  // C-ABI parameter struct (smg.h)

struct MGConfig {
    int n_levels = 2;
    int bx[4] = {4, 2, 2, 2}, by[4] = {4, 2, 2, 2}, k[4] = {8, 8, 8, 8};
    int null_iters = 20;
    int null_inner = 5;
    int pre = 3, post = 3;
    int w_cycle = 1;
    int coarse_maxit = 200;
    double coarse_tol = 1e-12;
    int smoother = 0;        // 0 MR, 1 Richardson (set by set_symmetric)
    int coarse_half = 0;     // FP32 cycle: FP16 coarse links
    // coarsest level: 0 CG (coarse_maxit, coarse_tol), 1 dense inverse, 2 CG in the FP64
    // cycles and, in the batched FP32 cycle with BF16 coarsest vectors (coarse_half 2), a
    // Chebyshev iteration of coarse_maxit steps on the Lanczos-estimated spectrum
    int coarse_solver = 0;
    int use_graph = 0;       // replay the cycle from a captured hipGraph
    int chiral = 0;          // chiral aggregation: k[l] = 2 * (null vectors of level l)
    int null_prec = 0;       // near-null MR sweeps: 0 FP64, 1 FP32 (vectors and orthogonalisation FP64)
};

template <typename T>
struct Cycle;
struct Eigenpairs;
struct RitzHarvest;

// A preconditioner z = M r on device vectors: the MG cycle, or a caller's callable
// (the reference's std::function<Vec(const Vec&)>, src/solvers.h:45-84).
struct Preconditioner {
    virtual ~Preconditioner() = default;
    virtual void precondition(const c64* b, c64* x, hipStream_t st) = 0;
    virtual void precondition(const c32* b, c32* x, hipStream_t st) {
        (void)b, (void)x, (void)st;
        throw Error(-1, "this preconditioner has no FP32 form");
    }
};

struct MG : Preconditioner {
    MGConfig cfg;
    const Dirac* D = nullptr;
    std::vector<std::unique_ptr<Aggregation>> P;
    std::vector<std::unique_ptr<CoarseOperator>> C;
    std::vector<std::unique_ptr<Field<double>>> nullv;  // per transfer, k vectors
    std::vector<double> omega;                          // Richardson omegas (0 = MR)
    std::unique_ptr<LinOp<double>> fine64;
    std::unique_ptr<LinOp<float>> fine32;
    std::unique_ptr<Cycle<double>> cyc64;
    std::unique_ptr<Cycle<float>> cyc32;
    std::vector<size_t> dims;   // per level
    std::vector<int> lnx, lny, lnd;
    // coarsest-level deflation space (MGHierarchy::levels[last].defl_vecs / defl_vals)
    std::vector<double> defl_vals;
    Field<double> defl_vecs;
    // setup_sparse_coarse: the coarsest solve of the FP64 cycle becomes the deflated CG
    // of SparseCoarseOp::solve with this space (two-level hierarchies)
    std::unique_ptr<Eigenpairs> sparse_defl;
    int sparse_maxit = 200;
    double sparse_tol = 1e-12;
    // dense inverse of the coarsest operator (cfg.coarse_solver == 1)
    Field<double> inv64;
    Field<float> inv32;
    // spectral interval [cheb_lo, cheb_hi] of the coarsest operator's Chebyshev iteration
    // (cfg.coarse_solver == 2): the extreme Ritz values of kChebLanczos Lanczos steps,
    // the upper one widened by kChebMargin so that it bounds the largest eigenvalue
    double cheb_lo = 0.0, cheb_hi = 0.0;
    static constexpr int kChebLanczos = 60;
    static constexpr double kChebMargin = 1.1;
    void cheb_bounds(hipStream_t st);
    // Chebyshev smoothing in the batched cycles (ratio > 0): level l's smoother is the
    // Chebyshev iteration on [hi_l / ratio, hi_l], hi_l = kChebMargin x the power-iteration
    // estimate of level l's largest eigenvalue; MR otherwise.  Fused: each step one launch
    // of the level's operator with the updates in its epilogue, where the operator has
    // that form; else the apply and separate vector passes
    double cheb_smooth_ratio = 0.0;
    bool cheb_smooth_fused = true;
    std::vector<double> cheb_smooth_hi;
    void set_cheb_smoother(double ratio, bool fused, hipStream_t st);
    // captured cycles (cfg.use_graph): fixed input/output buffers per precision
    struct Graph {
        hipGraph_t g = nullptr;
        hipGraphExec_t exec = nullptr;
    } graph[2];
    Field<double> gin64, gout64;
    Field<float> gin32, gout32;
    void reset_graphs();
    void build_dense_inverse(hipStream_t st);
    // starts: every level's k start vectors (host or device FP64), level 0 first
    MG(const Dirac& d, const MGConfig& c, const c64* starts, bool starts_on_device, bool schur = false);
    ~MG();
    void set_symmetric(double damping, const c64* starts, bool on_device);
    // the Galerkin operators of every level re-projected for the current gauge field of
    // the fine operator, with the existing prolongators (Ac.build at level 0, then
    // MGHierarchy::rebuild_deeper_levels, src/multigrid.cpp:163-194)
    void rebuild(hipStream_t st);
    void galerkin(size_t level, hipStream_t st);
    void near_null(size_t level, LinOp<double>& A, int nv, int iters, c64* vecs, hipStream_t st);
    // An extension with no reference counterpart (the reference's warm start,
    // build_full_mg in src/mg_builder.cpp:3-31, restarts the fine level only and
    // redraws the deeper levels): every level's near-null vectors iterated iters more
    // times from their current values on the current operator, then the aggregation
    // and the Galerkin operators rebuilt
    void refresh(int iters, hipStream_t st);
    void precondition(const c64* b, c64* x, hipStream_t st) override;
    void precondition(const c32* b, c32* x, hipStream_t st) override;
    template <typename T>
    Cycle<T>& cycle();
    LinOp<double>& level_op64(int l);
};

// Even-odd multigrid on the Schur operator M^dag M of the odd half-lattice
// (EvenOddMG, src/eo_multigrid.h:26-287): near-null vectors of M^dag M by the generic
// inverse iteration, aggregation of the odd sites of bx x by blocks, a dense Galerkin
// operator P^dag M^dag M P probed column by column and inverted once, Richardson
// smoothing with omega = 0.8 / lambda_max.  The odd sites of a bx x by block (bx even)
// are, in the half-lattice's (L/2) x L grid of odd sites (half-index y*L/2 + x/2), the
// (bx/2) x by block in the same row-major order, so the prolongator is the geometric
// Aggregation on that grid.
struct EOMG {
    const Dirac& D;
    int bx, by, k, pre, post;
    SchurOp<double> A;
    Aggregation P;
    size_t dim = 0;
    Field<double> mat, inv;   // P^dag A P and its inverse, row-major dim x dim
    double lmax = 0.0, omega = 0.0;
    Field<double> r, rc, ec;
    // starts: k host vectors of the half-lattice (V entries), the null-space starts;
    // lam_start: the power iteration's start vector (host)
    EOMG(const Dirac& d, int bx, int by, int k, int pre, int post, int null_iters, const c64* starts,
         const c64* lam_start);
    // x = M b with the smoothing operator Aop (the setup's A, or a caller's)
    void precondition(LinOp<double>& Aop, const c64* b, c64* x, hipStream_t st);
};

// ------------------------------------------------------------------ Krylov
struct SolveResult {
    int iterations = 0;
    double final_residual = 0.0;   // solver's residual estimate (reference semantics)
    double true_residual = 0.0;    // |b - A x| / |b| recomputed in FP64
    int restarts = 0;
};

// cg_solve_core (src/solvers.cpp:16-72); x0 may be null; mg may be null.
SolveResult cg_solve(LinOp<double>& A, const c64* b, c64* x, const c64* x0, Preconditioner* mg, int maxit, double tol,
                     hipStream_t st);
// fcg_solve (src/solvers.cpp:223-300) in precision T; ref_projection selects the
// reference's conjugated projection coefficient.
template <typename T>
SolveResult gcr_solve(LinOp<T>& A, const cplx<T>* b, cplx<T>* x, Preconditioner* mg, int maxit, double tol, int trunc,
                      bool ref_projection, hipStream_t st);
// Mixed precision: FP64 iterative refinement (true residual in FP64) around FP32
// GCR(trunc) preconditioned by the FP32 MG cycle.
// With use_x0 the solve starts from the x it is given (its FP64 true residual first).
SolveResult gcr_solve_mixed(LinOp<double>& A64, LinOp<float>& A32, const c64* b, c64* x, Preconditioner* mg, int maxit,
                            double tol, int trunc, double inner_tol, hipStream_t st, bool use_x0 = false);
// fgmres_solve_generic (src/solvers.cpp:555-772): no Ritz harvest when ritz is null,
// else the harvest and intra-solve deflation that ritz carries
SolveResult fgmres_solve(LinOp<double>& A, const c64* b, c64* x, Preconditioner* mg, int restart, int maxit, double tol,
                         hipStream_t st, RitzHarvest* ritz = nullptr);

double true_residual(LinOp<double>& A, const c64* b, const c64* x, hipStream_t st);
double dev_norm(const c64* v, size_t n, hipStream_t st);
template <typename T>
c64 dev_dot(const cplx<T>* a, const cplx<T>* b, size_t n, hipStream_t st);

}  // namespace smg
