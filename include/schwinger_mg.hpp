// schwinger_mg.hpp -- header-only C++17 re-statement of the reference's operator
// API (cpviolator/Schwinger_MG src/types.h, lattice.h, gauge.h, dirac.h,
// prolongator.h, coarse_op.h, smoother.h, multigrid.h, solvers.h, hmc.h) on top of
// the C-ABI in smg.h.  Every computation runs on the GPU through libsmg.so;
// vectors cross the boundary as the reference's host Vec (std::vector of
// std::complex<double>).  Link with -lsmg.
//
// Differences from the reference, all at the boundary:
//  * Names live in namespace schwinger_mg (`using namespace schwinger_mg;`
//    makes a reference caller compile unchanged).
//  * DiracOp copies the gauge links to the device when it is constructed; after
//    changing gauge.U call DiracOp::update_gauge() (the reference reads the
//    GaugeField through its reference on every apply).
//  * The Krylov solvers, smoothers and the coarse build take any OpApply and any
//    preconditioner std::function.  Those this header creates
//    (DiracOp::as_DdagD_op, SparseCoarseOp::as_op, MGHierarchy::as_precond)
//    run on the GPU as they are; any other callable (a lambda, as the
//    reference's callers write) is called back on the host for each application
//    while the Krylov vectors and reductions stay on the GPU.  The mixed-precision
//    solve, the eigensolvers and the HMC drivers take this header's objects only.
//  * A GPU failure throws std::runtime_error carrying smg_last_error() (the
//    reference asserts on shape errors and has no other error path).
#pragma once
#include <array>
#include <cmath>
#include <complex>
#include <exception>
#include <functional>
#include <iomanip>
#include <iostream>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "smg.h"

namespace schwinger_mg {

using cx = std::complex<double>;
using Vec = std::vector<cx>;
using RVec = std::vector<double>;
using OpApply = std::function<void(const Vec& src, Vec& dst)>;

namespace detail {
inline void check(int rc) {
    if (rc != 0) throw std::runtime_error(std::string("schwinger_mg: ") + smg_last_error());
}
inline const double* raw(const Vec& v) { return reinterpret_cast<const double*>(v.data()); }
inline double* raw(Vec& v) { return reinterpret_cast<double*>(v.data()); }
inline void fit(Vec& v, size_t n) {
    if (v.size() != n) v.assign(n, cx(0.0, 0.0));
}
}  // namespace detail

// random_vec, src/linalg.cpp:44-49: the reference writes cx(dist(rng), dist(rng)),
// whose two draws g++ evaluates right to left, so the imaginary part takes the
// first draw; written out here so that every compiler gives the g++ stream
inline Vec random_vec(int n, std::mt19937& rng) {
    std::normal_distribution<double> dist(0.0, 1.0);
    Vec v(n);
    for (auto& x : v) {
        const double im = dist(rng);
        const double re = dist(rng);
        x = cx(re, im);
    }
    return v;
}

// the host-vector vocabulary of src/linalg.h:5-11 (dot, norm, axpy, add, scale, zeros)
// for callers that combine solver results on the host, as the reference's do; sums run
// in index order
inline cx dot(const Vec& a, const Vec& b) {   // sum conj(a_i) b_i
    cx s(0.0, 0.0);
    for (size_t i = 0; i < a.size(); i++) s += std::conj(a[i]) * b[i];
    return s;
}
inline double norm(const Vec& v) { return std::sqrt(std::real(dot(v, v))); }
inline void axpy(cx a, const Vec& x, Vec& y) {
    for (size_t i = 0; i < y.size(); i++) y[i] += a * x[i];
}
inline Vec add(const Vec& a, const Vec& b, cx alpha = 1.0, cx beta = 1.0) {
    Vec r(a.size());
    for (size_t i = 0; i < a.size(); i++) r[i] = alpha * a[i] + beta * b[i];
    return r;
}
inline void scale(Vec& v, cx a) {
    for (auto& x : v) x *= a;
}
inline Vec zeros(int n) { return Vec(n, cx(0.0, 0.0)); }

// ------------------------------------------------------------------ Lattice, src/lattice.h
struct Lattice {
    int L, V, ndof, V_half;
    std::vector<int> even_sites, odd_sites, eo_index;
    explicit Lattice(int L_) : L(L_), V(L_ * L_), ndof(2 * L_ * L_), V_half(L_ * L_ / 2) { init_eo(); }
    int idx(int x, int y) const { return ((x % L + L) % L) + L * ((y % L + L) % L); }
    int xp(int s) const { return idx(s % L + 1, s / L); }
    int xm(int s) const { return idx(s % L - 1, s / L); }
    int yp(int s) const { return idx(s % L, s / L + 1); }
    int ym(int s) const { return idx(s % L, s / L - 1); }
    int parity(int s) const { return ((s % L) + (s / L)) % 2; }
    void init_eo() {
        even_sites.assign(V_half, 0);
        odd_sites.assign(V_half, 0);
        eo_index.assign(V, 0);
        detail::check(smg_eo_maps(L, even_sites.data(), odd_sites.data(), eo_index.data()));
    }
};

// ------------------------------------------------------------------ GaugeField, src/gauge.h
struct GaugeField {
    const Lattice& lat;
    std::array<Vec, 2> U;
    explicit GaugeField(const Lattice& lat_) : lat(lat_) {
        U[0].assign(lat.V, cx(1.0, 0.0));
        U[1].assign(lat.V, cx(1.0, 0.0));
    }
    void randomise(std::mt19937& rng, double width = M_PI) {
        std::uniform_real_distribution<double> dist(-width, width);
        for (int mu = 0; mu < 2; mu++)
            for (int s = 0; s < lat.V; s++) U[mu][s] = std::exp(cx(0, dist(rng)));
    }
    cx plaq(int s) const {
        return U[0][s] * U[1][lat.xp(s)] * std::conj(U[0][lat.yp(s)]) * std::conj(U[1][s]);
    }
    Vec links() const {   // U[mu*V + s], the C-ABI gauge layout
        Vec u(2 * (size_t)lat.V);
        std::copy(U[0].begin(), U[0].end(), u.begin());
        std::copy(U[1].begin(), U[1].end(), u.begin() + lat.V);
        return u;
    }
    double avg_plaq() const {
        double out = 0.0;
        const Vec u = links();
        detail::check(smg_avg_plaq(lat.L, detail::raw(u), &out));
        return out;
    }
    // src/gauge.h:38-57: int32 V then the links; false on any failure (as the reference)
    bool save(const std::string& path) const {
        const Vec u = links();
        return smg_gauge_write_raw(path.c_str(), lat.L, detail::raw(u)) == 0;
    }
    bool load(const std::string& path) {
        Vec u(2 * (size_t)lat.V);
        if (smg_gauge_read_raw(path.c_str(), lat.L, detail::raw(u)) != 0) return false;
        U[0].assign(u.begin(), u.begin() + lat.V);
        U[1].assign(u.begin() + lat.V, u.end());
        return true;
    }
};

// gauge files with the 'SWNG' header, src/hmc.h:19-29, src/hmc_forces.cpp:18-62
using GaugeHeader = smg_gauge_header;
inline bool save_gauge(const GaugeField& g, double beta, double mass, const std::string& filename) {
    const Vec u = g.links();
    const bool ok = smg_gauge_write(filename.c_str(), g.lat.L, detail::raw(u), beta, mass) == 0;
    if (!ok) std::cerr << "Cannot open " << filename << " for writing\n";
    return ok;
}
inline bool load_gauge(GaugeField& g, GaugeHeader& hdr, const std::string& filename) {
    Vec u(2 * (size_t)g.lat.V);
    if (smg_gauge_read(filename.c_str(), g.lat.L, detail::raw(u), &hdr) != 0) {
        std::cerr << smg_last_error() << "\n";
        return false;
    }
    g.U[0].assign(u.begin(), u.begin() + g.lat.V);
    g.U[1].assign(u.begin() + g.lat.V, u.end());
    return true;
}
// therm_cfg_path, src/hmc_utils.cpp:7-11
inline std::string therm_cfg_path(int L, double beta, int n_therm) {
    char buf[256];
    detail::check(smg_therm_cfg_path(L, beta, n_therm, buf, sizeof(buf)));
    return std::string(buf);
}

// ------------------------------------------------------------------ DiracOp, src/dirac.h:8-183
struct DiracOp;
struct DdagDApply {   // the OpApply of DiracOp::as_DdagD_op
    const DiracOp* D;
    void operator()(const Vec& s, Vec& d) const;
};
struct G5DApply {     // the OpApply of DiracOp::as_g5D_op
    const DiracOp* D;
    void operator()(const Vec& s, Vec& d) const;
};

struct DiracOp {
    const Lattice& lat;
    const GaugeField& gauge;
    double mass, r, c_sw, mu_t;
    RVec clover_field;   // F_01(s) (filled when c_sw != 0, as the reference)

    DiracOp(const Lattice& lat_, const GaugeField& g_, double m_, double r_ = 1.0, double csw_ = 0.0,
            double mu_t_ = 0.0)
        : lat(lat_), gauge(g_), mass(m_), r(r_), c_sw(csw_), mu_t(mu_t_) {
        const Vec u = gauge.links();
        smg_dirac* h = nullptr;
        detail::check(smg_dirac_create(&h, lat.L, detail::raw(u), mass, r, c_sw, mu_t));
        h_.reset(h, [](smg_dirac* p) { smg_dirac_destroy(p); });
        fetch_clover();
    }
    // re-read gauge.U (and recompute the clover field) after the caller changed it
    void update_gauge() {
        const Vec u = gauge.links();
        detail::check(smg_dirac_set_gauge(h_.get(), detail::raw(u)));
        fetch_clover();
    }
    void apply(const Vec& src, Vec& dst) const { run(SMG_D, src, dst); }
    void apply_dag(const Vec& src, Vec& dst) const { run(SMG_DDAG, src, dst); }
    void apply_DdagD(const Vec& src, Vec& dst) const { run(SMG_DDAGD, src, dst); }
    void apply_g5D(const Vec& src, Vec& dst) const { run(SMG_G5D, src, dst); }
    void apply_delta_D(const Vec& src, Vec& dst, const std::array<RVec, 2>& pi, double dt) const {
        detail::fit(dst, lat.ndof);
        RVec p(2 * (size_t)lat.V);
        std::copy(pi[0].begin(), pi[0].end(), p.begin());
        std::copy(pi[1].begin(), pi[1].end(), p.begin() + lat.V);
        detail::check(smg_dirac_delta(h_.get(), detail::raw(src), detail::raw(dst), p.data(), dt));
    }
    OpApply as_DdagD_op() const { return DdagDApply{this}; }
    OpApply as_g5D_op() const { return G5DApply{this}; }
    smg_dirac* handle() const { return h_.get(); }

   private:
    std::shared_ptr<smg_dirac> h_;
    void run(int kind, const Vec& src, Vec& dst) const {
        detail::fit(dst, lat.ndof);
        detail::check(smg_dirac_apply(h_.get(), kind, detail::raw(src), detail::raw(dst)));
    }
    void fetch_clover() {
        if (c_sw == 0.0) return;
        clover_field.assign(lat.V, 0.0);
        detail::check(smg_dirac_clover(h_.get(), clover_field.data()));
    }
};
inline void DdagDApply::operator()(const Vec& s, Vec& d) const { D->apply_DdagD(s, d); }
inline void G5DApply::operator()(const Vec& s, Vec& d) const { D->apply_g5D(s, d); }

// ------------------------------------------------------------------ EvenOddDiracOp, src/dirac.h:190-553
struct EvenOddDiracOp {
    const DiracOp& D;
    const Lattice& lat;
    explicit EvenOddDiracOp(const DiracOp& D_) : D(D_), lat(D_.lat) {}
    void apply_eo(const Vec& src_e, Vec& dst_o) const { run(SMG_EO_HOP_EO, src_e, dst_o); }
    void apply_oe(const Vec& src_o, Vec& dst_e) const { run(SMG_EO_HOP_OE, src_o, dst_e); }
    void apply_ee_inv(const Vec& src, Vec& dst) const { run(SMG_EO_EE_INV, src, dst); }
    void apply_oo_inv(const Vec& src, Vec& dst) const { run(SMG_EO_OO_INV, src, dst); }
    void apply_oo(const Vec& src, Vec& dst) const { run(SMG_EO_OO, src, dst); }
    void apply_schur(const Vec& src_o, Vec& dst_o) const { run(SMG_EO_SCHUR, src_o, dst_o); }
    void apply_schur_dag(const Vec& src_o, Vec& dst_o) const { run(SMG_EO_SCHUR_DAG, src_o, dst_o); }
    void apply_schur_dag_schur(const Vec& src_o, Vec& dst_o) const { run(SMG_EO_SCHUR_DAG_SCHUR, src_o, dst_o); }
    Vec scatter(const Vec& half_e, const Vec& half_o) const {
        Vec full(lat.ndof, 0.0);
        for (int i = 0; i < lat.V_half; i++) {
            const int se = lat.even_sites[i], so = lat.odd_sites[i];
            full[2 * se] = half_e[2 * i];
            full[2 * se + 1] = half_e[2 * i + 1];
            full[2 * so] = half_o[2 * i];
            full[2 * so + 1] = half_o[2 * i + 1];
        }
        return full;
    }
    Vec gather_even(const Vec& full) const { return gather(full, lat.even_sites); }
    Vec gather_odd(const Vec& full) const { return gather(full, lat.odd_sites); }
    Vec generate_pseudofermion_eo(std::mt19937& rng) const {
        const Vec eta_o = random_vec(2 * lat.V_half, rng);
        Vec phi_o(2 * (size_t)lat.V_half);
        apply_schur_dag(eta_o, phi_o);
        return phi_o;
    }
    Vec reconstruct_full(const Vec& x_o) const {
        Vec out(lat.ndof);
        detail::check(smg_eo_reconstruct(D.handle(), detail::raw(x_o), nullptr, detail::raw(out)));
        return out;
    }
    Vec reconstruct(const Vec& x_o, const Vec& phi) const {
        Vec out(lat.ndof);
        detail::check(smg_eo_reconstruct(D.handle(), detail::raw(x_o), detail::raw(phi), detail::raw(out)));
        return out;
    }
    double log_det_ee() const {
        double out = 0.0;
        detail::check(smg_eo_logdet(D.handle(), &out));
        return out;
    }

   private:
    void run(int kind, const Vec& src, Vec& dst) const {
        detail::fit(dst, 2 * (size_t)lat.V_half);
        detail::check(smg_eo_apply(D.handle(), kind, detail::raw(src), detail::raw(dst)));
    }
    Vec gather(const Vec& full, const std::vector<int>& sites) const {
        Vec half(2 * (size_t)lat.V_half);
        for (int i = 0; i < lat.V_half; i++) {
            half[2 * i] = full[2 * sites[i]];
            half[2 * i + 1] = full[2 * sites[i] + 1];
        }
        return half;
    }
};

// ------------------------------------------------------------------ Prolongator, src/prolongator.h:11-141
struct Prolongator {
    const Lattice& lat;
    int bx, by, k_vec, nbx, nby, nblocks, coarse_dim;
    Prolongator(const Lattice& lat_, int bx_, int by_, int k_)
        : lat(lat_), bx(bx_), by(by_), k_vec(k_), nbx(lat_.L / bx_), nby(lat_.L / by_) {
        if (lat.L % bx != 0 || lat.L % by != 0) throw std::invalid_argument("Prolongator: L % block != 0");
        nblocks = nbx * nby;
        coarse_dim = nblocks * k_vec;
        smg_prolongator* h = nullptr;
        detail::check(smg_prolongator_create(&h, lat.L, lat.L, 2, bx, by, k_vec, 0));
        h_.reset(h, [](smg_prolongator* p) { smg_prolongator_destroy(p); });
    }
    void build_from_vectors(const std::vector<Vec>& null_vecs) {
        if ((int)null_vecs.size() < k_vec) throw std::invalid_argument("build_from_vectors: too few vectors");
        Vec all((size_t)k_vec * lat.ndof);
        for (int i = 0; i < k_vec; i++) std::copy(null_vecs[i].begin(), null_vecs[i].end(), all.begin() + (size_t)i * lat.ndof);
        detail::check(smg_prolongator_build(h_.get(), detail::raw(all)));
    }
    Vec prolong(const Vec& coarse) const {
        Vec fine(lat.ndof);
        detail::check(smg_prolong(h_.get(), detail::raw(coarse), detail::raw(fine)));
        return fine;
    }
    Vec restrict_vec(const Vec& fine) const {
        Vec coarse(coarse_dim);
        detail::check(smg_restrict(h_.get(), detail::raw(fine), detail::raw(coarse)));
        return coarse;
    }
    // block_vecs[b][k] (length 2*bx*by), src/prolongator.h:23
    std::vector<std::vector<Vec>> block_vecs() const {
        const size_t bs = 2 * (size_t)bx * by;
        Vec all((size_t)nblocks * k_vec * bs);
        detail::check(smg_prolongator_get(h_.get(), detail::raw(all)));
        std::vector<std::vector<Vec>> out(nblocks, std::vector<Vec>(k_vec));
        for (int b = 0; b < nblocks; b++)
            for (int k = 0; k < k_vec; k++)
                out[b][k].assign(all.begin() + ((size_t)b * k_vec + k) * bs, all.begin() + ((size_t)b * k_vec + k + 1) * bs);
        return out;
    }
    smg_prolongator* handle() const { return h_.get(); }

   private:
    std::shared_ptr<smg_prolongator> h_;
};

// ------------------------------------------------------------------ coarse operators, src/coarse_op.h
struct SparseCoarseOp;
struct CoarseApply {   // the OpApply of SparseCoarseOp::as_op
    const SparseCoarseOp* C;
    void operator()(const Vec& s, Vec& d) const;
};

// SparseCoarseOp, src/coarse_op.h:154-548 (setup_deflation: the TRLM branch)
struct SparseCoarseOp {
    int nbx = 0, nby = 0, nblocks = 0, k_vec = 0, dim = 0;
    // Galerkin A_c = P^dag A P (any operator A on the prolongator's fine lattice)
    void build(const Prolongator& P, const OpApply& A, int fine_dim);
    void apply_to(const Vec& src, Vec& dst) const {
        dst.assign(dim, cx(0.0, 0.0));
        detail::check(smg_coarse_apply(h_.get(), SMG_F64, detail::raw(src), detail::raw(dst)));
    }
    OpApply as_op() const { return CoarseApply{this}; }
    // CG from x0 = 0, or from the deflated guess once setup_deflation has run
    Vec solve(const Vec& b, int max_iter = 200, double cg_tol = 1e-12) const {
        Vec x(dim);
        int it = 0;
        detail::check(smg_coarse_solve(h_.get(), detail::raw(b), detail::raw(x), max_iter, cg_tol, &it));
        return x;
    }
    // src/coarse_op.cpp:10-54; the deflation space lives on the device, and
    // defl_vecs / defl_vals below are host copies taken when it is set up
    void setup_deflation(int n_ev, int n_kr = 0, int max_restarts = 100, double tol = 1e-10,
                         const std::string& solver = "trlm", double feast_emax = 0.0) {
        (void)feast_emax;
        if (solver != "trlm") throw std::invalid_argument("setup_deflation: only the 'trlm' solver is provided");
        if (dim <= 0 || n_ev <= 0) return;
        std::mt19937 rng(42);
        const Vec start = random_vec(dim, rng);
        smg_trlm_info info{};
        detail::check(smg_coarse_setup_deflation(h_.get(), n_ev, n_kr, max_restarts, tol, detail::raw(start), &info));
        if (!info.converged)
            std::cerr << "SparseCoarseOp::setup_deflation: TRLM did not converge (restarts=" << info.restarts
                      << ")\n";
        pull_deflation();
    }
    // src/coarse_op.h:176-200: keeps the vectors whose fresh Rayleigh quotient is
    // positive and whose relative residual is below res_threshold
    int validate_deflation(double res_threshold = 0.1) {
        int kept = 0;
        for (size_t i = 0; i < defl_vecs.size(); i++) {
            Vec Av;
            apply_to(defl_vecs[i], Av);
            cx vav = 0.0;
            for (int j = 0; j < dim; j++) vav += std::conj(defl_vecs[i][j]) * Av[j];
            const double rq = vav.real();
            double rn = 0.0, an = 0.0;
            for (int j = 0; j < dim; j++) {
                rn += std::norm(Av[j] - rq * defl_vecs[i][j]);
                an += std::norm(Av[j]);
            }
            const double rel = an > 1e-60 ? std::sqrt(rn / an) : 1.0;
            if (rq > 1e-14 && rel < res_threshold) {
                if ((size_t)kept < i) defl_vecs[kept] = std::move(defl_vecs[i]);
                defl_vals[kept] = rq;
                kept++;
            }
        }
        defl_vecs.resize(kept);
        defl_vals.resize(kept);
        push_deflation();
        return kept;
    }
    // after editing defl_vecs / defl_vals on the host, hand them to the device solve
    void push_deflation() {
        const int nd = (int)defl_vals.size();
        Vec all((size_t)nd * dim);
        for (int i = 0; i < nd; i++) std::copy(defl_vecs[i].begin(), defl_vecs[i].end(), all.begin() + (size_t)i * dim);
        detail::check(smg_coarse_set_deflation(h_.get(), nd, nd ? defl_vals.data() : nullptr,
                                               nd ? detail::raw(all) : nullptr));
    }
    std::vector<Vec> defl_vecs;
    std::vector<double> defl_vals;
    // link[b][d][ki*k + kj] flattened, src/coarse_op.h:159-166
    Vec links() const {
        Vec out((size_t)nblocks * 9 * k_vec * k_vec);
        detail::check(smg_coarse_get_links(h_.get(), detail::raw(out)));
        return out;
    }
    smg_coarse* handle() const { return h_.get(); }

   private:
    void pull_deflation() {
        int nd = 0;
        detail::check(smg_coarse_deflation(h_.get(), &nd, nullptr, nullptr));
        defl_vals.assign(nd, 0.0);
        Vec all((size_t)nd * dim);
        if (nd) detail::check(smg_coarse_deflation(h_.get(), &nd, defl_vals.data(), detail::raw(all)));
        defl_vecs.assign(nd, Vec());
        for (int i = 0; i < nd; i++) defl_vecs[i].assign(all.begin() + (size_t)i * dim, all.begin() + (size_t)(i + 1) * dim);
    }
    std::shared_ptr<smg_coarse> h_;
};
inline void CoarseApply::operator()(const Vec& s, Vec& d) const { C->apply_to(s, d); }

// CoarseOp (dense), src/coarse_op.h:18-137: mat holds the columns of A_c
struct CoarseOp {
    int dim = 0;
    std::vector<Vec> mat;
    void build(const DiracOp& D, Prolongator& P) {
        sparse_.build(P, D.as_DdagD_op(), D.lat.ndof);
        dim = sparse_.dim;
        Vec all((size_t)dim * dim);
        detail::check(smg_coarse_dense(sparse_.handle(), detail::raw(all)));
        mat.assign(dim, Vec());
        for (int j = 0; j < dim; j++) mat[j].assign(all.begin() + (size_t)j * dim, all.begin() + (size_t)(j + 1) * dim);
    }
    Vec apply(const Vec& x) const {
        Vec y;
        sparse_.apply_to(x, y);
        return y;
    }
    void apply_to(const Vec& src, Vec& dst) const { dst = apply(src); }
    Vec solve(const Vec& b) const {
        Vec x(dim);
        detail::check(smg_coarse_dense_solve(sparse_.handle(), detail::raw(b), detail::raw(x)));
        return x;
    }

   private:
    SparseCoarseOp sparse_;
};

// ------------------------------------------------------------------ smoothers, src/smoother.h
namespace detail {
inline void smooth(const OpApply& A, int kind, Vec& x, const Vec& b, int niter, double omega, const char* who);
}  // namespace detail
inline void mr_smooth(const DiracOp& D, Vec& x, const Vec& b, int niter) {
    detail::check(smg_smooth_dirac(D.handle(), SMG_SMOOTH_MR, detail::raw(x), detail::raw(b), niter, 0.0));
}
inline void mr_smooth_op(const OpApply& A, Vec& x, const Vec& b, int niter) {
    detail::smooth(A, SMG_SMOOTH_MR, x, b, niter, 0.0, "mr_smooth_op");
}
inline void richardson_smooth_op(const OpApply& A, Vec& x, const Vec& b, int niter, double omega) {
    detail::smooth(A, SMG_SMOOTH_RICHARDSON, x, b, niter, omega, "richardson_smooth_op");
}
// power iteration from mt19937(12345) as src/smoother.cpp:43-59, any operator
inline double estimate_lambda_max(const OpApply& A, int n, int niter = 20);

// ------------------------------------------------------------------ multigrid, src/multigrid.h
struct MGHierarchy;
struct MGPrecond {   // the preconditioner of MGHierarchy::as_precond
    MGHierarchy* mg;
    Vec operator()(const Vec& b) const;
};

struct MGHierarchy {
    bool w_cycle = true;
    std::vector<int> dims;   // per level
    Vec precondition(const Vec& b) {
        Vec x(dims.at(0));
        detail::check(smg_mg_precondition(h_.get(), SMG_F64, detail::raw(b), detail::raw(x)));
        return x;
    }
    // src/multigrid.cpp:146-152: Richardson omega = damping / lambda_max, 30 power
    // iterations from mt19937(12345) per non-coarsest level
    void set_symmetric(double damping = 0.8) {
        Vec starts;
        for (size_t l = 0; l + 1 < dims.size(); l++) {
            std::mt19937 rng_local(12345);
            const Vec v = random_vec(dims[l], rng_local);
            starts.insert(starts.end(), v.begin(), v.end());
        }
        detail::check(smg_mg_set_symmetric(h_.get(), damping, detail::raw(starts)));
    }
    std::function<Vec(const Vec&)> as_precond() { return MGPrecond{this}; }
    // src/multigrid.cpp:111-141; the coarsest level's space (the reference's
    // levels[last].defl_vecs / defl_vals) is copied back into the two members below
    void update_coarsest_deflation(int n_defl, int lobpcg_iters = 3) {
        detail::check(smg_mg_update_coarsest_deflation(h_.get(), n_defl, lobpcg_iters));
        int nd = 0;
        detail::check(smg_mg_coarsest_deflation(h_.get(), &nd, nullptr, nullptr));
        const int dim = dims.back();
        coarsest_defl_vals.assign(nd, 0.0);
        Vec all((size_t)nd * dim);
        if (nd) detail::check(smg_mg_coarsest_deflation(h_.get(), &nd, coarsest_defl_vals.data(), detail::raw(all)));
        coarsest_defl_vecs.assign(nd, Vec());
        for (int i = 0; i < nd; i++)
            coarsest_defl_vecs[i].assign(all.begin() + (size_t)i * dim, all.begin() + (size_t)(i + 1) * dim);
    }
    std::vector<Vec> coarsest_defl_vecs;
    std::vector<double> coarsest_defl_vals;
    // the level-0 near-null vectors of the setup (src/multigrid.h, null_vecs_l0), kept
    // for warm-started rebuilds
    std::vector<Vec> null_vecs_l0;
    // levels[0].Ac.build(D, P0) followed by rebuild_deeper_levels (src/multigrid.cpp:
    // 163-194): call after the hierarchy's DiracOp has taken new links (update_gauge)
    void rebuild_deeper_levels() { detail::check(smg_mg_rebuild(h_.get())); }
    // src/multigrid.cpp:754-794 with solver "trlm"; fine_op is the hierarchy's own
    // operator, whose level-1 Galerkin operator the library already holds
    void setup_sparse_coarse(const OpApply& fine_op, int fine_dim, int n_defl, double cg_tol = 1e-12,
                             int max_cg_iter = 200, const std::string& solver = "trlm", double feast_emax = 0.0) {
        (void)fine_op;
        (void)fine_dim;
        (void)feast_emax;
        if (solver != "trlm") throw std::invalid_argument("setup_sparse_coarse: only the 'trlm' solver is provided");
        std::mt19937 rng(42);
        const Vec start = random_vec(dims.at(1), rng);
        detail::check(smg_mg_setup_sparse_coarse(h_.get(), n_defl, cg_tol, max_cg_iter, detail::raw(start)));
    }
    smg_mg* handle() const { return h_.get(); }
    std::shared_ptr<smg_mg> h_;
};
inline Vec MGPrecond::operator()(const Vec& b) const { return mg->precondition(b); }

// the identity preconditioner, recognised by lobpcg_update and fgmres_solve_generic
// (the reference's callers pass a lambda returning its argument)
struct IdentityPrecond {
    Vec operator()(const Vec& v) const { return v; }
};
inline std::function<Vec(const Vec&)> identity_precond() { return IdentityPrecond{}; }

// ------------------------------------------------------------------ any OpApply / preconditioner
namespace detail {
// A std::function handed to the library as a host callback (smg_op_fn).  An exception
// it throws is kept and rethrown once the library call has returned.
struct CallbackBox {
    const OpApply* op = nullptr;
    const std::function<Vec(const Vec&)>* prec = nullptr;
    size_t n = 0;
    std::exception_ptr err;
};
inline int op_trampoline(void* ctx, const double* src, double* dst) {
    auto* b = static_cast<CallbackBox*>(ctx);
    try {
        const cx* s = reinterpret_cast<const cx*>(src);
        Vec in(s, s + b->n), out(b->n);   // pre-sized: DiracOp's operators write without resizing
        (*b->op)(in, out);
        if (out.size() != b->n) throw std::length_error("OpApply returned a vector of another size");
        std::copy(out.begin(), out.end(), reinterpret_cast<cx*>(dst));
        return 0;
    } catch (...) {
        b->err = std::current_exception();
        return 1;
    }
}
inline int prec_trampoline(void* ctx, const double* src, double* dst) {
    auto* b = static_cast<CallbackBox*>(ctx);
    try {
        const cx* s = reinterpret_cast<const cx*>(src);
        const Vec out = (*b->prec)(Vec(s, s + b->n));
        if (out.size() != b->n) throw std::length_error("the preconditioner returned a vector of another size");
        std::copy(out.begin(), out.end(), reinterpret_cast<cx*>(dst));
        return 0;
    } catch (...) {
        b->err = std::current_exception();
        return 1;
    }
}
// the library's description of A: this header's operators by handle, anything else
// as a callback through box (which must outlive the call)
inline smg_op make_op(const OpApply& A, size_t n, CallbackBox& box) {
    if (!A) throw std::invalid_argument("schwinger_mg: empty OpApply");
    smg_op o{};
    o.n = n;
    if (const DdagDApply* f = A.target<DdagDApply>()) {
        o.dirac = f->D->handle();
    } else if (const CoarseApply* c = A.target<CoarseApply>()) {
        o.coarse = c->C->handle();
    } else {
        box.op = &A;
        box.n = n;
        o.fn = op_trampoline;
        o.ctx = &box;
    }
    return o;
}
inline smg_precond make_prec(const std::function<Vec(const Vec&)>* M, size_t n, CallbackBox& box) {
    smg_precond p{};
    if (!M || !*M || M->target<IdentityPrecond>()) return p;   // none
    if (const MGPrecond* g = M->target<MGPrecond>()) {
        p.mg = g->mg->handle();
        return p;
    }
    box.prec = M;
    box.n = n;
    p.fn = prec_trampoline;
    p.ctx = &box;
    return p;
}
// the status of a library call that may have called back: a callback's own exception
// first, then the library's error
inline void check_cb(int rc, const CallbackBox& a, const CallbackBox& m) {
    if (a.err) std::rethrow_exception(a.err);
    if (m.err) std::rethrow_exception(m.err);
    check(rc);
}
inline void check_size(const Vec& v, int n, const char* who) {
    if (n <= 0 || v.size() != (size_t)n)
        throw std::invalid_argument(std::string(who) + ": the vector's size is not n");
}
}  // namespace detail

inline void SparseCoarseOp::build(const Prolongator& P, const OpApply& A, int fine_dim) {
    if (fine_dim != P.lat.ndof) throw std::invalid_argument("SparseCoarseOp::build: fine_dim mismatch");
    smg_coarse* h = nullptr;
    if (const DdagDApply* f = A.target<DdagDApply>()) {
        detail::check(smg_coarse_create_galerkin(&h, f->D->handle(), P.handle()));
    } else {
        detail::CallbackBox box, none;
        const smg_op op = detail::make_op(A, (size_t)fine_dim, box);
        detail::check_cb(smg_coarse_create_galerkin_op(&h, &op, P.handle()), box, none);
    }
    h_.reset(h, [](smg_coarse* p) { smg_coarse_destroy(p); });
    nbx = P.nbx;
    nby = P.nby;
    nblocks = P.nblocks;
    k_vec = P.k_vec;
    dim = nblocks * k_vec;
}
namespace detail {
inline void smooth(const OpApply& A, int kind, Vec& x, const Vec& b, int niter, double omega, const char* who) {
    check_size(b, (int)x.size(), who);
    CallbackBox box, none;
    const smg_op op = make_op(A, x.size(), box);
    check_cb(smg_smooth_op(&op, kind, raw(x), raw(b), niter, omega), box, none);
}
}  // namespace detail
inline double estimate_lambda_max(const OpApply& A, int n, int niter) {
    std::mt19937 rng_local(12345);
    const Vec v0 = random_vec(n, rng_local);
    double out = 0.0;
    detail::CallbackBox box, none;
    const smg_op op = detail::make_op(A, (size_t)n, box);
    detail::check_cb(smg_lambda_max_op(&op, detail::raw(v0), niter, &out), box, none);
    return out;
}

// build_mg_hierarchy, src/multigrid.cpp:595-687.  The null-space start vectors are
// drawn from rng in the reference's order (k_null vectors per level, level 0
// first); with warm_start, its vectors replace the first level-0 draws, which are
// then not taken from rng (src/multigrid.cpp:443-447).  Levels >= 1 aggregate
// coarse_block_agg consecutive coarse dofs, which on the coarse lattice is an
// (agg/k_null) x 1 block; the coarsest level is solved exactly (dense inverse when
// its dimension is <= 8192, else CG to 1e-12).
inline MGHierarchy build_mg_hierarchy(const DiracOp& D, int n_levels, int block_size, int k_null,
                                      int coarse_block_agg, int null_iters, std::mt19937& rng, bool w_cycle = true,
                                      int pre_smooth = 3, int post_smooth = 3, bool verbose = true,
                                      const std::vector<Vec>* warm_start = nullptr) {
    (void)verbose;
    if (n_levels < 2 || n_levels > 5) throw std::invalid_argument("build_mg_hierarchy: 2..5 levels");
    if (coarse_block_agg <= 0) coarse_block_agg = 4 * k_null;
    smg_mg_params p{};
    p.n_levels = n_levels;
    p.null_iters = null_iters;
    p.pre_smooth = pre_smooth;
    p.post_smooth = post_smooth;
    p.w_cycle = w_cycle ? 1 : 0;
    p.coarse_maxit = 200;
    p.coarse_tol = 1e-12;
    MGHierarchy mg;
    mg.w_cycle = w_cycle;
    mg.dims.push_back(D.lat.ndof);
    int nx = D.lat.L;
    for (int l = 0; l + 1 < n_levels; l++) {
        int bx = block_size, by = block_size;
        if (l > 0) {
            const int prev = mg.dims.back();
            int blk = coarse_block_agg;
            while (prev % blk != 0 && blk > 1) blk--;
            if (blk % k_null != 0 || nx % (blk / k_null) != 0)
                throw std::invalid_argument("build_mg_hierarchy: coarse_block_agg must be a multiple of k_null "
                                            "dividing the coarse lattice row");
            bx = blk / k_null;
            by = 1;
        }
        p.bx[l] = bx;
        p.by[l] = by;
        p.k[l] = k_null;
        if (nx % bx != 0) throw std::invalid_argument("build_mg_hierarchy: block does not divide the lattice");
        const int ny = (l == 0 ? D.lat.L : mg.dims.back() / k_null / nx) / by;
        nx /= bx;
        mg.dims.push_back(nx * ny * k_null);
    }
    p.coarse_solver = mg.dims.back() <= 8192 ? 1 : 0;
    Vec starts;
    for (int l = 0; l + 1 < n_levels; l++)
        for (int i = 0; i < k_null; i++) {
            if (l == 0 && warm_start && i < (int)warm_start->size()) {
                const Vec& w = (*warm_start)[i];
                detail::check_size(w, mg.dims[0], "build_mg_hierarchy (warm_start)");
                starts.insert(starts.end(), w.begin(), w.end());
                continue;
            }
            const Vec v = random_vec(mg.dims[l], rng);
            starts.insert(starts.end(), v.begin(), v.end());
        }
    smg_mg* h = nullptr;
    detail::check(smg_mg_setup(&h, D.handle(), &p, detail::raw(starts)));
    mg.h_.reset(h, [](smg_mg* q) { smg_mg_destroy(q); });
    Vec nv((size_t)k_null * mg.dims[0]);
    detail::check(smg_mg_null_vecs(h, 0, detail::raw(nv)));
    for (int i = 0; i < k_null; i++)
        mg.null_vecs_l0.emplace_back(nv.begin() + (size_t)i * mg.dims[0], nv.begin() + (size_t)(i + 1) * mg.dims[0]);
    return mg;
}

// build_mg_hierarchy_warm, src/multigrid.cpp:688-701
inline MGHierarchy build_mg_hierarchy_warm(const DiracOp& D, int n_levels, int block_size, int k_null,
                                           int coarse_block_agg, int null_iters, std::mt19937& rng,
                                           const std::vector<Vec>& warm_null_vecs, bool w_cycle = true,
                                           int pre_smooth = 3, int post_smooth = 3, bool verbose = true) {
    return build_mg_hierarchy(D, n_levels, block_size, k_null, coarse_block_agg, null_iters, rng, w_cycle, pre_smooth,
                              post_smooth, verbose, &warm_null_vecs);
}

// MGConfig / SolverConfig (src/config.h:16-39) and build_full_mg (src/mg_builder.cpp:
// 3-31): 20 null-space iterations, 5 when warm-started, then the sparse coarse
// operator with TRLM deflation when n_defl > 0.  The FEAST null space is not provided.
struct MGConfig {
    int mg_levels = 1;
    int block_size = 4;
    int k_null = 4;
    int coarse_block = 0;   // 0 = k_null * 4
    bool w_cycle = true;
    int pre_smooth = 3;
    int post_smooth = 3;
    bool no_mg = false;
    bool symmetric_mg = false;
    int resolved_coarse_block() const { return coarse_block > 0 ? coarse_block : k_null * 4; }
};
struct SolverConfig {
    double tol = 1e-10;
    int krylov = 30;
    int max_iter = 300;
    std::string eigensolver = "trlm";
    double feast_emax = 0.0;
    int n_defl = 0;
};
inline MGHierarchy build_full_mg(const DiracOp& D, const MGConfig& mcfg, const SolverConfig& scfg,
                                 std::mt19937& rng, int n_defl = 0, bool verbose = true,
                                 const std::vector<Vec>* warm_start = nullptr) {
    if (scfg.eigensolver == "feast") throw std::invalid_argument("build_full_mg: the FEAST null space is not provided");
    const int null_iters = warm_start ? 5 : 20;
    MGHierarchy mg = build_mg_hierarchy(D, mcfg.mg_levels, mcfg.block_size, mcfg.k_null, mcfg.resolved_coarse_block(),
                                        null_iters, rng, mcfg.w_cycle, mcfg.pre_smooth, mcfg.post_smooth, verbose,
                                        warm_start);
    if (n_defl > 0) mg.setup_sparse_coarse(D.as_DdagD_op(), D.lat.ndof, n_defl, 1e-12, 200, scfg.eigensolver,
                                           scfg.feast_emax);
    return mg;
}

// ------------------------------------------------------------------ even-odd multigrid, src/eo_multigrid.h
// EvenOddMG (src/eo_multigrid.h:229-287) on the Schur operator M^dag M of the odd
// half-lattice: build draws k_vec start vectors from rng (random_vec, as
// compute_near_null_space_generic does) and estimate_lambda_max's mt19937(12345)
// start, then the library builds the near-null vectors, the aggregation of the odd
// sites of bx x by blocks (bx even), the dense Galerkin operator (copied back into
// Ac.mat, column j = P^dag M^dag M P e_j) and omega = 0.8 / lambda_max on the device.
// precondition runs the V-cycle with the caller's A_MdagM as the smoothing operator
// (called on the host, like any OpApply the library does not recognise).
struct EvenOddMG {
    struct Prolongator {   // EvenOddProlongator's shape (src/eo_multigrid.h:26-58)
        int bx, by, nbx, nby, nblocks, k_vec, coarse_dim;
    } P_eo;
    struct CoarseOp {      // EvenOddCoarseOp (src/eo_multigrid.h:158-223): dim and the columns
        int dim = 0;
        std::vector<Vec> mat;
    } Ac;
    int fine_half_dim;
    int pre_smooth, post_smooth;
    double richardson_omega = 0.0;

    EvenOddMG(const Lattice& lat, int bx, int by, int k_vec, int pre = 3, int post = 3)
        : P_eo{bx, by, lat.L / bx, lat.L / by, (lat.L / bx) * (lat.L / by), k_vec, (lat.L / bx) * (lat.L / by) * k_vec},
          fine_half_dim(2 * lat.V_half), pre_smooth(pre), post_smooth(post) {}

    void build(const DiracOp& D, int null_iters, std::mt19937& rng) {
        Vec starts;
        for (int i = 0; i < P_eo.k_vec; i++) {
            const Vec v = random_vec(fine_half_dim, rng);
            starts.insert(starts.end(), v.begin(), v.end());
        }
        std::mt19937 rng_local(12345);
        const Vec lam = random_vec(fine_half_dim, rng_local);
        smg_eo_mg* h = nullptr;
        detail::check(smg_eo_mg_setup(&h, D.handle(), P_eo.bx, P_eo.by, P_eo.k_vec, pre_smooth, post_smooth,
                                      null_iters, detail::raw(starts), detail::raw(lam)));
        h_.reset(h, [](smg_eo_mg* q) { smg_eo_mg_destroy(q); });
        double lmax = 0.0;
        detail::check(smg_eo_mg_info(h, &richardson_omega, &lmax, &Ac.dim));
        Vec m((size_t)Ac.dim * Ac.dim);
        detail::check(smg_eo_mg_coarse_matrix(h, detail::raw(m)));
        Ac.mat.assign(Ac.dim, Vec(Ac.dim));
        for (int i = 0; i < Ac.dim; i++)
            for (int j = 0; j < Ac.dim; j++) Ac.mat[j][i] = m[(size_t)i * Ac.dim + j];
        std::cout << "  E/O MG: fine_half_dim=" << fine_half_dim << " coarse_dim=" << Ac.dim
                  << " lambda_max=" << lmax << "\n";
    }

    Vec precondition(const Vec& b, const OpApply& A_MdagM) const {
        if (!h_) throw std::logic_error("EvenOddMG::precondition before build");
        detail::check_size(b, fine_half_dim, "EvenOddMG::precondition");
        Vec x(fine_half_dim);
        detail::CallbackBox box, none;
        const smg_op op = detail::make_op(A_MdagM, (size_t)fine_half_dim, box);
        detail::check_cb(smg_eo_mg_precondition(h_.get(), &op, detail::raw(b), detail::raw(x)), box, none);
        return x;
    }
    std::shared_ptr<smg_eo_mg> h_;
};

// ------------------------------------------------------------------ Krylov solvers, src/solvers.h
struct CGResult {
    Vec solution;
    int iterations;
    double final_residual;
};
using FCGResult = CGResult;
struct RitzPair {
    double value;
    Vec vector;
};
struct FGMRESResult {
    Vec solution;
    int iterations;
    double final_residual;
    std::vector<RitzPair> ritz_pairs;
};


namespace detail {
inline const DiracOp& dirac_of(const OpApply& A, const char* who) {
    const DdagDApply* f = A.target<DdagDApply>();
    if (!f) throw std::invalid_argument(std::string(who) + ": A must be DiracOp::as_DdagD_op()");
    return *f->D;
}
inline MGHierarchy* mg_of(const std::function<Vec(const Vec&)>& M, const char* who) {
    const MGPrecond* p = M.target<MGPrecond>();
    if (!p) throw std::invalid_argument(std::string(who) + ": the preconditioner must be MGHierarchy::as_precond()");
    return p->mg;
}
inline CGResult cg(const OpApply& A, int n, const Vec& rhs, const Vec* x0,
                   const std::function<Vec(const Vec&)>* precond, int max_iter, double tol, const char* who) {
    check_size(rhs, n, who);
    if (x0) check_size(*x0, n, who);
    CallbackBox ba, bm;
    const smg_op op = make_op(A, (size_t)n, ba);
    const smg_precond pc = make_prec(precond, (size_t)n, bm);
    CGResult r{Vec(n), 0, 0.0};
    smg_solve_info info{};
    check_cb(smg_cg_solve_op(&op, &pc, raw(rhs), x0 ? raw(*x0) : nullptr, raw(r.solution), max_iter, tol, &info),
             ba, bm);
    r.iterations = info.iterations;
    r.final_residual = info.final_residual;
    return r;
}
}  // namespace detail

// cg_solve family, src/solvers.cpp:16-96 (any operator and preconditioner)
inline CGResult cg_solve(const OpApply& A, int n, const Vec& rhs, int max_iter, double tol) {
    return detail::cg(A, n, rhs, nullptr, nullptr, max_iter, tol, "cg_solve");
}
inline CGResult cg_solve_x0(const OpApply& A, int n, const Vec& rhs, const Vec& x0, int max_iter, double tol) {
    return detail::cg(A, n, rhs, &x0, nullptr, max_iter, tol, "cg_solve_x0");
}
inline CGResult cg_solve_precond(const OpApply& A, int n, const Vec& rhs,
                                 const std::function<Vec(const Vec&)>& precond, int max_iter, double tol) {
    return detail::cg(A, n, rhs, nullptr, &precond, max_iter, tol, "cg_solve_precond");
}
inline CGResult cg_solve_x0_precond(const OpApply& A, int n, const Vec& rhs, const Vec& x0,
                                    const std::function<Vec(const Vec&)>& precond, int max_iter, double tol) {
    return detail::cg(A, n, rhs, &x0, &precond, max_iter, tol, "cg_solve_x0_precond");
}

enum class Precision { FP64 = SMG_F64, Mixed = SMG_F32 };

// fcg_solve / fcg_solve_mg, src/solvers.cpp:223-300,803-811 (the reference's
// projection coefficient, src/solvers.cpp:261, FP64)
inline FCGResult fcg_solve(const OpApply& A, int n, const Vec& rhs, const std::function<Vec(const Vec&)>& precond,
                           int max_iter, double tol, int trunc = 30) {
    detail::check_size(rhs, n, "fcg_solve");
    detail::CallbackBox ba, bm;
    const smg_op op = detail::make_op(A, (size_t)n, ba);
    const smg_precond pc = detail::make_prec(&precond, (size_t)n, bm);
    FCGResult r{Vec(n), 0, 0.0};
    smg_solve_info info{};
    detail::check_cb(smg_gcr_solve_op(&op, &pc, detail::raw(rhs), detail::raw(r.solution), max_iter, tol, trunc, 1,
                                      &info),
                     ba, bm);
    r.iterations = info.iterations;
    r.final_residual = info.final_residual;
    return r;
}
inline FCGResult fcg_solve_mg(const OpApply& A, int n, const Vec& rhs, MGHierarchy& mg, int max_iter, double tol,
                              int truncation = 20) {
    return fcg_solve(A, n, rhs, mg.as_precond(), max_iter, tol, truncation);
}
// GCR with the textbook projection, FP64 or mixed precision (FP64 refinement of
// the true residual around FP32 GCR and an FP32 cycle); final_residual is the
// FP64 true residual.  No reference counterpart.
inline FCGResult gcr_solve_mg(const OpApply& A, int n, const Vec& rhs, MGHierarchy& mg, int max_iter, double tol,
                              int trunc = 10, Precision prec = Precision::FP64, double inner_tol = 1e-5) {
    (void)n;
    const DiracOp& D = detail::dirac_of(A, "gcr_solve_mg");
    FCGResult r{Vec(D.lat.ndof), 0, 0.0};
    smg_solve_info info{};
    detail::check(smg_gcr_solve(D.handle(), mg.handle(), (int)prec, detail::raw(rhs), detail::raw(r.solution),
                                max_iter, tol, trunc, 0, inner_tol, &info));
    r.iterations = info.iterations;
    r.final_residual = info.true_residual;
    return r;
}
// fgmres_solve_generic / fgmres_solve_mg, src/solvers.cpp:555-811, with the Ritz
// harvest and intra-solve deflation of n_ritz_harvest pairs; precond:
// MGHierarchy::as_precond(), identity_precond() or an empty std::function
inline FGMRESResult fgmres_solve_generic(const OpApply& A, int n, const Vec& rhs,
                                         const std::function<Vec(const Vec&)>& precond, int restart, int max_iter,
                                         double tol, int n_ritz_harvest = 4) {
    detail::check_size(rhs, n, "fgmres_solve_generic");
    detail::CallbackBox ba, bm;
    const smg_op op = detail::make_op(A, (size_t)n, ba);
    const smg_precond pc = detail::make_prec(&precond, (size_t)n, bm);
    const int k = n_ritz_harvest > 0 ? n_ritz_harvest : 0;
    const size_t nd = (size_t)n;
    FGMRESResult r{Vec(nd), 0, 0.0, {}};
    smg_solve_info info{};
    std::vector<double> vals(2 * (size_t)std::max(k, 1));
    Vec vecs(2 * (size_t)std::max(k, 1) * nd);
    int npairs = 0;
    detail::check_cb(smg_fgmres_solve_op(&op, &pc, detail::raw(rhs), detail::raw(r.solution), restart, max_iter, tol,
                                         k, &info, &npairs, vals.data(), detail::raw(vecs)),
                     ba, bm);
    r.iterations = info.iterations;
    r.final_residual = info.final_residual;
    for (int q = 0; q < npairs; q++)
        r.ritz_pairs.push_back({vals[q], Vec(vecs.begin() + (ptrdiff_t)(q * nd), vecs.begin() + (ptrdiff_t)((q + 1) * nd))});
    return r;
}
inline FGMRESResult fgmres_solve_mg(const OpApply& A, int n, const Vec& rhs, MGHierarchy& mg, int restart,
                                    int max_iter, double tol, int n_ritz_harvest = 4) {
    return fgmres_solve_generic(A, n, rhs, mg.as_precond(), restart, max_iter, tol, n_ritz_harvest);
}

// ------------------------------------------------------------------ eigensolvers, src/eigensolver.h
struct TRLMResult {   // src/eigensolver.h:19-25
    std::vector<Vec> eigvecs;
    std::vector<double> eigvals;
    int iterations;
    int num_restarts;
    bool converged;
};
struct TRLMState;   // saving the Krylov space (state_out) is not provided by this library

// lanczos_eigen, src/eigensolver_core.cpp:103-265: A_cols[j][i] = A_ij; ascending
// evals; evecs[row][col] = component row of eigenvector col
inline void lanczos_eigen(std::vector<Vec>& A_cols, int n, RVec& evals, std::vector<Vec>& evecs) {
    Vec a((size_t)n * n), v((size_t)n * n);
    for (int j = 0; j < n; j++) std::copy(A_cols[j].begin(), A_cols[j].begin() + n, a.begin() + (size_t)j * n);
    evals.assign(n, 0.0);
    detail::check(smg_dense_eigh(n, detail::raw(a), evals.data(), detail::raw(v)));
    evecs.assign(n, Vec(n));
    for (int r = 0; r < n; r++) std::copy(v.begin() + (size_t)r * n, v.begin() + (size_t)(r + 1) * n, evecs[r].begin());
}

// trlm_eigensolver, src/eigensolver_core.cpp:691-1057, on DiracOp::as_DdagD_op() or
// SparseCoarseOp::as_op(); the Chebyshev filter (poly_deg > 0) and state_out are not
// provided
inline TRLMResult trlm_eigensolver(const OpApply& A, int n, int n_ev, int n_kr = 0, int max_restarts = 100,
                                   double tol = 1e-10, int poly_deg = 0, double a_min = 0.0, double a_max = 0.0,
                                   TRLMState* state_out = nullptr, const Vec* start_vec = nullptr) {
    (void)a_min;
    (void)a_max;
    if (poly_deg > 0) throw std::invalid_argument("trlm_eigensolver: the Chebyshev filter is not provided");
    if (state_out) throw std::invalid_argument("trlm_eigensolver: state_out is not provided");
    Vec start;
    if (start_vec && (int)start_vec->size() == n) {
        start = *start_vec;
    } else {
        std::mt19937 rng(42);
        start = random_vec(n, rng);
    }
    TRLMResult r{std::vector<Vec>(n_ev, Vec(n)), std::vector<double>(n_ev, 0.0), 0, 0, false};
    Vec all((size_t)n_ev * n);
    smg_trlm_info info{};
    if (const DdagDApply* f = A.target<DdagDApply>()) {
        if (n != f->D->lat.ndof) throw std::invalid_argument("trlm_eigensolver: n mismatch");
        detail::check(smg_dirac_trlm(f->D->handle(), SMG_OP_DDAGD, n_ev, n_kr, max_restarts, tol, detail::raw(start),
                                     r.eigvals.data(), detail::raw(all), &info));
    } else if (const CoarseApply* c = A.target<CoarseApply>()) {
        if (n != c->C->dim) throw std::invalid_argument("trlm_eigensolver: n mismatch");
        detail::check(smg_coarse_trlm(c->C->handle(), n_ev, n_kr, max_restarts, tol, detail::raw(start),
                                      r.eigvals.data(), detail::raw(all), &info));
    } else {
        throw std::invalid_argument("trlm_eigensolver: A must be DiracOp::as_DdagD_op() or SparseCoarseOp::as_op()");
    }
    for (int i = 0; i < n_ev; i++) r.eigvecs[i].assign(all.begin() + (size_t)i * n, all.begin() + (size_t)(i + 1) * n);
    r.iterations = info.iterations;
    r.num_restarts = info.restarts;
    r.converged = info.converged != 0;
    return r;
}

struct LOBPCGResult {   // src/eigensolver.h:6-10
    std::vector<Vec> eigvecs;
    std::vector<double> eigvals;
    int iterations;
};
// lobpcg_update, src/eigensolver_core.cpp:279-533, on DiracOp::as_DdagD_op() (precond:
// MGHierarchy::as_precond(), identity_precond() or an empty std::function) or on
// SparseCoarseOp::as_op() (identity only); k <= 32, run as one block on the device
inline LOBPCGResult lobpcg_update(const OpApply& A, int n, int k, const std::vector<Vec>& X0,
                                  const std::function<Vec(const Vec&)>& precond, int max_iter = 5,
                                  double tol = 1e-8) {
    const bool ident = !precond || precond.target<IdentityPrecond>() != nullptr;
    Vec x0((size_t)k * n), x((size_t)k * n);
    for (int i = 0; i < k; i++) std::copy(X0[i].begin(), X0[i].begin() + n, x0.begin() + (size_t)i * n);
    LOBPCGResult r{std::vector<Vec>(k), std::vector<double>(k, 0.0), 0};
    if (const DdagDApply* f = A.target<DdagDApply>()) {
        if (n != f->D->lat.ndof) throw std::invalid_argument("lobpcg_update: n mismatch");
        MGHierarchy* mg = ident ? nullptr : detail::mg_of(precond, "lobpcg_update");
        detail::check(smg_dirac_lobpcg(f->D->handle(), mg ? mg->handle() : nullptr, k, detail::raw(x0), max_iter,
                                       tol, r.eigvals.data(), detail::raw(x), &r.iterations));
    } else if (const CoarseApply* c = A.target<CoarseApply>()) {
        if (n != c->C->dim) throw std::invalid_argument("lobpcg_update: n mismatch");
        if (!ident) throw std::invalid_argument("lobpcg_update: the coarse operator takes the identity preconditioner");
        detail::check(smg_coarse_lobpcg(c->C->handle(), k, detail::raw(x0), max_iter, tol, r.eigvals.data(),
                                        detail::raw(x), &r.iterations));
    } else {
        throw std::invalid_argument("lobpcg_update: A must be DiracOp::as_DdagD_op() or SparseCoarseOp::as_op()");
    }
    for (int i = 0; i < k; i++) r.eigvecs[i].assign(x.begin() + (size_t)i * n, x.begin() + (size_t)(i + 1) * n);
    return r;
}

// cg_solve_deflated, src/solvers.cpp:111-213 (precond: MGHierarchy::as_precond() or
// the identity, an empty std::function); defl_vals is not read, as in the reference
inline CGResult cg_solve_deflated(const OpApply& A, int n, const Vec& rhs,
                                  const std::function<Vec(const Vec&)>& precond, const std::vector<Vec>& defl_vecs,
                                  const std::vector<double>& defl_vals, int max_iter, double tol) {
    (void)defl_vals;
    const DiracOp& D = detail::dirac_of(A, "cg_solve_deflated");
    MGHierarchy* mg = precond ? detail::mg_of(precond, "cg_solve_deflated") : nullptr;
    const int nd = (int)defl_vecs.size();
    Vec U((size_t)nd * n);
    for (int i = 0; i < nd; i++) std::copy(defl_vecs[i].begin(), defl_vecs[i].begin() + n, U.begin() + (size_t)i * n);
    CGResult r{Vec(n), 0, 0.0};
    smg_solve_info info{};
    detail::check(smg_cg_solve_deflated(D.handle(), SMG_OP_DDAGD, mg ? mg->handle() : nullptr, detail::raw(rhs), nd,
                                        nd ? detail::raw(U) : nullptr, detail::raw(r.solution), max_iter, tol, &info));
    r.iterations = info.iterations;
    r.final_residual = info.final_residual;
    return r;
}

// ------------------------------------------------------------------ actions and forces, src/hmc.h
namespace detail {
inline void to_force(const RVec& F, int V, std::array<RVec, 2>& force) {
    force[0].assign(F.begin(), F.begin() + V);
    force[1].assign(F.begin() + V, F.begin() + 2 * (size_t)V);
}
}  // namespace detail
inline double gauge_action(const GaugeField& g, double beta) {
    double out = 0.0;
    const Vec u = g.links();
    detail::check(smg_gauge_action(g.lat.L, detail::raw(u), beta, &out));
    return out;
}
inline void gauge_force(const GaugeField& g, double beta, std::array<RVec, 2>& force) {
    RVec F(2 * (size_t)g.lat.V);
    const Vec u = g.links();
    detail::check(smg_gauge_force(g.lat.L, detail::raw(u), beta, F.data()));
    detail::to_force(F, g.lat.V, force);
}
struct FermionActionResult {
    double action;
    Vec X;
    int cg_iters;
};
inline FermionActionResult fermion_action(const DiracOp& D, const Vec& phi, int max_iter, double tol) {
    FermionActionResult r{0.0, Vec(D.lat.ndof), 0};
    detail::check(smg_fermion_action(D.handle(), detail::raw(phi), max_iter, tol, &r.action, detail::raw(r.X),
                                     &r.cg_iters));
    return r;
}
// F = -dS_f/dtheta, src/hmc_forces.cpp:134-139 (chi = D X kept on the device)
inline void fermion_force(const DiracOp& D, const Vec& X, std::array<RVec, 2>& force) {
    RVec F(2 * (size_t)D.lat.V);
    detail::check(smg_fermion_force(D.handle(), detail::raw(X), F.data()));
    detail::to_force(F, D.lat.V, force);
}
inline void fermion_force_bilinear(const DiracOp& D, const Vec& chi, const Vec& X, std::array<RVec, 2>& force) {
    RVec F(2 * (size_t)D.lat.V);
    detail::check(smg_force_backward(D.handle(), detail::raw(chi), detail::raw(X), F.data()));
    detail::to_force(F, D.lat.V, force);
}
inline void hopping_force_bilinear(const DiracOp& D, const Vec& chi, const Vec& X, std::array<RVec, 2>& force) {
    RVec F(2 * (size_t)D.lat.V);
    detail::check(smg_hopping_force(D.handle(), detail::raw(chi), detail::raw(X), F.data()));
    detail::to_force(F, D.lat.V, force);
}
inline void logdet_ee_force(const DiracOp& D, std::array<RVec, 2>& force) {
    RVec F(2 * (size_t)D.lat.V);
    detail::check(smg_logdet_force(D.handle(), F.data()));
    detail::to_force(F, D.lat.V, force);
}
inline void eo_fermion_force(const DiracOp& D, const EvenOddDiracOp& eoD, const Vec& x_o, const Vec& y_o,
                             std::array<RVec, 2>& force) {
    (void)eoD;
    RVec F(2 * (size_t)D.lat.V);
    detail::check(smg_eo_fermion_force(D.handle(), detail::raw(x_o), detail::raw(y_o), F.data()));
    detail::to_force(F, D.lat.V, force);
}
// generate_pseudofermion, src/hmc_trajectory.cpp: phi = D^dag eta, eta from rng
inline void generate_pseudofermion(const DiracOp& D, std::mt19937& rng, Vec& phi) {
    const Vec eta = random_vec(D.lat.ndof, rng);
    D.apply_dag(eta, phi);
}

// ------------------------------------------------------------------ HMC, src/hmc.h, src/hmc_trajectory.cpp
// MomentumField, src/hmc.h:31-44 (randomise draws pi[0] then pi[1] from rng)
struct MomentumField {
    const Lattice& lat;
    std::array<RVec, 2> pi;
    explicit MomentumField(const Lattice& lat_) : lat(lat_) {
        pi[0].assign(lat.V, 0.0);
        pi[1].assign(lat.V, 0.0);
    }
    void randomise(std::mt19937& rng) {
        std::normal_distribution<double> dist(0.0, 1.0);
        for (int mu = 0; mu < 2; mu++)
            for (int s = 0; s < lat.V; s++) pi[mu][s] = dist(rng);
    }
    double kinetic_energy() const {
        double ke = 0.0;
        for (int mu = 0; mu < 2; mu++)
            for (int s = 0; s < lat.V; s++) ke += pi[mu][s] * pi[mu][s];
        return 0.5 * ke;
    }
};
struct HMCParams {   // src/hmc.h:107-119
    double beta;
    double tau;
    int n_steps;
    int cg_maxiter;
    double cg_tol;
    bool use_mg;
    double c_sw = 0.0;
    double mu_t = 0.0;
    bool use_eo = false;
    bool omelyan = false;
    bool force_accept = false;
};
struct HMCResult {   // src/hmc.h:120-128
    bool accepted;
    double dH, dKE, dSG, dSF, dLD;
    int total_cg_iters;
};
struct ReversibilityResult {   // src/hmc.h:402-408
    double gauge_delta, mom_delta, dH_forward, dH_backward;
    int total_cg;
};
struct TrackingState;   // eigenspace tracking is not provided by this library

namespace detail {
inline smg_hmc_params hmc_c(const HMCParams& p) {
    smg_hmc_params c{};
    c.beta = p.beta;
    c.tau = p.tau;
    c.n_steps = p.n_steps;
    c.omelyan = p.omelyan ? 1 : 0;
    c.cg_maxiter = p.cg_maxiter;
    c.cg_tol = p.cg_tol;
    c.force_accept = p.force_accept ? 1 : 0;
    c.use_eo = p.use_eo ? 1 : 0;
    return c;
}
inline RVec mom_flat(const MomentumField& m) {
    RVec p(2 * (size_t)m.lat.V);
    std::copy(m.pi[0].begin(), m.pi[0].end(), p.begin());
    std::copy(m.pi[1].begin(), m.pi[1].end(), p.begin() + m.lat.V);
    return p;
}
inline void mom_unflat(const RVec& p, MomentumField& m) {
    m.pi[0].assign(p.begin(), p.begin() + m.lat.V);
    m.pi[1].assign(p.begin() + m.lat.V, p.end());
}
// {KE, S_g, S_f, LD} on D's gauge field; returns the CG iterations
inline int hamiltonian(const DiracOp& D, bool eo, double beta, const MomentumField& mom, const Vec& phi, int maxit,
                       double tol, double* h) {
    const RVec p = mom_flat(mom);
    int it = 0;
    check(smg_hmc_hamiltonian(D.handle(), eo ? 1 : 0, beta, p.data(), raw(phi), maxit, tol, h, &it));
    return it;
}
inline void no_extras(const std::function<Vec(const Vec&)>* precond, TrackingState* tracking, const char* who) {
    if (precond || tracking)
        throw std::invalid_argument(std::string(who) + ": preconditioned or tracked MD is not supported");
}
}  // namespace detail

namespace detail {
// the MD evolution on the GPU (eo: the even-odd trajectory's leapfrog, phi = phi_o);
// gauge and mom are updated in place; returns the CG iterations
inline int md_evolve(GaugeField& gauge, MomentumField& mom, const Lattice& lat, double mass, double wilson_r,
                     const Vec& phi, const HMCParams& params, bool eo) {
    DiracOp D(lat, gauge, mass, wilson_r, params.c_sw, params.mu_t);
    smg_hmc_params p = hmc_c(params);
    p.use_eo = eo ? 1 : 0;
    RVec pi = mom_flat(mom);
    int it = 0;
    check(smg_hmc_evolve(D.handle(), &p, pi.data(), raw(phi), &it));
    Vec u(2 * (size_t)lat.V);
    check(smg_dirac_get_gauge(D.handle(), raw(u)));
    gauge.U[0].assign(u.begin(), u.begin() + lat.V);
    gauge.U[1].assign(u.begin() + lat.V, u.end());
    mom_unflat(pi, mom);
    return it;
}
}  // namespace detail

// plain_leapfrog_evolve, src/hmc_trajectory.cpp:243-409: gauge and mom evolve in place
// (on the GPU, copied back); returns the CG iterations
inline int plain_leapfrog_evolve(GaugeField& gauge, MomentumField& mom, const Lattice& lat, double mass,
                                 double wilson_r, const Vec& phi, const HMCParams& params,
                                 const std::function<Vec(const Vec&)>* precond = nullptr,
                                 TrackingState* tracking = nullptr) {
    detail::no_extras(precond, tracking, "plain_leapfrog_evolve");
    return detail::md_evolve(gauge, mom, lat, mass, wilson_r, phi, params, false);
}

// hmc_trajectory, src/hmc_trajectory.cpp:38-236: momenta, then the pseudofermion's
// Gaussians (phi = D^dag eta, or phi_o = M^dag eta_o with use_eo), then (only if
// needed) the Metropolis uniform are drawn from rng in the reference's order
inline HMCResult hmc_trajectory(GaugeField& gauge, const Lattice& lat, double mass, double wilson_r,
                                const HMCParams& params, std::mt19937& rng,
                                const std::function<Vec(const Vec&)>* precond = nullptr,
                                TrackingState* tracking = nullptr) {
    detail::no_extras(precond, tracking, "hmc_trajectory");
    const std::array<Vec, 2> old = gauge.U;
    MomentumField mom(lat);
    mom.randomise(rng);
    Vec phi;
    double h0[4], h1[4];
    int cg = 0;
    {
        DiracOp D0(lat, gauge, mass, wilson_r, params.c_sw, params.mu_t);
        if (params.use_eo) phi = EvenOddDiracOp(D0).generate_pseudofermion_eo(rng);
        else generate_pseudofermion(D0, rng, phi);
        cg += detail::hamiltonian(D0, params.use_eo, params.beta, mom, phi, params.cg_maxiter, params.cg_tol, h0);
    }
    cg += detail::md_evolve(gauge, mom, lat, mass, wilson_r, phi, params, params.use_eo);
    {
        DiracOp D1(lat, gauge, mass, wilson_r, params.c_sw, params.mu_t);
        cg += detail::hamiltonian(D1, params.use_eo, params.beta, mom, phi, params.cg_maxiter, params.cg_tol, h1);
    }
    const double dH = (h1[0] + h1[1] + h1[2] + h1[3]) - (h0[0] + h0[1] + h0[2] + h0[3]);
    std::uniform_real_distribution<double> uniform(0.0, 1.0);
    const bool accept = params.force_accept || (dH < 0) || (uniform(rng) < std::exp(-dH));
    if (!accept) gauge.U = old;
    return {accept, dH, h1[0] - h0[0], h1[1] - h0[1], h1[2] - h0[2], h1[3] - h0[3], cg};
}

// ------------------------------------------------------------------ U(1) heatbath (an addition: the reference has none)
// n_sweeps checkerboard heatbath sweeps of the Wilson action on gauge.U (L even), Philox
// keyed by seed, sweeps numbered from first_sweep (see smg_heatbath)
inline void heatbath(GaugeField& gauge, double beta, int n_sweeps, unsigned long long seed, unsigned first_sweep = 0) {
    Vec u = gauge.links();
    detail::check(smg_heatbath(gauge.lat.L, detail::raw(u), beta, n_sweeps, seed, first_sweep));
    gauge.U[0].assign(u.begin(), u.begin() + gauge.lat.V);
    gauge.U[1].assign(u.begin() + gauge.lat.V, u.end());
}

// ------------------------------------------------------------------ multi-timescale MG HMC, src/hmc_multiscale.cpp
// CoarseDeflState, src/hmc.h:308-324
struct CoarseDeflState {
    std::vector<Vec> eigvecs;
    std::vector<double> eigvals;
    Vec lowmode_solution(const Vec& phi, const Prolongator& P) const {
        Vec pc = P.restrict_vec(phi);
        const int nc = (int)pc.size();
        Vec xc(nc, cx(0.0, 0.0));
        for (size_t i = 0; i < eigvecs.size(); i++)
            if (eigvals[i] > 1e-14) {
                cx c = 0.0;
                for (int t = 0; t < nc; t++) c += std::conj(eigvecs[i][t]) * pc[t];
                c /= eigvals[i];
                for (int t = 0; t < nc; t++) xc[t] += c * eigvecs[i][t];
            }
        return P.prolong(xc);
    }
};
enum class OuterIntegrator { Leapfrog, Omelyan, FGI, FGI_QPQPQ };   // src/hmc.h:326-335
struct MGMultiScaleParams {   // src/hmc.h:346-365
    double beta;
    double tau = 1.0;
    int n_outer = 4;
    int n_inner = 5;
    int cg_maxiter = 500;
    double cg_tol = 1e-10;
    int inner_smooth = 3;
    OuterIntegrator outer_type = OuterIntegrator::Leapfrog;
    int defl_refresh = 0;
    int mg_perturb_freq = 0;
    double c_sw = 0.0;
    double mu_t = 0.0;
    int inner_tracking = 0;
    bool use_eo = false;
};
struct MGMultiScaleResult {   // src/hmc.h:367-376
    bool accepted;
    double dH;
    double dKE, dSG, dSF;
    int highmode_cg_iters;
    int lowmode_force_evals;
    double highmode_time;
    double lowmode_time;
};
struct EigenForecastState;   // forecasting is not provided by this library

namespace detail {
// an smg_coarse handle on P's coarse space holding the deflation state (its operator
// is not used)
inline std::shared_ptr<smg_coarse> defl_handle(const Prolongator& P, const CoarseDeflState& cd) {
    Vec zero((size_t)P.nblocks * 9 * P.k_vec * P.k_vec, cx(0.0, 0.0));
    smg_coarse* h = nullptr;
    check(smg_coarse_create_links(&h, P.nbx, P.nby, P.k_vec, raw(zero)));
    std::shared_ptr<smg_coarse> C(h, [](smg_coarse* p) { smg_coarse_destroy(p); });
    const int nd = (int)cd.eigvals.size();
    Vec all((size_t)nd * P.coarse_dim);
    for (int i = 0; i < nd; i++) std::copy(cd.eigvecs[i].begin(), cd.eigvecs[i].end(), all.begin() + (size_t)i * P.coarse_dim);
    check(smg_coarse_set_deflation(h, nd, nd ? cd.eigvals.data() : nullptr, nd ? raw(all) : nullptr));
    return C;
}
inline void defl_read(smg_coarse* C, int dim, CoarseDeflState& cd) {
    int nd = 0;
    check(smg_coarse_deflation(C, &nd, nullptr, nullptr));
    cd.eigvals.assign(nd, 0.0);
    Vec all((size_t)nd * dim);
    if (nd) check(smg_coarse_deflation(C, &nd, cd.eigvals.data(), raw(all)));
    cd.eigvecs.assign(nd, Vec());
    for (int i = 0; i < nd; i++) cd.eigvecs[i].assign(all.begin() + (size_t)i * dim, all.begin() + (size_t)(i + 1) * dim);
}
}  // namespace detail

// coarse_lowmode_force, src/hmc_multiscale.cpp:237-276
inline void coarse_lowmode_force(const DiracOp& D, const CoarseDeflState& cdefl, const Prolongator& P, const Vec& phi,
                                 std::array<RVec, 2>& force, int smooth_iters = 0) {
    auto C = detail::defl_handle(P, cdefl);
    RVec F(2 * (size_t)D.lat.V);
    detail::check(smg_coarse_lowmode_force(D.handle(), P.handle(), C.get(), detail::raw(phi), smooth_iters, F.data()));
    detail::to_force(F, D.lat.V, force);
}
// evolve_coarse_deflation without forecast, src/hmc_multiscale.cpp:278-338
inline void evolve_coarse_deflation(CoarseDeflState& cdefl, const SparseCoarseOp& Ac_new,
                                    EigenForecastState* forecast = nullptr) {
    if (forecast) throw std::invalid_argument("evolve_coarse_deflation: forecasting is not provided");
    if (cdefl.eigvecs.empty()) return;
    const int nd = (int)cdefl.eigvals.size();
    Vec zero((size_t)Ac_new.nblocks * 9 * Ac_new.k_vec * Ac_new.k_vec, cx(0.0, 0.0));
    smg_coarse* h = nullptr;
    detail::check(smg_coarse_create_links(&h, Ac_new.nbx, Ac_new.nby, Ac_new.k_vec, detail::raw(zero)));
    std::shared_ptr<smg_coarse> C(h, [](smg_coarse* p) { smg_coarse_destroy(p); });
    Vec all((size_t)nd * Ac_new.dim);
    for (int i = 0; i < nd; i++) std::copy(cdefl.eigvecs[i].begin(), cdefl.eigvecs[i].end(), all.begin() + (size_t)i * Ac_new.dim);
    detail::check(smg_coarse_set_deflation(h, nd, cdefl.eigvals.data(), detail::raw(all)));
    detail::check(smg_coarse_evolve_deflation(h, Ac_new.handle()));
    detail::defl_read(h, Ac_new.dim, cdefl);
}

// hmc_trajectory_mg_multiscale, src/hmc_multiscale.cpp:340-888: leapfrog or Omelyan outer
// integrator; mg_precond is MGHierarchy::as_precond() (its FP64 cycle preconditions the
// outer solves), identity_precond() or empty.  The reference's rng order is kept: the
// momenta, two pseudofermions (the second is used), then the Metropolis uniform when
// dH >= 0.  Forecasting, pre_solve, prolongator perturbation, tracking and the FGI
// integrators are not provided.
inline MGMultiScaleResult hmc_trajectory_mg_multiscale(GaugeField& gauge, const Lattice& lat, double mass,
                                                       double wilson_r, const MGMultiScaleParams& params,
                                                       CoarseDeflState& cdefl, Prolongator& P,
                                                       std::function<Vec(const Vec&)>& mg_precond, std::mt19937& rng,
                                                       EigenForecastState* forecast = nullptr,
                                                       const std::function<void()>* pre_solve = nullptr,
                                                       MGHierarchy* mg_hierarchy = nullptr,
                                                       TrackingState* tracking = nullptr) {
    if (forecast || pre_solve || mg_hierarchy || tracking || params.mg_perturb_freq || params.inner_tracking ||
        params.use_eo)
        throw std::invalid_argument("hmc_trajectory_mg_multiscale: forecasting, pre_solve, perturbation, tracking "
                                    "and even-odd are not provided");
    if (params.outer_type != OuterIntegrator::Leapfrog && params.outer_type != OuterIntegrator::Omelyan)
        throw std::invalid_argument("hmc_trajectory_mg_multiscale: the FGI integrators are not provided");
    const bool ident = !mg_precond || mg_precond.target<IdentityPrecond>() != nullptr;
    MGHierarchy* mg = ident ? nullptr : detail::mg_of(mg_precond, "hmc_trajectory_mg_multiscale");
    const std::array<Vec, 2> old = gauge.U;
    MomentumField mom(lat);
    mom.randomise(rng);
    (void)random_vec(lat.ndof, rng);   // the first generate_pseudofermion, overwritten (:380-383)
    const Vec eta = random_vec(lat.ndof, rng);
    DiracOp D(lat, gauge, mass, wilson_r, params.c_sw, params.mu_t);
    auto C = detail::defl_handle(P, cdefl);
    smg_ms_params cp{params.beta, params.tau, params.n_outer, params.n_inner,
                     params.outer_type == OuterIntegrator::Omelyan ? 1 : 0, params.inner_smooth, params.defl_refresh,
                     params.cg_maxiter, params.cg_tol, 1};
    const RVec pi0 = detail::mom_flat(mom);
    smg_ms_result r{};
    detail::check(smg_hmc_ms_trajectory(D.handle(), mg ? mg->handle() : nullptr, P.handle(), C.get(), &cp, pi0.data(),
                                        detail::raw(eta), 0.0, &r));
    detail::defl_read(C.get(), P.coarse_dim, cdefl);
    std::uniform_real_distribution<double> uniform(0.0, 1.0);
    const bool accept = (r.dH < 0) || (uniform(rng) < std::exp(-r.dH));
    if (accept) {
        Vec u(2 * (size_t)lat.V);
        detail::check(smg_dirac_get_gauge(D.handle(), detail::raw(u)));
        gauge.U[0].assign(u.begin(), u.begin() + lat.V);
        gauge.U[1].assign(u.begin() + lat.V, u.end());
    } else {
        gauge.U = old;
    }
    return {accept, r.dH, r.dKE, r.dSG, r.dSF, r.highmode_cg_iters, r.lowmode_force_evals, 0.0, 0.0};
}

// reversibility_test_plain, src/hmc_trajectory.cpp:416-509 (the gauge field is restored)
inline ReversibilityResult reversibility_test_plain(GaugeField& gauge, const Lattice& lat, double mass,
                                                    double wilson_r, const HMCParams& params, std::mt19937& rng,
                                                    const std::function<Vec(const Vec&)>* precond = nullptr) {
    detail::no_extras(precond, nullptr, "reversibility_test_plain");
    const std::array<Vec, 2> save = gauge.U;
    Vec phi;
    {
        DiracOp D0(lat, gauge, mass, wilson_r, params.c_sw, params.mu_t);
        generate_pseudofermion(D0, rng, phi);
    }
    MomentumField mom(lat);
    mom.randomise(rng);
    const std::array<RVec, 2> pi_save = mom.pi;
    auto H = [&](double* h) {
        DiracOp D(lat, gauge, mass, wilson_r, params.c_sw, params.mu_t);
        detail::hamiltonian(D, false, params.beta, mom, phi, params.cg_maxiter, params.cg_tol, h);
        return h[0] + h[1] + h[2];
    };
    double h[4];
    const double H0 = H(h);
    const int cg_fwd = plain_leapfrog_evolve(gauge, mom, lat, mass, wilson_r, phi, params);
    const double H1 = H(h);
    for (int mu = 0; mu < 2; mu++)
        for (auto& p : mom.pi[mu]) p = -p;
    const int cg_bwd = plain_leapfrog_evolve(gauge, mom, lat, mass, wilson_r, phi, params);
    const double H2 = H(h);
    double gn = 0, gd = 0, mn = 0, md = 0;
    for (int mu = 0; mu < 2; mu++)
        for (int s = 0; s < lat.V; s++) {
            gn += std::norm(save[mu][s]);
            gd += std::norm(gauge.U[mu][s] - save[mu][s]);
            mn += pi_save[mu][s] * pi_save[mu][s];
            const double d = -mom.pi[mu][s] - pi_save[mu][s];
            md += d * d;
        }
    gauge.U = save;
    return {std::sqrt(gd / std::max(gn, 1e-30)), std::sqrt(md / std::max(mn, 1e-30)), H1 - H0, H2 - H1,
            cg_fwd + cg_bwd};
}

// verify_forces, src/hmc_forces.cpp:869-981: central differences (eps = 1e-5) of the
// gauge and fermion actions on 20 links against the analytic forces, all on the GPU;
// true if both relative errors are below err_threshold
inline bool verify_forces(const GaugeField& g, double beta, double mass, double wilson_r, int max_iter, double tol,
                          double c_sw = 0.0, bool use_eo = false, double err_threshold = 1e-4, double mu_t = 0.0) {
    const Lattice& lat = g.lat;
    const double eps = 1e-5;
    std::mt19937 rng(12345);
    const int n_tests = 20;
    std::cout << "\n=== Force Verification (" << (use_eo ? "E/O" : "Full") << ", c_sw=" << c_sw
              << ", mu_t=" << mu_t << ") ===\n";
    DiracOp D0(lat, g, mass, wilson_r, c_sw, mu_t);
    std::array<RVec, 2> gf_ana, ff_ana;
    gauge_force(g, beta, gf_ana);
    const int nh = 2 * lat.V_half;
    Vec phi, phi_o;
    smg_solve_info info{};
    if (!use_eo) {
        const Vec eta = random_vec(lat.ndof, rng);
        D0.apply_dag(eta, phi);
        Vec X(lat.ndof);
        detail::check(smg_cg_solve(D0.handle(), SMG_OP_DDAGD, nullptr, detail::raw(phi), nullptr, detail::raw(X),
                                   max_iter, tol, &info));
        fermion_force(D0, X, ff_ana);
    } else {
        EvenOddDiracOp eo(D0);
        const Vec eta = random_vec(nh, rng);
        eo.apply_schur_dag(eta, phi_o);
        Vec x(nh), y(nh);
        detail::check(smg_cg_solve(D0.handle(), SMG_OP_SCHUR, nullptr, detail::raw(phi_o), nullptr, detail::raw(x),
                                   max_iter, tol, &info));
        eo.apply_schur(x, y);
        eo_fermion_force(D0, eo, x, y, ff_ana);
    }
    auto fermion_S = [&](const GaugeField& gf) {
        DiracOp Dp(lat, gf, mass, wilson_r, c_sw, mu_t);
        if (!use_eo) {
            Vec X(lat.ndof);
            detail::check(smg_cg_solve(Dp.handle(), SMG_OP_DDAGD, nullptr, detail::raw(phi), nullptr,
                                       detail::raw(X), max_iter, tol, &info));
            std::complex<double> acc(0.0, 0.0);
            for (int i = 0; i < lat.ndof; i++) acc += std::conj(phi[i]) * X[i];
            return acc.real();
        }
        EvenOddDiracOp eo(Dp);
        Vec x(nh);
        detail::check(smg_cg_solve(Dp.handle(), SMG_OP_SCHUR, nullptr, detail::raw(phi_o), nullptr, detail::raw(x),
                                   max_iter, tol, &info));
        std::complex<double> acc(0.0, 0.0);
        for (int i = 0; i < nh; i++) acc += std::conj(phi_o[i]) * x[i];
        double S = acc.real();
        if (c_sw != 0.0) S -= 2.0 * eo.log_det_ee();
        return S;
    };
    double max_gf_err = 0, max_ff_err = 0, max_gf_val = 0, max_ff_val = 0;
    for (int test = 0; test < n_tests; test++) {
        const int mu = test % 2, s = (test * 7 + test / 2 * 13) % lat.V;
        GaugeField gp(lat), gm(lat);
        gp.U = g.U;
        gm.U = g.U;
        gp.U[mu][s] *= std::exp(cx(0, eps));
        gm.U[mu][s] *= std::exp(cx(0, -eps));
        const double gf_num = -(gauge_action(gp, beta) - gauge_action(gm, beta)) / (2 * eps);
        const double ff_num = -(fermion_S(gp) - fermion_S(gm)) / (2 * eps);
        max_gf_err = std::max(max_gf_err, std::abs(gf_ana[mu][s] - gf_num));
        max_gf_val = std::max(max_gf_val, std::abs(gf_ana[mu][s]));
        max_ff_err = std::max(max_ff_err, std::abs(ff_ana[mu][s] - ff_num));
        max_ff_val = std::max(max_ff_val, std::abs(ff_ana[mu][s]));
        std::cout << "  link(" << mu << "," << std::setw(3) << s << "): gf_ana=" << std::setw(12) << gf_ana[mu][s]
                  << " gf_num=" << std::setw(12) << gf_num << " | ff_ana=" << std::setw(12) << ff_ana[mu][s]
                  << " ff_num=" << std::setw(12) << ff_num << "\n";
    }
    const double gf_rel = max_gf_err / std::max(max_gf_val, 1e-30);
    const double ff_rel = max_ff_err / std::max(max_ff_val, 1e-30);
    std::cout << "Gauge force:   max_err=" << std::scientific << max_gf_err << " rel=" << gf_rel
              << (gf_rel < err_threshold ? " PASS" : " FAIL") << "\n";
    std::cout << "Fermion force: max_err=" << std::scientific << max_ff_err << " rel=" << ff_rel
              << (ff_rel < err_threshold ? " PASS" : " FAIL") << "\n";
    return gf_rel < err_threshold && ff_rel < err_threshold;
}

}  // namespace schwinger_mg
