// runtime_lat.cpp -- multigrid-preconditioned GCR over many lattices of different sizes.
//
// The hierarchies are built lattice by lattice with the single-lattice setup (near-null
// vectors, aggregation, Galerkin operators, src/multigrid.cpp:595-687) and copied into
// one concatenated storage per level; from then on every step of the V-cycle
// (src/multigrid.cpp:43-102) and of GCR (src/solvers.cpp:223-300) is one launch over
// all lattices, each lattice keeping its own smoother, coarsest-CG and GCR scalars.
#include "runtime_lat.h"

#include <algorithm>
#include <cmath>
#include <deque>

namespace smg {

namespace {
// per-lattice scalar block, in doubles (G lattices)
struct ScLayout {
    size_t w, alpha, beta, cgaux, galpha, dots, apap, rr, nrm, nrmb, ints, bytes;
    explicit ScLayout(int G) {
        const size_t g = (size_t)G;
        w = 0;                                  // MR step [2G]
        alpha = w + 2 * g;                      // coarsest CG alpha [2G]
        beta = alpha + 2 * g;                   // coarsest CG beta [2G]
        cgaux = beta + 2 * g;                   // coarsest CG state [16G]
        galpha = cgaux + 16 * g;                // GCR alpha [2G]
        dots = galpha + 2 * g;                  // GCR projections [2 kSegWindow G]
        apap = dots + 2 * kSegWindow * g;       // |AP|^2 of every window slot [kSegWindow G]
        rr = apap + kSegWindow * g;             // |r|^2 [G]
        nrm = rr + g;                           // {., ., |b - A x|^2} [3G]
        nrmb = nrm + 3 * g;                     // {., ., |b|^2} [3G]
        ints = nrmb + 3 * g;                    // int: done, late, active[G]
        bytes = ints * sizeof(double) + (2 + g) * sizeof(int) + 64;
    }
};

// a device table and its workgroup count; fills wg0 and nwg of every entry (the
// lattices share a budget of workgroups, each at most kSegCap)
int upload(std::vector<SegDesc>& h, DevMem& d, hipStream_t st) {
    std::vector<long long> n;
    for (const SegDesc& s : h) n.push_back(s.n);
    const std::vector<int> w = lattice_workgroups(n);
    int wg = 0;
    for (size_t g = 0; g < h.size(); g++) {
        h[g].wg0 = wg;
        h[g].nwg = std::min(w[g], kSegCap);
        wg += h[g].nwg;
    }
    d.alloc(h.size() * sizeof(SegDesc));
    SMG_HIP(hipMemcpyAsync(d.p, h.data(), h.size() * sizeof(SegDesc), hipMemcpyHostToDevice, st));
    return wg;
}
}  // namespace

LatFin LatticeMG::fin(int op, double* out, int ostride, double* aux, int astride) const {
    const Fin f = Context::get().fin(op, out, aux);
    LatFin l;
    l.op = op;
    l.out = out;
    l.ostride = ostride;
    l.aux = aux;
    l.astride = astride;
    l.part = f.part;
    l.counter = f.counter;
    return l;
}

LatticeMG::LatticeMG(const std::vector<const Dirac*>& dirac, const LatDesc* fine_table, int nwg_sites_,
                     const MGConfig& c, const c64* starts)
    : G((int)dirac.size()), nlev(c.n_levels), cfg(c), fine(fine_table), nwg_sites(nwg_sites_) {
    SMG_REQUIRE(G >= 1 && G <= kSegMaxLat, "many-lattice multigrid: 1..128 lattices");
    SMG_REQUIRE(nlev >= 2 && nlev <= 5, "n_levels must be 2..5");
    SMG_REQUIRE(c.coarse_solver == 0, "many-lattice multigrid: the coarsest solve is CG");
    for (int l = 0; l + 1 < nlev; l++) {
        const int k = c.k[l];
        SMG_REQUIRE(k >= 2 && k <= 64 && (k & (k - 1)) == 0,
                    "many-lattice multigrid: k must be a power of two in 2..64 at every level");
        SMG_REQUIRE(!c.chiral || k % 2 == 0, "chiral aggregation needs an even k at every level");
    }
    // every lattice's level extents, in the order the single-lattice setup makes them
    kdim.assign(nlev, 2);
    for (int l = 1; l < nlev; l++) kdim[l] = c.k[l - 1];
    std::vector<std::vector<int>> nx(G, std::vector<int>(nlev));
    for (int g = 0; g < G; g++) {
        nx[g][0] = dirac[g]->L;
        for (int l = 0; l + 1 < nlev; l++) {
            SMG_REQUIRE(c.bx[l] == c.by[l] && c.bx[l] >= 1 && nx[g][l] % c.bx[l] == 0,
                        "many-lattice multigrid: square aggregates that tile every lattice");
            nx[g][l + 1] = nx[g][l] / c.bx[l];
        }
    }
    hipStream_t st = Context::get().stream;
    // offsets of every lattice in the concatenated storage
    hent.assign(nlev, {});
    len.assign(nlev, 0);
    std::vector<size_t> bvlen(nlev - 1, 0), lklen(nlev - 1, 0);
    for (int g = 0; g < G; g++)
        for (int l = 0; l < nlev; l++) {
            SegDesc s{};
            s.nx = nx[g][l];
            s.nd = kdim[l];
            s.n = (long long)s.nx * s.nx * s.nd;
            s.off = (long long)len[l];
            len[l] += (size_t)s.n;
            if (l + 1 < nlev) {
                s.bv_off = (long long)bvlen[l];
                bvlen[l] += (size_t)s.n * c.k[l];
            }
            if (l >= 1) {
                s.link_off = (long long)lklen[l - 1];
                lklen[l - 1] += (size_t)s.nx * s.nx * 9 * s.nd * s.nd;
            }
            hent[l].push_back(s);
        }
    for (int l = 0; l + 1 < nlev; l++) bv.emplace_back(bvlen[l]);
    for (int l = 1; l < nlev; l++) {
        links.emplace_back(lklen[l - 1]);
        links16.emplace_back(cfg.coarse_half ? lklen[l - 1] * sizeof(ch16) : 0);
    }
    // build each lattice's hierarchy and copy it in
    size_t soff = 0;
    for (int g = 0; g < G; g++) {
        MG mg(*dirac[g], c, starts + soff, false);
        for (int l = 0; l + 1 < nlev; l++) {
            const int nv = c.chiral ? c.k[l] / 2 : c.k[l];
            soff += (size_t)nv * hent[l][g].n;
            SMG_REQUIRE(mg.dims[l] == (size_t)hent[l][g].n && mg.P[l]->bv64.n == (size_t)hent[l][g].n * c.k[l],
                        "many-lattice multigrid: level sizes differ from the single-lattice hierarchy");
            SMG_HIP(hipMemcpyAsync(bv[l].p(hent[l][g].bv_off), mg.P[l]->bv64.p(), mg.P[l]->bv64.n * sizeof(c64),
                                   hipMemcpyDeviceToDevice, st));
            const CoarseOperator& C = *mg.C[l];
            SMG_REQUIRE(C.nlinks() == (size_t)hent[l + 1][g].n * hent[l + 1][g].nd * 9,
                        "many-lattice multigrid: coarse link count differs");
            SMG_HIP(hipMemcpyAsync(links[l].p(hent[l + 1][g].link_off), C.l64.p(), C.nlinks() * sizeof(c64),
                                   hipMemcpyDeviceToDevice, st));
            if (cfg.coarse_half)   // the single-lattice FP16 copy (made exactly Hermitian where it can be)
                SMG_HIP(hipMemcpyAsync(links16[l].as<ch16>() + hent[l + 1][g].link_off, C.l16.as<ch16>(),
                                       C.nlinks() * sizeof(ch16), hipMemcpyDeviceToDevice, st));
        }
        SMG_HIP(hipStreamSynchronize(st));   // before the single-lattice hierarchy is freed
    }
    // launch tables
    ent.resize(nlev);
    lanes.resize(nlev);
    nwg_ent.assign(nlev, 0);
    nwg_lanes.assign(nlev, 0);
    for (int l = 0; l < nlev; l++) {
        nwg_ent[l] = upload(hent[l], ent[l], st);
        if (l >= 1) {
            std::vector<SegDesc> hl = hent[l];
            for (SegDesc& s : hl) s.n = (long long)s.nx * s.nx * s.nd * s.nd / seg_coarse_run(s.nd);
            nwg_lanes[l] = upload(hl, lanes[l], st);
        }
    }
    SMG_REQUIRE((size_t)std::max(*std::max_element(nwg_ent.begin(), nwg_ent.end()),
                                 *std::max_element(nwg_lanes.begin(), nwg_lanes.end())) *
                        2 * kSegWindow <= (size_t(1) << 21),
                "many-lattice multigrid: too many workgroups for the partial sums");
    SMG_REQUIRE(nwg_sites * 3 <= (1 << 21), "too many sites for one launch's partial sums");
    for (int l = 0; l < nlev; l++) {
        xs.emplace_back(len[l]);
        bs.emplace_back(len[l]);
        rs.emplace_back(len[l]);
        Ars.emplace_back(len[l]);
        const bool w = cfg.w_cycle && l >= 1 && l + 1 < nlev;
        ts.emplace_back(w ? len[l] : 1);
        e2s.emplace_back(w ? len[l] : 1);
    }
    cp = Field<double>(len[nlev - 1]);
    cAp = Field<double>(len[nlev - 1]);
    t0 = Field<double>(len[0]);
    sc.alloc(ScLayout(G).bytes);
    SMG_HIP(hipStreamSynchronize(st));
}

LatticeMG::~LatticeMG() { reset_graph(); }

void LatticeMG::reset_graph() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    gexec = nullptr;
    graph = nullptr;
}

void LatticeMG::apply(int l, c64* dst, const c64* src, Epi epi, const c64* bvec, const c64* aux, const LatFin& f,
                      hipStream_t st) {
    if (l == 0) {
        // D^dag D = (g5 D g5) D: t = D src, dst = g5 D(mu -> -mu) g5 t
        launch_ml_dslash(fine, G, nwg_sites, t0.p(), src, 1.0, 0, 0, EPI_NONE, nullptr, LatFin(), st);
        launch_ml_dslash(fine, G, nwg_sites, dst, t0.p(), -1.0, 1, 1, epi == EPI_DOTS ? EPI_DOTS : EPI_NONE, aux,
                         epi == EPI_DOTS ? f : LatFin(), st);
        if (epi == EPI_RESID) launch_axpby<double>(dst, bvec, c64{1.0, 0.0}, c64{-1.0, 0.0}, len[0], st);
        return;
    }
    if (cfg.coarse_half)
        launch_seg_coarse(lanes[l].as<SegDesc>(), G, nwg_lanes[l], kdim[l], links16[l - 1].as<ch16>(), src, dst, epi,
                          bvec, aux, f, st);
    else
        launch_seg_coarse(lanes[l].as<SegDesc>(), G, nwg_lanes[l], kdim[l], links[l - 1].p(), src, dst, epi, bvec,
                          aux, f, st);
}

// MR sweeps (src/smoother.cpp:20-45) with each lattice's own step
void LatticeMG::smooth(int l, c64* x, const c64* b, int n, bool zero_init, hipStream_t st) {
    const ScLayout Ls(G);
    double* w = sc.as<double>() + Ls.w;
    c64* r = rs[l].p();
    c64* Ar = Ars[l].p();
    if (zero_init) {
        launch_zero<double>(x, len[l], st);
        launch_copy<double>(r, b, len[l], st);
    } else {
        apply(l, r, x, EPI_RESID, b, nullptr, LatFin(), st);
    }
    const LatFin fmr = fin(FIN_MR, w, 2, nullptr, 0);
    for (int it = 0; it < n; it++) {
        apply(l, Ar, r, EPI_DOTS, nullptr, r, fmr, st);
        launch_seg_mr_update(ent[l].as<SegDesc>(), G, nwg_ent[l], x, r, Ar, w, st);
    }
}

// CG on every lattice's coarsest operator (src/coarse_op.h:508-541): a fixed number of
// launches, each lattice stopping on its own tolerance, all of them skipped once every
// lattice has stopped -- nothing is read back, so the cycle captures as a graph
void LatticeMG::coarsest(const c64* b, c64* x, hipStream_t st) {
    const int l = nlev - 1;
    const ScLayout Ls(G);
    double* s = sc.as<double>();
    double* alpha = s + Ls.alpha;
    double* beta = s + Ls.beta;
    double* aux = s + Ls.cgaux;
    int* done = reinterpret_cast<int*>(s + Ls.ints);
    int* late = done + 1;
    const SegDesc* d = ent[l].as<SegDesc>();
    launch_ml_cg_state(aux, G, cfg.coarse_tol * cfg.coarse_tol, done, late, st);
    launch_seg_cg_init(d, G, nwg_ent[l], b, x, rs[l].p(), cp.p(), fin(FIN_CG_INIT, nullptr, 2, aux, 16), st);
    LatFin fa = fin(FIN_CG_ALPHA, alpha, 2, aux, 16), fb = fin(FIN_CG_BETA, beta, 2, aux, 16);
    fa.skip = done;
    fb.skip = done;
    fb.done = done;
    for (int it = 0; it < cfg.coarse_maxit; it++) {
        apply(l, cAp.p(), cp.p(), EPI_DOTS, nullptr, cp.p(), fa, st);
        launch_seg_cg_r(d, G, nwg_ent[l], rs[l].p(), cAp.p(), alpha, fb, late, st);
        launch_seg_cg_xp(d, G, nwg_ent[l], x, cp.p(), rs[l].p(), alpha, beta, aux, late, st);
    }
}

// src/multigrid.cpp:43-102
void LatticeMG::run(int l, const c64* b, c64* x, hipStream_t st) {
    const int last = nlev - 1;
    if (l == last) {
        coarsest(b, x, st);
        return;
    }
    smooth(l, x, b, cfg.pre, true, st);
    // r = b - A x is carried by the smoother
    const SegDesc* fd = ent[l].as<SegDesc>();
    const SegDesc* cd = ent[l + 1].as<SegDesc>();
    launch_seg_restrict(fd, cd, G, nwg_ent[l + 1], bv[l].p(), rs[l].p(), bs[l + 1].p(), cfg.bx[l], cfg.by[l],
                        cfg.chiral, st);
    run(l + 1, bs[l + 1].p(), xs[l + 1].p(), st);
    if (cfg.w_cycle && l + 1 < last) {
        apply(l + 1, ts[l + 1].p(), xs[l + 1].p(), EPI_RESID, bs[l + 1].p(), nullptr, LatFin(), st);
        run(l + 1, ts[l + 1].p(), e2s[l + 1].p(), st);
        launch_axpby<double>(xs[l + 1].p(), e2s[l + 1].p(), c64{1.0, 0.0}, c64{1.0, 0.0}, len[l + 1], st);
    }
    launch_seg_prolong(fd, cd, G, nwg_ent[l], bv[l].p(), xs[l + 1].p(), x, 1, cfg.bx[l], cfg.by[l], cfg.chiral, st);
    smooth(l, x, b, cfg.post, false, st);
}

void LatticeMG::precondition(const c64* b, c64* x, hipStream_t st) {
    if (!cfg.use_graph) {
        run(0, b, x, st);
        return;
    }
    if (!gexec) {
        gin = Field<double>(len[0]);
        gout = Field<double>(len[0]);
        SMG_HIP(hipStreamSynchronize(st));
        SMG_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        try {
            run(0, gin.p(), gout.p(), st);
        } catch (...) {
            hipGraph_t tmp = nullptr;
            (void)hipStreamEndCapture(st, &tmp);
            if (tmp) (void)hipGraphDestroy(tmp);
            throw;
        }
        SMG_HIP(hipStreamEndCapture(st, &graph));
        SMG_HIP(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
    }
    launch_copy<double>(gin.p(), b, len[0], st);
    SMG_HIP(hipGraphLaunch(gexec, st));
    launch_copy<double>(x, gout.p(), len[0], st);
}

// src/solvers.cpp:223-300 (truncated GCR, classical Gram-Schmidt against the kept
// directions) on every lattice at once: one window of directions shared by all
// lattices, each with its own projections and step; a lattice that has converged
// keeps its x and r (its update is masked) while the others go on.
int LatticeMG::gcr(const c64* b, c64* x, int maxit, double tol, int trunc, int* iters, double* res, hipStream_t st) {
    SMG_REQUIRE(trunc >= 1 && trunc + 1 <= kSegWindow, "many-lattice GCR: truncation window 1..7");
    const size_t n = len[0];
    const SegDesc* d0 = ent[0].as<SegDesc>();
    const int nw = nwg_ent[0];
    const ScLayout Ls(G);
    double* s = sc.as<double>();
    double* galpha = s + Ls.galpha;
    double* dots = s + Ls.dots;
    double* apap = s + Ls.apap;
    double* rr = s + Ls.rr;
    int* active = reinterpret_cast<int*>(s + Ls.ints) + 2;
    auto read = [&](const double* dev, std::vector<double>& h) {
        SMG_HIP(hipMemcpyAsync(h.data(), dev, h.size() * sizeof(double), hipMemcpyDeviceToHost, st));
        SMG_HIP(hipStreamSynchronize(st));
    };
    launch_zero<double>(x, n, st);
    Field<double> r(n);
    launch_copy<double>(r.p(), b, n, st);
    std::vector<double> hb(3 * (size_t)G), hr((size_t)G);
    launch_seg_dots(d0, G, nw, b, b, fin(FIN_STORE, s + Ls.nrmb, 3, nullptr, 0), st);
    read(s + Ls.nrmb, hb);
    std::vector<double> bn(G), rn(G);
    for (int g = 0; g < G; g++) rn[g] = bn[g] = std::sqrt(hb[3 * (size_t)g + 2]);
    std::vector<Field<double>> P, AP;
    for (int i = 0; i <= trunc; i++) {
        P.emplace_back(n);
        AP.emplace_back(n);
    }
    std::deque<int> order;
    std::vector<int> freeslots;
    for (int i = trunc; i >= 0; i--) freeslots.push_back(i);
    std::vector<int> act(G, 0), its(G, 0);
    int iter = 0;
    while (iter < maxit) {
        int any = 0;
        for (int g = 0; g < G; g++) {
            act[g] = bn[g] >= 1e-30 && rn[g] / bn[g] >= tol;
            any |= act[g];
        }
        if (!any) break;
        SMG_HIP(hipMemcpyAsync(active, act.data(), G * sizeof(int), hipMemcpyHostToDevice, st));
        const int slot = freeslots.back();
        freeslots.pop_back();
        c64* p = P[slot].p();
        c64* Ap = AP[slot].p();
        precondition(r.p(), p, st);
        apply(0, Ap, p, EPI_NONE, nullptr, nullptr, LatFin(), st);
        const int nd = (int)order.size();
        if (nd > 0) {
            SegVecs Pv{}, APv{};
            SegRows ap{};
            for (int i = 0; i < nd; i++) {
                Pv.p[i] = P[order[i]].p();
                APv.p[i] = AP[order[i]].p();
                ap.p[i] = apap + (size_t)order[i] * G;
            }
            launch_seg_multi_dot(d0, G, nw, APv, nd, Ap, fin(FIN_STORE, dots, 2 * kSegWindow, nullptr, 0), st);
            launch_seg_gcr_orth(d0, G, nw, p, Ap, Pv, APv, nd, dots, ap, st);
        }
        launch_seg_dots(d0, G, nw, Ap, r.p(), fin(FIN_GCR_ALPHA, galpha, 2, apap + (size_t)slot * G, 1), st);
        launch_seg_update(d0, G, nw, x, r.p(), p, Ap, galpha, active, fin(FIN_STORE, rr, 1, nullptr, 0), st);
        read(rr, hr);
        for (int g = 0; g < G; g++)
            if (act[g]) {
                its[g]++;
                rn[g] = std::sqrt(hr[g]);
            }
        iter++;
        if ((int)order.size() >= trunc) {
            freeslots.push_back(order.front());
            order.pop_front();
        }
        order.push_back(slot);
    }
    // FP64 true residuals per lattice
    apply(0, Ars[0].p(), x, EPI_RESID, b, nullptr, LatFin(), st);
    std::vector<double> hn(3 * (size_t)G);
    launch_seg_dots(d0, G, nw, Ars[0].p(), Ars[0].p(), fin(FIN_STORE, s + Ls.nrm, 3, nullptr, 0), st);
    read(s + Ls.nrm, hn);
    for (int g = 0; g < G; g++) {
        if (iters) iters[g] = its[g];
        if (res) res[g] = bn[g] > 0.0 ? std::sqrt(hn[3 * (size_t)g + 2]) / bn[g] : 0.0;
    }
    return iter;
}

}  // namespace smg
