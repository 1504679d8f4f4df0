"""Benchmark: MG-GCR solves/s on a 512 x 512 Wilson-clover Schwinger lattice.

One step = a batch of N (default 32) Gaussian right-hand sides, each solved to a
true residual |phi - A x|/|phi| <= 1e-10 (A = D^dag D, the reference's system,
src/solvers.cpp:55-59) by mixed-precision GCR + multigrid run on all N at once:
FP32 GCR and an FP32 V-cycle with FP16 coarse links (bf16x3 matrix-core products)
inside FP64 iterative refinement on the FP64 true residual.  value = N / (seconds
per step).  The same hierarchy solved in plain FP64 (FP64 GCR + FP64 cycle, also
batched) is timed beside it, as is the one-right-hand-side solve; the
Wilson-clover stencil is timed alone (HIP events on the library's stream), and
the oracle's FP64 CPU MG-GCR solve of one right-hand side on the same hierarchy
is timed on the host cores.  Inputs are synthetic (seeded U(1) links
theta ~ N(0, 0.4^2), Gaussian sources) and resident in HBM when the timed region
starts.  roofline.traffic is the stencil's HBM bytes per launch from the
FETCH_SIZE / WRITE_SIZE counter runs summarised in profiles/.  Prints one JSON
line.

    python bench.py [--gpus 1] [--steps 5] [--warmup 1]

With --workload hmc it times trajectories of the nested Omelyan HMC at the same L
(beta 6, m 0.02, the fermion force on one outer step, the gauge force on 8 inner
steps, every solve the mixed-precision MG) and prints that line instead.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from schwinger_mg_amd.fields import gaussian_vec, gaussian_vecs, random_gauge  # noqa: E402

HBM_PEAK_GBS = 8000.0


def marker(torch, name, start):
    """roctx range around a timed step (a no-op where torch has no marker support)."""
    try:
        if start:
            torch.cuda.nvtx.range_push(name)
        else:
            torch.cuda.nvtx.range_pop()
    except Exception:
        pass


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--L", type=int, default=512)
    ap.add_argument("--mass", type=float, default=0.05)
    ap.add_argument("--sigma", type=float, default=0.4)
    ap.add_argument("--levels", type=int, default=3)
    ap.add_argument("--k", type=int, default=32, help="coarse dof per aggregate (2 x null vectors when chiral)")
    ap.add_argument("--chiral", type=int, default=1)
    ap.add_argument("--w-cycle", type=int, default=0)
    ap.add_argument("--null-iters", type=int, default=10)
    ap.add_argument("--pre", type=int, default=1, help="mixed batched cycle: pre-smoothing steps")
    ap.add_argument("--post", type=int, default=2, help="mixed batched cycle: post-smoothing steps")
    ap.add_argument("--mr-pre", type=int, default=3,
                    help="MR pre-smoothing sweeps of the FP64, single-right-hand-side, lattice-batch and CPU cycles")
    ap.add_argument("--mr-post", type=int, default=3, help="their MR post-smoothing sweeps")
    ap.add_argument("--coarse-maxit", type=int, default=30)
    ap.add_argument("--coarse-tol", type=float, default=1e-1)
    ap.add_argument("--coarse-solver", type=int, default=2,
                    help="0 CG, 1 dense inverse, 2 CG in FP64 and a Chebyshev iteration in the mixed cycle")
    ap.add_argument("--cheb-smooth", type=float, default=12.0,
                    help="Chebyshev smoothing in the mixed batched cycle on [hi / R, hi] (0: MR)")
    ap.add_argument("--cheb-lo-scale", type=float, default=1.0,
                    help="coarse-solver 2: the Chebyshev interval's lower end times this")
    ap.add_argument("--cheb-hi-scale", type=float, default=1.0,
                    help="coarse-solver 2: the Chebyshev interval's upper end times this")
    ap.add_argument("--cheb-steps", type=int, default=32,
                    help="coarse-solver 2: Chebyshev steps of the mixed cycle's coarsest solve")
    ap.add_argument("--coarse-bf16", type=int, default=1,
                    help="BF16 vectors in the coarsest CG of the FP32 cycle (FP16 links either way)")
    ap.add_argument("--graph", type=int, default=1)
    ap.add_argument("--trunc", type=int, default=10)
    ap.add_argument("--inner-tol", type=float, default=1e-5)
    ap.add_argument("--fp64-steps", type=int, default=1)
    ap.add_argument("--nrhs", type=int, default=32)
    ap.add_argument("--cpu-baseline", type=int, default=1)
    ap.add_argument("--single", type=int, default=1, help="also time one right-hand side at a time")
    ap.add_argument("--fp64", type=int, default=1, help="also time the FP64 batched solve")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--lattice-batch", type=int, default=1,
                    help="also time the 40-lattice batch of the workload (one launch over all lattices)")
    ap.add_argument("--lattice-mg", type=int, default=1,
                    help="lattice batch: also the three-level MG over all 40 lattices with GCR")
    ap.add_argument("--summarize-trace", default=None, metavar="DIR",
                    help="print the per-range kernel table of a rocprofv3 kernel+marker trace and exit")
    ap.add_argument("--range", default="mixed_batch", help="--summarize-trace: the marker range name")
    ap.add_argument("--summarize-pmc", nargs="+", default=None, metavar="DIR",
                    help="print per-kernel counter averages of rocprofv3 --pmc runs and exit")
    ap.add_argument("--summarize-kernels", default=None, metavar="DIR",
                    help="print per-kernel, per-grid durations of a rocprofv3 kernel trace and exit")
    ap.add_argument("--kernel-filter", default=None,
                    help="--summarize-pmc / --summarize-kernels: kernels whose name contains this")
    ap.add_argument("--kernels-select", default=None,
                    help="kernels workload: time only the kernels whose name contains this")
    ap.add_argument("--tune-load", default=None, metavar="PATH", help="kernels workload: read this tuner table")
    ap.add_argument("--tune-save", default=None, metavar="PATH", help="kernels workload: write the tuner table")
    ap.add_argument("--workload", choices=("solve", "hmc", "kernels"), default="solve",
                    help="solve: the batched MG solve (default); hmc: trajectories of the nested Omelyan HMC; "
                         "kernels: the hierarchy's setup, then only the stencil and coarse-operator launches (and a "
                         "plain copy of 128 MiB) that the counter runs summarised in profiles/ measure")
    ap.add_argument("--tune", type=int, default=1,
                    help="run the launch-shape autotuner on the hierarchy before timing (its time is reported)")
    ap.add_argument("--beta", type=float, default=6.0, help="hmc: gauge coupling")
    ap.add_argument("--hmc-mass", type=float, default=0.02, help="hmc: fermion mass")
    ap.add_argument("--n-outer", type=int, default=1, help="hmc: outer (fermion-force) steps")
    ap.add_argument("--n-gauge", type=int, default=8, help="hmc: gauge-force steps per outer position update")
    ap.add_argument("--therm", type=int, default=200, help="hmc: heatbath sweeps before the trajectories")
    ap.add_argument("--hmc-null-iters", type=int, default=40, help="hmc: null-space iterations of the hierarchy")
    ap.add_argument("--hmc-coarse-maxit", type=int, default=100, help="hmc: coarsest CG iterations")
    ap.add_argument("--hmc-coarse-tol", type=float, default=1e-2, help="hmc: coarsest CG tolerance")
    ap.add_argument("--mg-refresh", type=int, default=1,
                    help="hmc: refresh the hierarchy every N outer position updates (0: Galerkin rebuild between "
                         "trajectories only)")
    ap.add_argument("--mg-refresh-iters", type=int, default=2, help="hmc: near-null iterations per refresh")
    ap.add_argument("--chrono", type=int, default=1,
                    help="hmc: every solve after the first starts from the previous solution")
    ap.add_argument("--lobpcg", type=int, default=16,
                    help="hmc: low modes of D^dag D tracked by LOBPCG (one block) after every trajectory (0: none)")
    return ap.parse_args()


TRAFFIC_PROFILES = {
    "dslash_f32": os.path.join(ROOT, "profiles", "k_b_dslash_f32_L512_N32_hbm_traffic.md"),
    "ddagd_f32": os.path.join(ROOT, "profiles", "k_b_ddagd_f32_L512_N32_hbm_traffic.md"),
    "coarse": os.path.join(ROOT, "profiles", "k_b_coarse_L512_N32_hbm_traffic.md"),
}


def profiled_traffic(L, nrhs, kernel="dslash_f32", key="traffic_bytes_per_launch"):
    """HBM bytes per launch of a kernel at L=512, N=32 from its counter summary (a line
    '<key>: <int>'); None if absent or if the run is at another size."""
    path = TRAFFIC_PROFILES[kernel]
    if L != 512 or nrhs != 32 or not os.path.exists(path):
        return None
    for line in open(path):
        if line.strip().startswith(key + ":"):
            return int(line.split(":")[1].strip().strip("`"))
    return None


def stencil_bytes_per_site(prec, nrhs, csw):
    sc = 16 if prec == "f64" else 8      # complex
    sr = 8 if prec == "f64" else 4       # real
    return nrhs * 2 * 2 * sc + 2 * sc + (sr if csw else 0)


def time_stencil(smg, torch, D, prec, nrhs, V, csw, launches=48, warm=6, same=True, kind=None):
    """Average duration of one Wilson-clover launch on the batched layout (RHS
    innermost, the layout of the solve): a HIP event pair on the library's stream (the
    stream the kernel runs on) around `launches` back-to-back launches, divided by their
    number.  The launches form a chain over three vectors, each reading the vector the
    launch before it wrote, as the solve's launches do (its input then comes from HBM, not
    from the last-level cache that a re-read of one unchanged input would hit); the
    duration of launches that all read one unchanged vector is reported beside it
    (same_input_us).  kind: the Dirac kind (default D; D^dag D runs as one fused
    launch where the tuner chose it)."""
    dt = torch.complex128 if prec == "f64" else torch.complex64
    bufs = [torch.randn(2 * V, nrhs, dtype=dt, device="cuda") for _ in range(3)]
    p = smg.F64 if prec == "f64" else smg.F32
    sp = smg.library_stream()
    st = torch.cuda.ExternalStream(sp)
    import ctypes
    sarg = ctypes.c_void_p(sp)
    torch.cuda.synchronize()

    def launch(i, chain):
        src, dst = (bufs[i % 3], bufs[(i + 1) % 3]) if chain else (bufs[0], bufs[1])
        smg._chk(smg.lib().smg_dirac_apply_batched_dev(D.h, smg.D_ if kind is None else kind, p, src.data_ptr(),
                                                       dst.data_ptr(), nrhs, sarg))

    def run(chain):
        for i in range(warm):
            launch(i, chain)
        smg.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for i in range(launches):
            launch(i, chain)
        e1.record(st)
        e1.synchronize()
        for b in bufs:   # the chain grows its vectors by the operator's norm: rescale
            b.mul_(1.0 / max(float(b.abs().max()), 1e-30))
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / launches

    ms_same = run(False) if same else None
    ms = run(True)
    byt = stencil_bytes_per_site(prec, nrhs, csw) * V
    out = {"prec": prec, "nrhs": nrhs, "us_per_launch": ms * 1e3, "bytes_per_launch": byt,
           "GBs": byt / (ms * 1e-3) / 1e9}
    if same:
        out.update(same_input_us=ms_same * 1e3, same_input_GBs=byt / (ms_same * 1e-3) / 1e9)
    return out


PEAKS = {"hbm": 8000.0, "bf16": 2500.0, "f64": 78.6, "f32": 157.3}   # GB/s; TFLOP/s (dense MFMA)


def time_coarse(smg, torch, mg, level, prec, nrhs, k, launches=50, warm=5, cheb=False):
    """Average duration of one launch of the hierarchy's level-`level` coarse operator on
    the batched layout with the fused dot products (the launch shape of the smoother and
    the coarse CG), or with cheb (BF16 vectors) one step of the Chebyshev coarsest solve:
    a HIP event pair around `launches` back-to-back launches on the library's stream,
    divided by their number.
    Algorithmic bytes per launch Vc (9 k^2 s_L + 2 k s_psi N) (links of s_L bytes per
    complex entry, sources read and results written once), plus 4 k s_psi N for the
    Chebyshev step's residual and iterate (each read and written once); flops 72 k^2 Vc N."""
    import ctypes
    n = mg.level_dim(level)
    vc = n // k
    sl, sp, dt = {smg.F64: (16, 16, torch.complex128), smg.F16: (4, 8, torch.complex64),
                  smg.BF16V: (4, 4, torch.int32)}[prec]
    if prec == smg.BF16V:     # bf16 pairs: random bits of moderate exponent
        src = (torch.randint(0, 1 << 14, (n * nrhs,), dtype=torch.int32, device="cuda") | 0x3c003c00)
    else:
        src = torch.randn(n * nrhs, dtype=dt, device="cuda")
    dst = torch.empty_like(src)
    sp_ = smg.library_stream()
    st = torch.cuda.ExternalStream(sp_)
    sarg = ctypes.c_void_p(sp_)
    torch.cuda.synchronize()

    def launch():
        smg._chk(smg.lib().smg_mg_coarse_apply_dev(mg.h, level, prec, nrhs, src.data_ptr(), dst.data_ptr(),
                                                   2 if cheb else 1, sarg))

    for _ in range(warm):
        launch()
    smg.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(launches):
        launch()
    e1.record(st)
    e1.synchronize()
    us = e0.elapsed_time(e1) / launches * 1e3
    byt = vc * (9 * k * k * sl + (6 if cheb else 2) * k * sp * nrhs)
    flop = 72 * k * k * vc * nrhs
    return {"level": level, "cheb": cheb, "prec": {smg.F64: "f64", smg.F16: "f16 links, f32 vectors",
                                                   smg.BF16V: "f16 links, bf16 vectors"}[prec],
            "sites": vc, "k": k, "nrhs": nrhs, "us_per_launch": us, "bytes_per_launch": byt,
            "GBs": byt / (us * 1e-6) / 1e9, "TFLOPs": flop / (us * 1e-6) / 1e12}


def lattice_batch(smg, torch, gaussian_vec, random_gauge, args):
    """The configuration-3 batch of 40 lattices (L in {64, 96, 128, 192, 256}, 8 each,
    theta ~ N(0, 0.4^2), independent seeds, m = 0.1, c_sw = 1) as one Lattices handle:
    the FP64 Wilson-clover D over all 40 in one launch (event-timed on the library
    stream over 50 back-to-back launches), the lockstep CG on D^dag D of all 40 to 1e-10, and the
    three-level MG (4x4 then 2x2 aggregates, 16 near-null vectors per level, chiral:
    k = 32) over all 40 with GCR to 1e-10, every kernel one launch for all lattices
    and the V-cycle replayed from a hipGraph (solves: host-array entries, wall time
    including their PCIe copies of b and x; the cycle event-timed on device arrays)."""
    import ctypes
    Ls = [L for L in (64, 96, 128, 192, 256) for _ in range(8)]
    Us = [random_gauge(L, 0.4, 800 + i) for i, L in enumerate(Ls)]
    B = smg.Lattices(Ls, Us, 0.1, 1.0, 1.0, 0.0)
    n = int(B.offsets[-1])
    sites = n // 2
    src = torch.randn(n, dtype=torch.complex128, device="cuda")
    dst = torch.empty_like(src)
    sp = smg.library_stream()
    st = torch.cuda.ExternalStream(sp)
    sarg = ctypes.c_void_p(sp)
    torch.cuda.synchronize()

    def launch():
        smg._chk(smg.lib().smg_lattices_apply_dev(B.h, smg.D_, src.data_ptr(), dst.data_ptr(), sarg))

    for _ in range(5):
        launch()
    smg.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(50):
        launch()
    e1.record(st)
    e1.synchronize()
    us = e0.elapsed_time(e1) / 50 * 1e3
    byt = stencil_bytes_per_site("f64", 1, True) * sites
    bs = [gaussian_vec(2 * L * L, 900 + i) for i, L in enumerate(Ls)]
    B.cg(bs, 5000, 1e-10)   # warm-up
    smg.synchronize()
    t0 = time.perf_counter()
    marker(torch, "lattice_cg", True)
    _, its, tr = B.cg(bs, 5000, 1e-10)
    marker(torch, "lattice_cg", False)
    dt = time.perf_counter() - t0
    out = {"lattices": len(Ls), "L": [64, 96, 128, 192, 256], "sites": sites,
           "dslash_f64_us": round(us, 2), "dslash_f64_GBs": round(byt / (us * 1e-6) / 1e9, 1),
           "cg_batch_s": round(dt, 4), "cg_lattice_solves_per_s": round(len(Ls) / dt, 2),
           "cg_iterations": [int(its.min()), int(its.max())], "max_true_residual": float(tr.max())}
    if not args.lattice_mg:
        return out
    kw = dict(n_levels=3, blocks=((4, 4), (2, 2)), k=(32, 32), null_iters=args.null_iters, pre_smooth=args.mr_pre,
              post_smooth=args.mr_post, coarse_maxit=args.coarse_maxit, coarse_tol=args.coarse_tol, chiral=True,
              coarse_half=bool(args.coarse_bf16))
    starts = []
    for g, L in enumerate(Ls):
        sz = smg.mg_start_sizes(L, kw["n_levels"], kw["blocks"], kw["k"], kw["chiral"])
        starts.append(np.concatenate([gaussian_vecs(nv, m, 1000 + 10 * g + l).ravel()
                                      for l, (nv, m) in enumerate(sz)]))
    smg.synchronize()
    t0 = time.perf_counter()
    M = smg.LatticesMG(B, starts, use_graph=True, **kw)
    smg.synchronize()
    setup_s = time.perf_counter() - t0
    del starts
    trunc = 7
    M.gcr(bs, 500, 1e-10, trunc)   # warm-up (captures the cycle)
    smg.synchronize()
    t0 = time.perf_counter()
    marker(torch, "lattice_mg_gcr", True)
    _, mits, mtr, _ = M.gcr(bs, 500, 1e-10, trunc)
    marker(torch, "lattice_mg_gcr", False)
    mdt = time.perf_counter() - t0
    x = torch.empty_like(src)

    def cycle():
        smg._chk(smg.lib().smg_lattices_mg_precondition_dev(M.h, src.data_ptr(), x.data_ptr(), sarg))

    for _ in range(3):
        cycle()
    smg.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(10)]
    for e0, e1 in ev:
        e0.record(st)
        cycle()
        e1.record(st)
    ev[-1][1].synchronize()
    cyc_ms = sum(e0.elapsed_time(e1) for e0, e1 in ev) / 10
    out.update({"mg": "3 levels, 4x4 then 2x2 aggregates, 16 near-null vectors (chiral, k=32), FP64 vectors, "
                      f"{'FP16' if args.coarse_bf16 else 'FP64'} coarse links, "
                      f"MR({args.mr_pre},{args.mr_post}), coarsest CG {args.coarse_maxit} x tol {args.coarse_tol}, "
                      f"GCR({trunc}), graph-replayed cycle",
                "mg_setup_s": round(setup_s, 2), "mg_cycle_ms": round(cyc_ms, 3),
                "mg_gcr_batch_s": round(mdt, 4), "mg_lattice_solves_per_s": round(len(Ls) / mdt, 2),
                "mg_iterations": [int(mits.min()), int(mits.max())], "mg_max_true_residual": float(mtr.max())})
    del M
    return out


def cpu_baseline(args, U, B, mg, threads):
    """The oracle's FP64 MG-GCR solve of one right-hand side on the host cores, on
    the same hierarchy: the block vectors and coarse links of the GPU setup are
    copied into the oracle (its own setup, minutes at L=512, is not timed, as the
    GPU setup is not part of value).  A bounded sample: right-hand sides are solved
    one at a time (each capped at 30 iterations) until about 10 s have passed."""
    import oracle as O
    threads = O.set_threads(threads)   # the OpenMP runtime may predate any environment change
    L = args.L
    D = O.Dirac(L, U, args.mass, 1.0, 1.0, 0.0)
    A = D.op()
    nt = args.levels - 1
    blocks = ((4, 4),) + ((2, 2),) * (nt - 1)
    parts = (np.concatenate([mg.block_vecs(l).ravel() for l in range(nt)]),
             np.concatenate([mg.coarse_links(l).ravel() for l in range(nt)]))
    omg = O.MG(A, L, L, 2, None, n_levels=args.levels, block=blocks, k=(args.k,) * nt,
               pre=args.mr_pre, post=args.mr_post, w_cycle=bool(args.w_cycle),
               coarse_solver=1 if args.coarse_solver == 1 else 0,
               coarse_maxit=args.coarse_maxit, coarse_tol=args.coarse_tol, chiral=bool(args.chiral), parts=parts)
    its, tres, dt = [], [], 0.0
    for rhs in B:   # solves one after another until about 10 s of CPU work
        t0 = time.time()
        x, it, res = O.gcr(A, rhs, 30, 1e-10, args.trunc, omg, ref_projection=False)
        dt += time.time() - t0
        its.append(it)
        tres.append(np.linalg.norm(rhs - A(x)) / np.linalg.norm(rhs))
        if dt >= 10.0:
            break
    n = len(its)
    return {"value": round(n / dt, 5), "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": f"{n} right-hand side(s) solved one at a time, FP64 MG-GCR({args.trunc}) at L={L} on the same "
                      f"{args.levels}-level hierarchy (parts copied from the GPU setup), {sum(its)} iterations in "
                      f"{dt:.1f} s, max true residual {max(tres):.2e}"}


def build_mg(smg, args, D, L, seed):
    V = L * L
    dims = [2 * V]
    nx = L
    for lev in range(args.levels - 1):
        b = 4 if lev == 0 else 2
        nx //= b
        dims.append(nx * nx * args.k)
    nvec = args.k // 2 if args.chiral else args.k
    starts = np.concatenate([gaussian_vecs(nvec, n, seed + 10 + i).ravel() for i, n in enumerate(dims[:-1])])
    blocks = ((4, 4),) + ((2, 2),) * (args.levels - 2)
    return smg.MGHierarchy(D, starts, n_levels=args.levels, blocks=blocks, k=(args.k,) * (args.levels - 1),
                           null_iters=args.null_iters, pre_smooth=args.mr_pre, post_smooth=args.mr_post,
                           coarse_maxit=args.coarse_maxit, coarse_tol=args.coarse_tol,
                           coarse_half=2 if args.coarse_bf16 else 1,
                           coarse_solver=args.coarse_solver, use_graph=bool(args.graph), chiral=bool(args.chiral),
                           w_cycle=bool(args.w_cycle))


def hmc_workload(args):
    """Trajectories of the nested HMC: Omelyan with the fermion force on the outer scale
    (n_outer steps) and the gauge force on the fine scale (n_gauge Omelyan steps per
    outer position update), tau = 1, every solve the mixed-precision MG GCR to 1e-10
    inside FP64 refinement, FP64 dH and Metropolis; the hierarchy's Galerkin operators
    are rebuilt between trajectories.  The links start from heatbath-thermalised
    quenched configurations."""
    import torch
    import schwinger_mg_amd as smg
    L, V, seed = args.L, args.L * args.L, 2024
    D = smg.DiracOp(L, random_gauge(L, args.sigma, seed), args.hmc_mass, 1.0, 1.0, 0.0)
    t0 = time.time()
    smg.dirac_heatbath(D, args.beta, args.therm, seed)
    smg.synchronize()
    therm_s = time.time() - t0
    plaq0 = smg.avg_plaq(L, smg.gauge_of(D))
    log(f"heatbath {args.therm} sweeps {therm_s:.2f} s, plaquette {plaq0:.6f}")
    # the thermalised field at small mass wants a better near-null space and coarsest
    # solve than the random-field solve workload
    args.null_iters, args.coarse_maxit, args.coarse_tol = args.hmc_null_iters, args.hmc_coarse_maxit, \
        args.hmc_coarse_tol
    args.coarse_solver = 0   # the trajectories' solves are single-right-hand-side: coarsest CG
    t0 = time.time()
    mg = build_mg(smg, args, D, L, seed)
    smg.synchronize()
    setup_s = time.time() - t0
    log(f"MG setup {setup_s:.2f} s")
    p = smg.nested_params(args.beta, 1.0, args.n_outer, args.n_gauge, True, True, smg.F32, 500, 1e-10,
                          args.inner_tol, mg_refresh=args.mg_refresh, mg_refresh_iters=args.mg_refresh_iters,
                          chrono_guess=bool(args.chrono))
    X = None
    if args.lobpcg:   # the low modes, tracked as one block of right-hand sides
        t0 = time.time()
        lr = smg.lobpcg_update(D, gaussian_vecs(args.lobpcg, 2 * V, seed + 7), precond=mg, max_iter=10, tol=1e-8,
                               exact_ax=True)
        X = lr.eigvecs
        log(f"LOBPCG {args.lobpcg} modes from random starts: {time.time() - t0:.2f} s, {lr.iterations} iterations, "
            f"lowest {lr.eigvals[0]:.4e}")
    rng = np.random.default_rng(seed + 5)
    recs, times = [], []
    for t in range(args.warmup + args.steps):
        pi0 = rng.normal(size=2 * V)
        eta = rng.normal(size=2 * V) + 1j * rng.normal(size=2 * V)
        u = rng.uniform()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        res = smg.hmc_trajectory_nested(D, p, pi0, eta, u, mg=mg)
        if not args.mg_refresh:
            mg.rebuild()
        elif res.accepted:   # a rejected trajectory has already refreshed on the kept field
            mg.refresh(args.mg_refresh_iters)
        smg.synchronize()
        dt = time.perf_counter() - t1
        plaq = smg.avg_plaq(L, smg.gauge_of(D))
        rec = {"dH": res.dH, "accepted": res.accepted, "plaquette": plaq, "solver_iters": res.total_cg_iters,
               "max_true_residual": res.max_true_residual, "s": round(dt, 4)}
        if X is not None:   # warm-started from the previous trajectory's modes
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            lr = smg.lobpcg_update(D, X, precond=mg, max_iter=3, tol=1e-8, exact_ax=True)
            rec["lobpcg_s"] = round(time.perf_counter() - t1, 4)
            rec["lobpcg_iters"] = lr.iterations
            rec["lowest_modes"] = [float(v) for v in lr.eigvals[:4]]
            X = lr.eigvecs
        log(f"trajectory {t}: {rec}")
        if t >= args.warmup:
            recs.append(rec)
            times.append(dt)
    total = sum(times)
    dHs = np.array([r["dH"] for r in recs])
    out = {
        "metric": f"HMC trajectories/sec at L={L} (nested Omelyan, mixed-precision {args.levels}-level MG)",
        "value": round(len(recs) / total, 4),
        "unit": "trajectories/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(total / len(recs) * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 MD and dH; solves f32 (FP16 coarse links) in f64 refinement",
        "data": "synthetic",
        "config": {"workload": f"L={L} beta={args.beta} m={args.hmc_mass} c_sw=1, tau=1, Omelyan outer "
                               f"n={args.n_outer} (fermion force), Omelyan inner n={args.n_gauge} (gauge force), "
                               f"{args.therm} heatbath sweeps, {args.levels}-level MG ({args.null_iters} null "
                               f"iterations, coarsest CG to {args.coarse_tol:g}) GCR({args.trunc}) to 1e-10, "
                               + (f"hierarchy refreshed ({args.mg_refresh_iters} near-null iterations) every "
                                  f"{args.mg_refresh} outer position updates" if args.mg_refresh
                                  else "Galerkin rebuild between trajectories"),
                   "L": L, "beta": args.beta, "mass": args.hmc_mass, "n_outer": args.n_outer,
                   "n_gauge": args.n_gauge, "mg_refresh": args.mg_refresh,
                   "mg_refresh_iters": args.mg_refresh_iters},
        "acceptance": float(np.mean([r["accepted"] for r in recs])),
        "mean_dH": float(dHs.mean()),
        "mean_exp_minus_dH": float(np.exp(-dHs).mean()),
        "plaquette_start": plaq0,
        "plaquette_end": recs[-1]["plaquette"],
        "setup_s": round(setup_s, 3),
        "therm_s": round(therm_s, 3),
        "lobpcg_modes": args.lobpcg,
        "trajectories": recs,
    }
    print(json.dumps(out), flush=True)


def summarize_trace(d, range_name):
    """Markdown table of the kernels whose start lies inside each marker range named
    range_name, from a rocprofv3 --kernel-trace --marker-trace csv directory (per range:
    launches, ms and average us of every kernel)."""
    import collections
    import csv
    import glob
    kt, mk = [], []
    for p in glob.glob(os.path.join(d, "**", "*kernel_trace.csv"), recursive=True):
        kt += list(csv.DictReader(open(p)))
    for p in glob.glob(os.path.join(d, "**", "*marker_api_trace.csv"), recursive=True):
        mk += list(csv.DictReader(open(p)))
    rng = [(int(r["Start_Timestamp"]), int(r["End_Timestamp"])) for r in mk
           if range_name in r.get("Function", "") + r.get("Name", "") + r.get("Message", "")]
    agg = collections.defaultdict(lambda: [0, 0.0])
    busy = 0.0
    for r in kt:
        s0, e0 = int(r["Start_Timestamp"]), int(r["End_Timestamp"])
        if any(a <= s0 <= b for a, b in rng):
            g = agg[r["Kernel_Name"].split("(")[0].replace("void ", "")]
            g[0] += 1
            g[1] += (e0 - s0) / 1e3
            busy += (e0 - s0) / 1e6
    nb = max(1, len(rng))
    span = sum(b - a for a, b in rng) / 1e6 / nb
    out = [f"{len(rng)} ranges named {range_name}; per range: span {span:.2f} ms, kernel busy {busy / nb:.2f} ms, "
           f"{sum(v[0] for v in agg.values()) / nb:.0f} launches", "",
           "| kernel | launches / range | ms / range | average us |", "|---|---|---|---|"]
    for k, (n, us) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        out.append(f"| `{k}` | {n / nb:.0f} | {us / nb / 1e3:.2f} | {us / n:.1f} |")
    return "\n".join(out)


def summarize_kernels(d, pattern=None):
    """Markdown table of launches, average and minimum duration per kernel and grid from
    a rocprofv3 --kernel-trace csv directory."""
    import collections
    import csv
    import glob
    agg = collections.defaultdict(list)
    for p in glob.glob(os.path.join(d, "**", "*kernel_trace.csv"), recursive=True):
        for r in csv.DictReader(open(p)):
            k = (r["Kernel_Name"].split("(")[0].replace("void ", ""), r.get("Grid_Size", ""))
            if pattern and pattern not in k[0]:
                continue
            agg[k].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
    out = ["| kernel | grid | launches | average us | min us |", "|---|---|---|---|---|"]
    for (k, g), ts in sorted(agg.items()):
        out.append(f"| `{k}` | {g} | {len(ts)} | {sum(ts) / len(ts):.2f} | {min(ts):.2f} |")
    return "\n".join(out)


def summarize_pmc(dirs, pattern=None):
    """Markdown table of per-dispatch counter averages per kernel (and grid) from
    rocprofv3 --pmc csv directories; counters named *_SIZE are in KiB."""
    import collections
    import csv
    import glob
    agg = collections.defaultdict(lambda: collections.defaultdict(list))
    names = set()
    for d in dirs:
        for p in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            for r in csv.DictReader(open(p)):
                k = (r["Kernel_Name"].split("(")[0].replace("void ", ""), r.get("Grid_Size", ""))
                if pattern and pattern not in k[0]:
                    continue
                agg[k][r["Counter_Name"]].append(float(r["Counter_Value"]))
                names.add(r["Counter_Name"])
    cols = sorted(names)
    out = ["| kernel | grid | dispatches | " + " | ".join(cols) + " |", "|---|---|---|" + "---|" * len(cols)]
    for (k, g), cs in sorted(agg.items()):
        n = max(len(v) for v in cs.values())
        vals = [f"{sum(cs[c]) / len(cs[c]):.6g}" if cs.get(c) else "" for c in cols]
        out.append(f"| `{k}` | {g} | {n} | " + " | ".join(vals) + " |")
    return "\n".join(out)


def kernels_workload(args):
    """The hierarchy of the solve workload, tuned (or with the table of --tune-load), then
    20 launches of each timed kernel (the FP32 and FP64 stencils at N RHS, every coarse
    level's apply with dots in FP16 links / FP32 vectors, BF16 vectors at the coarsest,
    FP64; --kernels-select keeps those whose name contains the text), each inside a
    marker range of its name, and a plain 128 MiB copy whose counters calibrate the HBM
    byte counters.  Prints one line per kernel."""
    import torch
    import schwinger_mg_amd as smg
    L, V, N = args.L, args.L * args.L, args.nrhs
    D = smg.DiracOp(L, random_gauge(L, args.sigma, 2024), args.mass, 1.0, 1.0, 0.0)
    mg = build_mg(smg, args, D, L, 2024)
    if args.tune_load:
        smg.tune_load(args.tune_load)
    elif args.tune:
        smg.tune_mg(mg, N)
    if args.tune_save:
        smg.tune_save(args.tune_save)
    sel = args.kernels_select or ""
    for p in ("f32", "f64"):
        for name, kind in ((f"stencil_{p}", None), (f"ddagd_{p}", smg.DDAGD)):
            if sel not in name:
                continue
            marker(torch, name, True)
            s = time_stencil(smg, torch, D, p, N, V, True, launches=20, warm=2, same=False, kind=kind)
            marker(torch, name, False)
            print(f"{name} x{N}: {s['us_per_launch']:.1f} us, {s['GBs']:.0f} GB/s", flush=True)
    for lev in range(1, args.levels):
        for prec in (smg.F16, smg.F64) + ((smg.BF16V,) if lev == args.levels - 1 else ()):
            name = f"coarse_l{lev}_" + {smg.F16: "f16", smg.F64: "f64", smg.BF16V: "bf16v"}[prec]
            if sel not in name:
                continue
            marker(torch, name, True)
            c = time_coarse(smg, torch, mg, lev, prec, N, args.k, launches=20, warm=2)
            marker(torch, name, False)
            print(f"{name} ({c['prec']}): {c['us_per_launch']:.1f} us, {c['GBs']:.0f} GB/s, "
                  f"{c['TFLOPs']:.1f} TFLOP/s", flush=True)
        name = f"coarse_l{lev}_cheb"
        if lev == args.levels - 1 and args.coarse_solver == 2 and sel in name:
            marker(torch, name, True)
            c = time_coarse(smg, torch, mg, lev, smg.BF16V, N, args.k, launches=20, warm=2, cheb=True)
            marker(torch, name, False)
            print(f"{name} (Chebyshev step, {c['prec']}): {c['us_per_launch']:.1f} us, {c['GBs']:.0f} GB/s, "
                  f"{c['TFLOPs']:.1f} TFLOP/s", flush=True)
    a = torch.empty(1 << 25, dtype=torch.float32, device="cuda").uniform_()
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print("copy 128 MiB x3 done", flush=True)


def main():
    args = parse()
    if args.summarize_trace:
        print(summarize_trace(args.summarize_trace, args.range))
        return
    if args.summarize_pmc:
        print(summarize_pmc(args.summarize_pmc, args.kernel_filter))
        return
    if args.summarize_kernels:
        print(summarize_kernels(args.summarize_kernels, args.kernel_filter))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (every box here has one GPU)",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "hmc":
        hmc_workload(args)
        return
    if args.workload == "kernels":
        kernels_workload(args)
        return
    import torch
    import schwinger_mg_amd as smg

    L, V = args.L, args.L * args.L
    seed = 2024
    U = random_gauge(L, args.sigma, seed)
    D = smg.DiracOp(L, U, args.mass, 1.0, 1.0, 0.0)
    t0 = time.time()
    mg = build_mg(smg, args, D, L, seed)
    smg.synchronize()
    setup_s = time.time() - t0
    log(f"MG setup {setup_s:.2f} s")
    N = args.nrhs
    tune_s, tuned, table = None, None, None
    if args.tune and N in (16, 32, 64):
        t0 = time.time()
        tuned = smg.tune_mg(mg, N)
        tune_s = time.time() - t0
        import tempfile
        with tempfile.TemporaryDirectory() as td:   # the table as the library writes it
            p = os.path.join(td, "tune.txt")
            smg.tune_save(p)
            table = [l.strip() for l in open(p) if not l.startswith("#")]
        log(f"autotuner: {tuned} launch shapes in {tune_s:.2f} s: {table}")
    B = gaussian_vecs(N, 2 * V, seed + 1)
    Bd = torch.from_numpy(B).cuda()
    Xd = torch.zeros_like(Bd)
    info = smg.SolveInfo()
    tres = np.zeros(N)
    import ctypes

    def solve_batch(prec):
        smg._chk(smg.lib().smg_gcr_solve_batched_dev(D.h, mg.h, prec, N, Bd.data_ptr(), Xd.data_ptr(), 500, 1e-10,
                                                     min(args.trunc, 16), args.inner_tol,
                                                     tres.ctypes.data_as(ctypes.c_void_p), ctypes.byref(info), None))

    def timed(fn, steps, warm, name):
        for _ in range(warm):
            fn()
        torch.cuda.synchronize(); smg.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            marker(torch, name, True)    # a marker range per timed step (kernel traces select by it)
            fn()
            marker(torch, name, False)
        torch.cuda.synchronize(); smg.synchronize()
        return (time.perf_counter() - t0) / steps

    cheb = args.coarse_solver == 2

    def cycle_for(mixed_batched):
        """The mixed batched cycle: pre / post smoothing steps (Chebyshev with cheb_smooth)
        and, with coarse_solver 2, the coarsest Chebyshev iteration's cheb_steps; the FP64
        and single-right-hand-side cycles: MR(mr_pre, mr_post) and CG (coarse_maxit)."""
        if mixed_batched:
            mg.set_cycle(args.pre, args.post, args.cheb_steps if cheb else args.coarse_maxit, args.coarse_tol,
                         use_graph=bool(args.graph))
        else:
            mg.set_cycle(args.mr_pre, args.mr_post, args.coarse_maxit, args.coarse_tol, use_graph=bool(args.graph))

    if cheb and (args.cheb_lo_scale != 1.0 or args.cheb_hi_scale != 1.0):
        lo, hi = mg.cheb_interval()
        mg.set_cheb_interval(lo * args.cheb_lo_scale, hi * args.cheb_hi_scale)
    if args.cheb_smooth > 0:
        smooth_hi = mg.set_cheb_smoother(args.cheb_smooth)
        log(f"Chebyshev smoothing intervals: " + ", ".join(f"level {l} [{h / args.cheb_smooth:.4g}, {h:.4g}]"
                                                          for l, h in enumerate(smooth_hi)))
    cycle_for(True)
    t_mixed = timed(lambda: solve_batch(smg.F32), args.steps, args.warmup, "mixed_batch")
    mixed = {"iterations": info.iterations, "restarts": info.restarts, "max_true_residual": float(tres.max())}
    log(f"mixed x{N}: {t_mixed * 1e3:.1f} ms/batch = {N / t_mixed:.1f} solves/s, {mixed}")
    X_mixed = Xd.cpu().numpy()
    t_f64, fp64, err_vs_f64 = float("nan"), None, None
    cycle_for(False)
    fp64_cheb = None
    if args.fp64:
        if args.cheb_smooth > 0:
            mg.set_cheb_smoother(0.0)   # the plain FP64 solve: MR
        t_f64 = timed(lambda: solve_batch(smg.F64), args.fp64_steps, 1, "fp64_batch")
        fp64 = {"iterations": info.iterations, "max_true_residual": float(tres.max())}
        log(f"fp64 x{N}: {t_f64 * 1e3:.1f} ms/batch = {N / t_f64:.1f} solves/s, {fp64}")
        X_f64 = Xd.cpu().numpy()
        err_vs_f64 = float(np.max(np.linalg.norm(X_mixed - X_f64, axis=1) / np.linalg.norm(X_f64, axis=1)))
        if args.cheb_smooth > 0:
            # the FP64 solve with the mixed cycle's smoother (the coarsest solve stays CG)
            mg.set_cheb_smoother(args.cheb_smooth)
            mg.set_cycle(args.pre, args.post, args.coarse_maxit, args.coarse_tol, use_graph=bool(args.graph))
            t_f64c = timed(lambda: solve_batch(smg.F64), args.fp64_steps, 1, "fp64_cheb_batch")
            fp64_cheb = {"solves_per_s": round(N / t_f64c, 3), "iterations": info.iterations,
                         "max_true_residual": float(tres.max()),
                         "smoother": f"Chebyshev({args.pre},{args.post}) on [hi/{args.cheb_smooth:g}, hi]",
                         "speedup_mixed_vs_this": round(t_f64c / t_mixed, 3)}
            log(f"fp64 with the Chebyshev smoother x{N}: {fp64_cheb}")
            mg.set_cheb_smoother(0.0)
            cycle_for(False)

    # one right-hand side at a time (the reference's mode of use)
    b1 = Bd[0].contiguous()
    x1 = torch.zeros_like(b1)

    def solve_one(prec):
        smg._chk(smg.lib().smg_gcr_solve_dev(D.h, mg.h, prec, b1.data_ptr(), x1.data_ptr(), 500, 1e-10, args.trunc,
                                             args.inner_tol, ctypes.byref(info), None))

    one = None
    if args.single:
        t_one_mixed = timed(lambda: solve_one(smg.F32), 2, 1, "mixed_single")
        one = {"mixed_solves_per_s": round(1.0 / t_one_mixed, 3), "mixed_iterations": info.iterations,
               "true_residual": info.true_residual}
        t_one_f64 = timed(lambda: solve_one(smg.F64), 1, 1, "fp64_single")
        one["fp64_solves_per_s"] = round(1.0 / t_one_f64, 3)
        log(f"single rhs: {one}")

    stencil = [time_stencil(smg, torch, D, p, n, V, True) for p, n in
               [("f32", 1), ("f64", 1), ("f32", N), ("f64", N)]]
    ddagd = [dict(time_stencil(smg, torch, D, p, N, V, True, same=False, kind=smg.DDAGD), kernel="DdagD")
             for p in ("f32", "f64")]
    coarse = []
    if N in (16, 32, 64) and args.k in (8, 16, 32):
        for lev in range(1, args.levels):
            coarse.append(time_coarse(smg, torch, mg, lev, smg.F16, N, args.k))
        if args.coarse_bf16:
            coarse.append(time_coarse(smg, torch, mg, args.levels - 1, smg.BF16V, N, args.k, cheb=cheb))
        if args.fp64:
            for lev in range(1, args.levels):
                coarse.append(time_coarse(smg, torch, mg, lev, smg.F64, N, args.k))
        for c in coarse:
            log(f"coarse level {c['level']} ({c['prec']}): {c['us_per_launch']:.1f} us, {c['GBs']:.0f} GB/s, "
                f"{c['TFLOPs']:.1f} TFLOP/s")
    lbatch = None
    if args.lattice_batch:
        lbatch = lattice_batch(smg, torch, gaussian_vec, random_gauge, args)
        log(f"lattice batch: {lbatch}")
    log("stencil: " + ", ".join(f"{s['prec']}x{s['nrhs']} {s['GBs']:.0f} GB/s" for s in stencil))

    def line(name, t, traffic, per_cycle, mfma_peak=None):
        r = {"kernel": name, "bound": "hbm", "achieved": round(t["GBs"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
             "frac": round(t["GBs"] / HBM_PEAK_GBS, 4), "traffic": traffic,
             "algorithmic_bytes_per_launch": t["bytes_per_launch"], "avg_launch_us": round(t["us_per_launch"], 2),
             "launches_per_cycle": per_cycle, "ms_per_cycle": round(per_cycle * t["us_per_launch"] / 1e3, 3)}
        if mfma_peak:
            r["mfma_TFLOPs"] = round(t["TFLOPs"], 2)
            r["mfma_frac"] = round(t["TFLOPs"] / mfma_peak, 4)
        return r

    # the mixed solve's kernels; launches per V-cycle: D^dag D (one fused launch, or
    # two stencil launches) in the pre + post smoothing steps (Chebyshev or MR), the
    # post-smoothing's residual and the GCR apply, each coarse level's apply in its
    # smoothing steps and residual, the coarsest in every step.  The Chebyshev steps'
    # launches carry the vector updates in their epilogues; the lines time the plain apply
    fused = N in (16, 32, 64) and L % 16 == 0
    if fused:
        lines = [line(f"k_b_ddagd<float> (D^dag D in one launch, {N} RHS, L={L})", ddagd[0],
                      profiled_traffic(L, N, "ddagd_f32"), args.pre + args.post + 2)]
    else:
        lines = [line(f"k_b_dslash<float> (Wilson-clover D, {N} RHS, L={L})", stencil[2],
                      profiled_traffic(L, N, "dslash_f32"), 2 * (args.pre + args.post + 2))]
    last = args.levels - 1
    for c in coarse:
        if c["prec"].endswith("f32 vectors") and not (c["level"] == last and args.coarse_bf16):
            lines.append(line(f"k_b_coarse_bf16 (FP16 links, FP32 vectors, level {c['level']}, "
                              f"{c['sites']} sites, k={c['k']}, {N} RHS)", c,
                              profiled_traffic(L, N, "coarse", f"level{c['level']}_f16"),
                              args.pre + args.post + 1, PEAKS["bf16"]))
        elif c["prec"].endswith("bf16 vectors"):
            what = "Chebyshev step" if c["cheb"] else "CG"
            lines.append(line(f"k_b_coarse_bf16 (FP16 links, BF16 vectors, level {c['level']} coarsest {what}, "
                              f"{c['sites']} sites, k={c['k']}, {N} RHS)", c,
                              profiled_traffic(L, N, "coarse", f"level{c['level']}_" + ("cheb" if c["cheb"] else "bf16v")),
                              args.cheb_steps if c["cheb"] else args.coarse_maxit, PEAKS["bf16"]))
    fp64_lines = [line(f"k_b_coarse_mfma<double> (FP64, level {c['level']}, {c['sites']} sites, k={c['k']}, "
                       f"{N} RHS)", c, profiled_traffic(L, N, "coarse", f"level{c['level']}_f64"),
                       args.coarse_maxit if c["level"] == last else args.mr_pre + args.mr_post + 1,
                       PEAKS["f64"]) for c in coarse if c["prec"] == "f64"]
    roofline = max(lines, key=lambda r: r["ms_per_cycle"])   # the mixed solve's dominant kernel

    cpu = None
    if args.cpu_baseline:
        threads = args.cpu_threads or int(os.environ.get("OMP_NUM_THREADS", 0)) or min(16, os.cpu_count() or 1)
        log(f"cpu baseline on {threads} threads ...")
        cpu = cpu_baseline(args, U, B[:8], mg, threads)
        log(f"cpu baseline: {cpu}")

    out = {
        "metric": "MG-GCR solves/sec at L=512, 1e-10 resid: mixed vs FP64 GPU; Dslash GB/s",
        "value": round(N / t_mixed, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_mixed * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (FP16 coarse links" + (", BF16 coarsest vectors" if args.coarse_bf16 else "")
                 + ") in f64 refinement",
        "data": "synthetic",
        "config": {"workload": f"L={L} Wilson-clover m={args.mass} c_sw=1, theta~N(0,{args.sigma}^2), {N} RHS, "
                               f"{args.levels}-level MG (4x4 then 2x2 aggregates, "
                               + (f"{args.k // 2} null vectors split by chirality, k={args.k}" if args.chiral
                                  else f"k={args.k}")
                               + f", {'W' if args.w_cycle else 'V'}-cycle, "
                               + (f"Chebyshev({args.pre},{args.post}) smoothing on [hi/{args.cheb_smooth:g}, hi] in the "
                                  f"mixed cycle, MR({args.mr_pre},{args.mr_post}) in the FP64 cycle"
                                  if args.cheb_smooth > 0 else f"MR({args.pre},{args.post}) in the mixed cycle, "
                                  f"MR({args.mr_pre},{args.mr_post}) in the FP64 cycle")
                               + ", coarsest "
                               + ("dense inverse" if args.coarse_solver == 1
                                  else (f"{args.cheb_steps}-step Chebyshev in the mixed cycle, CG to "
                                        f"{args.coarse_tol:g} (<= {args.coarse_maxit} it) in the FP64 cycle" if cheb
                                        else f"CG to {args.coarse_tol:g} (<= {args.coarse_maxit} it)"))
                               + f"), GCR({args.trunc}), tol 1e-10 true residual",
                   "L": L, "nrhs": N, "levels": args.levels, "k": args.k, "pre": args.pre, "post": args.post,
                   "smoother": (f"Chebyshev on [hi/{args.cheb_smooth:g}, hi]" if args.cheb_smooth > 0 else "MR"),
                   "fp64_smoother": f"MR({args.mr_pre},{args.mr_post})",
                   "coarse_maxit": args.coarse_maxit, "coarse_tol": args.coarse_tol,
                   "coarse_solver": {0: "CG", 1: "dense inverse", 2: "Chebyshev (mixed), CG (FP64)"}[args.coarse_solver],
                   "cheb_steps": args.cheb_steps if cheb else None, "graph": bool(args.graph),
                   "coarsest_cg_vectors": "bf16" if args.coarse_bf16 else "fp32"},
        "mixed": mixed,
        "fp64_solves_per_s": round(N / t_f64, 3) if fp64 else None,
        "fp64": fp64,
        "speedup_mixed_vs_fp64": round(t_f64 / t_mixed, 3) if fp64 else None,
        "fp64_with_mixed_smoother": fp64_cheb,
        "single_rhs": one,
        "mixed_vs_fp64_solution_relerr": err_vs_f64,
        "setup_s": round(setup_s, 3),
        "tune_s": round(tune_s, 3) if tune_s is not None else None,
        "tuned_shapes": tuned,
        "tuned_table": table,
        "stencil": stencil,
        "ddagd": ddagd,
        "lattice_batch": lbatch,
        "roofline": roofline,
        "roofline_kernels": lines,
        "roofline_fp64_kernels": fp64_lines,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
