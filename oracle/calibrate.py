"""Times the oracle (the FP64 CPU port that bench.py's cpu_baseline leg runs) beside
the compiled reference (oracle/_ref/libref.so, built from the reference's own
sources) on the same inputs: CG on D^dag D to 1e-10 (per iteration), 50 MR sweeps on
D^dag D, and the clover fermion force, at L = 16, 32, 64.  The ratio port/reference
says how far the CPU baseline's timing stands for the reference's own code.

TEST INFRASTRUCTURE ONLY; runs where the reference sources are (not on the GPU box).
The thread count is OMP_NUM_THREADS, read when the libraries start, so run one
process per count:

    make -C oracle && OMP_NUM_THREADS=1 python oracle/calibrate.py
    OMP_NUM_THREADS=8 python oracle/calibrate.py
"""
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
import oracle.ref as R  # noqa: E402
from schwinger_mg_amd.fields import gaussian_vec, random_gauge  # noqa: E402


def median_time(f, reps):
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        f()
        ts.append(time.perf_counter() - t0)
    return statistics.median(ts)


def main():
    assert R.available(), "oracle/_ref/libref.so is not built"
    threads = O.set_threads(int(os.environ.get("OMP_NUM_THREADS", "1")))
    print(f"threads {threads} (the oracle and the reference share the OpenMP runtime)")
    print("| L | operation | reference | port | port / reference |")
    print("|---|---|---|---|---|")
    for L in (16, 32, 64, 256):
        V = L * L
        U = random_gauge(L, 0.4, 10 + L)
        b = gaussian_vec(2 * V, 20 + L)
        x0 = gaussian_vec(2 * V, 30 + L)
        Dr = R.Dirac(L, U, 0.1, 1.0, 1.0, 0.0)
        Do = O.Dirac(L, U, 0.1, 1.0, 1.0, 0.0)
        A = Do.op()
        reps = 7 if L < 64 else 3
        _, itr, _ = Dr.krylov(0, b, 5000, 1e-10)
        _, ito, _ = O.cg(A, b, 5000, 1e-10)
        tr = median_time(lambda: Dr.krylov(0, b, 5000, 1e-10), reps) / itr
        to = median_time(lambda: O.cg(A, b, 5000, 1e-10), reps) / ito
        print(f"| {L} | CG on D^dag D, per iteration ({itr} / {ito} iterations) | {tr * 1e6:.1f} us | "
              f"{to * 1e6:.1f} us | {to / tr:.2f} |")
        tr = median_time(lambda: Dr.mr_smooth(x0, b, 50), reps)
        to = median_time(lambda: O.mr_smooth(A, x0, b, 50), reps)
        print(f"| {L} | 50 MR sweeps on D^dag D | {tr * 1e3:.2f} ms | {to * 1e3:.2f} ms | {to / tr:.2f} |")
        tr = median_time(lambda: [Dr.force(0, x0) for _ in range(10)], reps) / 10
        to = median_time(lambda: [Do.force(x0) for _ in range(10)], reps) / 10
        print(f"| {L} | fermion force (clover) | {tr * 1e6:.1f} us | {to * 1e6:.1f} us | {to / tr:.2f} |")


if __name__ == "__main__":
    main()
