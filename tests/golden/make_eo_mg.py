"""Writes tests/golden/eo_mg_vectors.npz from the reference's own EvenOddMG
(src/eo_multigrid.h), compiled from its sources by oracle/Makefile into
oracle/_ref/libref.so: for seeded links and right-hand sides, the Richardson weight,
the dense Galerkin matrix and one preconditioner application.  Run in the container
that holds the reference:

    make -C oracle && python tests/golden/make_eo_mg.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle.ref as R  # noqa: E402
from schwinger_mg_amd.fields import gaussian_vec, random_gauge  # noqa: E402

# (L, csw, mu, bx, by, k, null_iters, pre, post, seed)
CASES = [(16, 1.0, 0.0, 4, 4, 4, 10, 3, 3, 7), (16, 0.0, 0.1, 4, 2, 3, 8, 2, 4, 8), (24, 1.0, 0.05, 4, 4, 6, 5, 3, 3, 9)]


def main():
    assert R.available(), "oracle/_ref/libref.so is not built"
    out = {}
    for i, (L, csw, mu, bx, by, k, ni, pre, post, seed) in enumerate(CASES):
        V = L * L
        U = random_gauge(L, 0.5, seed=100 + i)
        b = gaussian_vec(V, seed=200 + i)
        x, om, M = R.eo_mg(L, U, 0.1, 1.0, csw, mu, bx, by, k, ni, pre, post, seed, b)
        out[f"c{i}_params"] = np.array([L, csw, mu, bx, by, k, ni, pre, post, seed], np.float64)
        # the start vectors the build draws: random_vec(V, rng) k times from mt19937(seed),
        # and estimate_lambda_max's random_vec(V, mt19937(12345))
        out[f"c{i}_starts"] = R.random_vec(k * V, seed).reshape(k, V)
        out[f"c{i}_lam"] = R.random_vec(V, 12345)
        out[f"c{i}_U"], out[f"c{i}_b"], out[f"c{i}_x"] = U, b, x
        out[f"c{i}_omega"], out[f"c{i}_mat"] = np.array([om]), M
        print(f"case {i}: L={L} dim={M.shape[0]} omega={om:.6e} |x|={np.linalg.norm(x):.6e}")
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "eo_mg_vectors.npz"), **out)


if __name__ == "__main__":
    main()
