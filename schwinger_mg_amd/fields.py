"""Synthetic inputs for the Schwinger-model hot path.

All inputs are generated on the host with numpy's Philox4x32-10 counter-based
generator, so the CPU oracle and the GPU library consume byte-identical arrays.

Layouts follow the reference: gauge ``U[mu*V + s]`` with ``s = x + L*y``
(src/gauge.h:12, src/lattice.h:19); spinor ``psi[2*s + spin]`` (src/dirac.h:99).
"""
import numpy as np


def rng(seed, stream=0):
    return np.random.Generator(np.random.Philox(key=[int(seed), int(stream)]))


def random_gauge(L, sigma, seed, stream=0):
    """U = exp(i theta), theta ~ N(0, sigma^2) per link, complex128 [2*L*L]."""
    th = rng(seed, stream).normal(0.0, sigma, size=2 * L * L)
    return np.exp(1j * th)


def gaussian_vec(n, seed, stream=1):
    """re, im ~ N(0, 1) per entry (the distribution of src/linalg.cpp:44-49)."""
    g = rng(seed, stream).normal(0.0, 1.0, size=(n, 2))
    return g[:, 0] + 1j * g[:, 1]


def gaussian_vecs(count, n, seed, stream=2):
    g = rng(seed, stream).normal(0.0, 1.0, size=(count, n, 2))
    return g[..., 0] + 1j * g[..., 1]


def antiperiodic_y(U, L):
    """Fermion links for antiperiodic boundary conditions in y: U_y(x, L-1) -> -U_y.

    Every plaquette holds 0 or 2 flipped links, so plaquettes, the clover term and
    the gauge force are unchanged; the periodic reference operator fed these links
    is the antiperiodic operator.
    """
    V = L * L
    Uf = np.array(U, dtype=np.complex128, copy=True)
    Uf[V + (L - 1) * L: V + L * L] *= -1.0
    return Uf


def uniform_flux_gauge(L, n):
    """Links with every plaquette equal to exp(i*alpha), alpha = 2*pi*n/V.

    U_x(x, y) = exp(-i alpha y) for y < L-1 handled by a Landau-like gauge with a
    compensating twist on the boundary so that all V plaquettes are equal.
    """
    V = L * L
    alpha = 2.0 * np.pi * n / V
    U = np.ones(2 * V, np.complex128)
    for y in range(L):
        for x in range(L):
            s = x + L * y
            # U_y(x, y) = exp(i alpha x); at the x-boundary U_x(L-1, y) = exp(-i alpha L y)
            U[V + s] = np.exp(1j * alpha * x)
            if x == L - 1:
                U[s] = np.exp(-1j * alpha * L * y)
    return U, alpha
