"""Shared helpers for the parity tests."""
import numpy as np

CONFIGS = [(0.0, 0.0), (1.0, 0.0), (1.0, 0.1), (0.0, 0.1)]


def relerr(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def banner(msg):
    """Name the GPU case before it runs (flushed, so it survives a lost box)."""
    print(f"\n[gpu case] {msg}", flush=True)
