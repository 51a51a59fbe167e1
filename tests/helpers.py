"""Shared helpers for the parity tests."""
import numpy as np


def rel_max(a, b):
    """max |a - b| / max |b| (the exitance error measure of the benchmark)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / (den if den > 0 else 1.0))


def hetero_maps(shape, seed, mua=0.01, mus=1.0, sigma=0.3):
    rng = np.random.default_rng(seed)
    N = int(np.prod(shape))
    return mua * np.exp(sigma * rng.standard_normal(N)), mus * np.exp(sigma * rng.standard_normal(N))


def bounds_for(shape, d=(0.1, 0.1, 0.1), z0=0.0):
    nx, ny, nz = shape
    return (-nx * d[0] / 2, nx * d[0] / 2, -ny * d[1] / 2, ny * d[1] / 2, z0, z0 + nz * d[2])
