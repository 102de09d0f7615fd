"""Synthetic fields of the benchmark configurations (seeded, numpy PCG64)."""
import numpy as np


def bumps_field(dims, seed=0, nbumps=8, noise=1e-3, dtype=np.float32):
    """sin(2pi 3x) cos(2pi 5y) sin(2pi 2z) + Gaussian bumps + N(0, noise) on the
    unit cube, dims slowest-first (x along dims[0]). Draw order: A, w, c, noise."""
    rng = np.random.default_rng(seed)
    nd = len(dims)
    A = rng.uniform(0.5, 2.0, nbumps)
    w = rng.uniform(0.02, 0.1, nbumps)
    c = rng.uniform(0.0, 1.0, (nbumps, nd))
    axes = [np.linspace(0.0, 1.0, n) for n in dims]
    grids = np.meshgrid(*axes, indexing="ij")
    base = np.sin(2 * np.pi * 3 * grids[0])
    if nd > 1:
        base = base * np.cos(2 * np.pi * 5 * grids[1])
    if nd > 2:
        base = base * np.sin(2 * np.pi * 2 * grids[2])
    u = base.astype(np.float64)
    for b in range(nbumps):
        r2 = sum((g - c[b, d]) ** 2 for d, g in enumerate(grids))
        u += A[b] * np.exp(-r2 / (2 * w[b] ** 2))
    u += rng.normal(0.0, noise, u.shape)
    return u.astype(dtype).ravel()
