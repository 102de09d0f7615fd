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


def fourier_field_gpu(dims, seed=2, nmodes=32, kmax=64, noise=1e-8, device="cuda"):
    """Sum of nmodes random Fourier modes on the unit box, float64, built on
    the GPU (torch). Integer wavevectors with |k| drawn from a k^(-5/3) law
    on [1, kmax] (inverse CDF), random directions and phases, amplitudes
    |k|^(-5/6), plus N(0, noise). Draw order: |k|, directions, phases, then
    the noise from torch's generator seeded with `seed`."""
    import torch
    rng = np.random.default_rng(seed)
    a = -5.0 / 3.0 + 1.0
    r = rng.random(nmodes)
    kmag = ((kmax ** a - 1.0) * r + 1.0) ** (1.0 / a)
    dirs = rng.normal(size=(nmodes, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    kvec = np.rint(dirs * kmag[:, None])
    kvec[np.all(kvec == 0, axis=1), 0] = 1.0
    phase = rng.uniform(0, 2 * np.pi, nmodes)
    amp = np.linalg.norm(kvec, axis=1) ** (-5.0 / 6.0)
    axes = [torch.linspace(0.0, 1.0, n, dtype=torch.float64, device=device) for n in dims]
    u = torch.zeros(dims, dtype=torch.float64, device=device)
    for m in range(nmodes):
        ph = (2 * np.pi * kvec[m, 0] * axes[0])[:, None, None] + (2 * np.pi * kvec[m, 1] * axes[1])[None, :, None] \
            + (2 * np.pi * kvec[m, 2] * axes[2])[None, None, :] + phase[m]
        u += amp[m] * torch.sin(ph)
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    u += noise * torch.randn(dims, dtype=torch.float64, device=device, generator=g)
    return u.reshape(-1)
