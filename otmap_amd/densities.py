"""Seeded synthetic densities for the five benchmark/parity workloads.

All generators return a flat float64 array of n*n cell values in the
reference's face order ``j + i*n`` (i = x index, j = y index; cell centre
((i+1/2)/n, (j+1/2)/n)).  An image ``img[row, col]`` maps to the solver input
through column-major vectorisation (reference ``apps/common/eigen_addons.h``
``vec``), i.e. face (i=col, j=row); the generators here build the field
directly in face order.  Randomness comes from numpy's PCG64 with a fixed
seed per workload, so inputs are bit-identical on every machine with the same
numpy.

Workloads:
  gaussian_mixture   K anisotropic Gaussians (centres U[0.2,0.8]^2, sigmas
                     U[0.03,0.15], angle U[0,pi), weights U[0.5,2]),
                     + 1e-3*max floor, mean 1.                 (256^2 single solve)
  smoothed_shapes    union of 3 discs/rings (radii U[0.05,0.3], ring inner
                     radius r*U[0.5,0.8]), Gaussian blur sigma=5px, +1e-3*max
                     floor, mean 1.                            (1024^2 solve, 512^2 barycentres)
  batch_specs        64 (n, K) pairs, n in {128,...,1024}, K in [2,12]
  hdr_lognormal      exp(2*G) * mixture(K=5), G = white noise blurred at 16px
                     and scaled to unit variance, contrast capped at 1e4,
                     + 1e-4*max floor, mean 1.                 (2048^2 stippling)
"""
from __future__ import annotations

import numpy as np

SEED_GAUSSIAN_256 = 1000
SEED_SHAPES_1024 = 1001
SEED_BATCH = 1002
SEED_BARYCENTRE = 1003
SEED_HDR_2048 = 1004
BATCH_SIZES = (128, 192, 256, 384, 512, 768, 1024)


def _centres(n):
    c = (np.arange(n, dtype=np.float64) + 0.5) / n
    X, Y = np.meshgrid(c, c, indexing="ij")  # X[i, j] = x of face (i, j)
    return X, Y


def _finish(rho, floor_rel):
    rho = rho + floor_rel * rho.max()
    rho = rho / rho.mean()
    return np.ascontiguousarray(rho.reshape(-1), dtype=np.float64)


def _mixture(n, rng, K):
    X, Y = _centres(n)
    rho = np.zeros((n, n))
    for _ in range(K):
        cx, cy = rng.uniform(0.2, 0.8, size=2)
        s1, s2 = rng.uniform(0.03, 0.15, size=2)
        th = rng.uniform(0.0, np.pi)
        w = rng.uniform(0.5, 2.0)
        c, s = np.cos(th), np.sin(th)
        dx, dy = X - cx, Y - cy
        u = c * dx + s * dy
        v = -s * dx + c * dy
        rho += w * np.exp(-0.5 * ((u / s1) ** 2 + (v / s2) ** 2))
    return rho


def gaussian_mixture(n: int = 256, seed: int = SEED_GAUSSIAN_256, K: int = 5, floor_rel: float = 1e-3):
    rng = np.random.default_rng(seed)
    return _finish(_mixture(n, rng, K), floor_rel)


def gaussian_kernel(sigma: float, truncate: float = 3.0):
    r = int(np.ceil(truncate * sigma))
    x = np.arange(-r, r + 1, dtype=np.float64)
    k = np.exp(-0.5 * (x / sigma) ** 2)
    return k / k.sum()


def blur(img, sigma: float, truncate: float = 3.0):
    """Separable Gaussian blur, mirror borders (edge sample repeated)."""
    k = gaussian_kernel(sigma, truncate)
    r = (k.size - 1) // 2
    out = np.asarray(img, dtype=np.float64)
    for axis in (0, 1):
        pad = [(0, 0), (0, 0)]
        pad[axis] = (r, r)
        p = np.pad(out, pad, mode="symmetric")
        acc = np.zeros_like(out)
        m = out.shape[axis]
        for t in range(k.size):
            sl = [slice(None), slice(None)]
            sl[axis] = slice(t, t + m)
            acc += k[t] * p[tuple(sl)]
        out = acc
    return out


def _shapes(n, rng, count=3):
    X, Y = _centres(n)
    ind = np.zeros((n, n), dtype=bool)
    for _ in range(count):
        cx, cy = rng.uniform(0.2, 0.8, size=2)
        r = rng.uniform(0.05, 0.3)
        ring = rng.uniform() < 0.5
        inner = r * rng.uniform(0.5, 0.8)
        d = np.hypot(X - cx, Y - cy)
        ind |= (d <= r) & ((d >= inner) if ring else True)
    return ind.astype(np.float64)


def smoothed_shapes(n: int = 1024, seed: int = SEED_SHAPES_1024, sigma_px: float = 5.0, floor_rel: float = 1e-3):
    rng = np.random.default_rng(seed)
    ind = _shapes(n, rng)
    return _finish(blur(ind, sigma_px), floor_rel)


def barycentre_inputs(n: int = 512, seed: int = SEED_BARYCENTRE, count: int = 4):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        ind = _shapes(n, rng)
        out.append(_finish(blur(ind, 5.0), 1e-3))
    return out


def batch_specs(count: int = 64, seed: int = SEED_BATCH):
    rng = np.random.default_rng(seed)
    ns = rng.choice(np.array(BATCH_SIZES), size=count)
    Ks = rng.integers(2, 13, size=count)
    seeds = rng.integers(0, 2**31 - 1, size=count)
    return [(int(n), int(k), int(s)) for n, k, s in zip(ns, Ks, seeds)]


def batch_densities(count: int = 64, seed: int = SEED_BATCH):
    return [gaussian_mixture(n, s, K) for n, K, s in batch_specs(count, seed)]


def hdr_lognormal(n: int = 2048, seed: int = SEED_HDR_2048, corr_px: float = 16.0, contrast: float = 1e4,
                  floor_rel: float = 1e-4):
    rng = np.random.default_rng(seed)
    noise = rng.standard_normal((n, n))
    g = blur(noise, corr_px)
    g = (g - g.mean()) / g.std()
    rho = np.exp(2.0 * g) * _mixture(n, rng, 5)
    rho = np.maximum(rho, rho.max() / contrast)
    return _finish(rho, floor_rel)


def uniform_points(m: int, seed: int):
    """m query points uniform in [0,1]^2, interleaved (m, 2)."""
    rng = np.random.default_rng(seed)
    return rng.uniform(0.0, 1.0, size=(m, 2))
