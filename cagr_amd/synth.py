"""Seeded synthetic point clouds of the benchmark configurations.

Every generator returns float64 numpy arrays; ``normalize`` applies the
reference's normalisation (CUDA-based/main_CAGR.py:42-45): centre the
bounding box, scale its longest side to 2/1.1.
"""
from __future__ import annotations

import numpy as np


def normalize(points: np.ndarray) -> np.ndarray:
    lo, hi = points.min(0), points.max(0)
    center = (lo + hi) / 2.0
    length = (hi - lo).max()
    return (points - center) * (2.0 / (length * 1.1))


def sphere(n: int, seed: int = 0, sigma: float = 0.002) -> np.ndarray:
    """Uniform on the unit sphere plus N(0, sigma^2) noise per coordinate."""
    rng = np.random.default_rng(seed)
    g = rng.normal(size=(n, 3))
    p = g / np.linalg.norm(g, axis=1, keepdims=True)
    return p + rng.normal(scale=sigma, size=p.shape)


def torus(n: int, seed: int = 0, R: float = 1.0, r: float = 0.3, sigma: float = 0.001,
          outliers: float = 0.05) -> np.ndarray:
    """Area-uniform torus samples (rejection on the tube angle), noise, and a
    fraction of points replaced by uniform outliers in the bounding box."""
    rng = np.random.default_rng(seed)
    out = np.empty((0, 3))
    while out.shape[0] < n:
        m = 2 * (n - out.shape[0]) + 64
        th = rng.uniform(0, 2 * np.pi, m)
        ph = rng.uniform(0, 2 * np.pi, m)
        keep = rng.uniform(0, 1, m) < (R + r * np.cos(ph)) / (R + r)
        th, ph = th[keep], ph[keep]
        p = np.stack([(R + r * np.cos(ph)) * np.cos(th), (R + r * np.cos(ph)) * np.sin(th),
                      r * np.sin(ph)], 1)
        out = np.concatenate([out, p])
    out = out[:n] + rng.normal(scale=sigma, size=(n, 3))
    k = int(round(outliers * n))
    if k:
        idx = rng.choice(n, k, replace=False)
        lo, hi = out.min(0), out.max(0)
        out[idx] = rng.uniform(lo, hi, size=(k, 3))
    return out


def _ellipsoid_area(a, b, c, p=1.6075):
    return 4 * np.pi * (((a * b) ** p + (a * c) ** p + (b * c) ** p) / 3) ** (1 / p)


def _ellipsoid_samples(rng, n, axes, center):
    a, b, c = axes
    out = np.empty((0, 3))
    gmax = max(a * b, a * c, b * c)
    while out.shape[0] < n:
        m = 2 * (n - out.shape[0]) + 64
        g = rng.normal(size=(m, 3))
        u = g / np.linalg.norm(g, axis=1, keepdims=True)
        # area element ratio of the map sphere -> ellipsoid, for rejection
        dens = np.sqrt((b * c * u[:, 0]) ** 2 + (a * c * u[:, 1]) ** 2 + (a * b * u[:, 2]) ** 2)
        keep = rng.uniform(0, 1, m) < dens / gmax
        out = np.concatenate([out, u[keep] * np.array([a, b, c])])
    return out[:n] + center


def ellipsoids(n: int, seed: int = 0, k: int = 6, sigma: float = 5e-4) -> np.ndarray:
    """Union of k seeded ellipsoids (semi-axes U[0.2,0.6], centres U[-0.5,0.5]^3),
    points split by surface area, area-uniform on each, plus noise."""
    rng = np.random.default_rng(seed)
    axes = rng.uniform(0.2, 0.6, size=(k, 3))
    centers = rng.uniform(-0.5, 0.5, size=(k, 3))
    areas = np.array([_ellipsoid_area(*ax) for ax in axes])
    counts = np.floor(n * areas / areas.sum()).astype(int)
    counts[: n - counts.sum()] += 1
    parts = [_ellipsoid_samples(rng, int(cn), axes[i], centers[i]) for i, cn in enumerate(counts)]
    p = np.concatenate(parts)
    return p + rng.normal(scale=sigma, size=p.shape)


def displaced_sphere(n: int, seed: int = 0, terms: int = 4, amp: float = 0.1,
                     sigma: float = 0.0) -> np.ndarray:
    """r(u) = 1 + amp * sum_k sin(f_k * polar(u) + phi_k), f_k in {2..6}."""
    rng = np.random.default_rng(seed)
    f = rng.integers(2, 7, size=terms)
    phi = rng.uniform(0, 2 * np.pi, size=terms)
    g = rng.normal(size=(n, 3))
    u = g / np.linalg.norm(g, axis=1, keepdims=True)
    polar = np.arccos(np.clip(u[:, 2], -1, 1))
    r = 1 + amp * np.sin(f[None, :] * polar[:, None] + phi[None, :]).sum(1)
    p = u * r[:, None]
    if sigma:
        p = p + rng.normal(scale=sigma, size=p.shape)
    return p


def batch_sizes(count: int = 64, seed: int = 0, lo: float = 2e3, hi: float = 2e5) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return np.round(np.exp(rng.uniform(np.log(lo), np.log(hi), count))).astype(np.int64)


def batch_clouds(count: int = 64, seed: int = 0, sigma: float = 0.001):
    """count clouds, sizes log-uniform in [2e3, 2e5], shape sphere/torus/ellipsoid by seed."""
    sizes = batch_sizes(count, seed)
    rng = np.random.default_rng(seed + 1)
    clouds = []
    for i, n in enumerate(sizes):
        kind = int(rng.integers(0, 3))
        s = int(rng.integers(0, 2**31 - 1))
        if kind == 0:
            p = sphere(int(n), s, sigma)
        elif kind == 1:
            p = torus(int(n), s, sigma=sigma, outliers=0.0)
        else:
            p = ellipsoids(int(n), s, k=1, sigma=sigma)
        clouds.append(normalize(p))
    return clouds


def random_normals(n: int, seed: int = 0) -> np.ndarray:
    """Unit random directions times U[0.5, 1.5] magnitudes (non-zero weights)."""
    rng = np.random.default_rng(seed)
    g = rng.normal(size=(n, 3))
    return g / np.linalg.norm(g, axis=1, keepdims=True) * rng.uniform(0.5, 1.5, size=(n, 1))
