"""FP64 restatement of CAGR's orientation solve -- TEST INFRASTRUCTURE ONLY.

Follows CUDA-based/main_CAGR.py:42-45 (normalisation), :50-61 (b, widths),
:74-82 (width presets), :90-174 (the 40-iteration schedule: 31 steepest-descent
steps on the normal equations, the residual reset at i = N, 9 CG steps, then 8
WNNC G-steps with rescaling), with every treecode call done by the oracle in
FP64 (the reference runs the same schedule in FP32 on its CUDA module).
Reference quirks are kept: the wasted initial rk1 (:99-101), A_r summed into
one [N,1] before squaring (:114-116), the width changing inside the CG phase
(:103), lambda0 = 0.
"""
from __future__ import annotations

import numpy as np

from . import apply, build_tree

PRESETS = {"l0": (0.002, 0.016), "l1": (0.01, 0.04), "l2": (0.02, 0.08), "l3": (0.03, 0.12),
           "l4": (0.04, 0.16), "l5": (0.05, 0.2)}


def normalize(points_unnormalized: np.ndarray) -> np.ndarray:
    """main_CAGR.py:42-45, then the .float() cast of :47."""
    lo, hi = points_unnormalized.min(0), points_unnormalized.max(0)
    center = (lo + hi) / 2.0
    length = (hi - lo).max()
    return ((points_unnormalized - center) * (2 / (length * 1.1))).astype(np.float32)


def cagr_schedule(points: np.ndarray, wsmin: float, wsmax: float, iters: int = 40,
                  n_sd: int = 30, wnnc_steps: int = 8, leaf_size: int = 1, theta: float = 0.5,
                  an: float = 1.0):
    """Returns (normals [N,3], residual norms {'r': ||r||, 'rk1': ||rk1||}).
    an is the convection strength of A2k / A2Tk (the source's An = 1)."""
    p = np.asarray(points, np.float64)
    n = p.shape[0]
    tree = build_tree(p, 15, leaf_size)
    ones = np.ones(n)
    b = np.full((n, 1), 0.5)

    def A2(k, mu, ws):
        return apply(f"A2{k}", tree, p, mu, ones * ws, theta=theta, an=an)

    def A2T(k, s, ws):
        return apply(f"A2T{k}", tree, p, s, ones * ws, theta=theta, an=an)

    def G(mu, ws):
        return apply("G", tree, p, mu, ones * ws, theta=theta)

    normals = np.zeros((n, 3))
    ws = wsmin + 0.5 * (wsmax - wsmin)
    rk1 = A2T(1, b, ws) + A2T(2, b, ws) + A2T(3, b, ws)
    pk = rk1
    r = None
    for i in range(iters):
        ws = wsmin + ((iters - 1 - i) / (iters - 1)) * (wsmax - wsmin)
        if i <= n_sd:
            a1, a2, a3 = -A2(1, normals, ws), -A2(2, normals, ws), -A2(3, normals, ws)
            atam = A2T(1, a1, ws) + A2T(2, a2, ws) + A2T(3, a3, ws)
            r = A2T(1, b, ws) + A2T(2, b, ws) + A2T(3, b, ws) - atam - 0.0 * normals
            a_r = -(A2(1, r, ws) + A2(2, r, ws) + A2(3, r, ws))
            alpha = (r * r).sum() / (a_r * a_r).sum()
            normals = normals + alpha * r
            if i == n_sd:
                a1, a2, a3 = -A2(1, normals, ws), -A2(2, normals, ws), -A2(3, normals, ws)
                atam = A2T(1, a1, ws) + A2T(2, a2, ws) + A2T(3, a3, ws)
                rk1 = A2T(1, b, ws) + A2T(2, b, ws) + A2T(3, b, ws) - atam - 0.0 * normals
                pk = rk1
        else:
            p1, p2, p3 = -A2(1, pk, ws), -A2(2, pk, ws), -A2(3, pk, ws)
            t1, t2, t3 = A2T(1, p1, ws), A2T(2, p2, ws), A2T(3, p3, ws)
            alpha = (rk1 * rk1).sum() / ((t3 * pk) + (t2 * pk) + (t1 * pk)).sum()
            normals = normals + alpha * pk
            rk20 = rk1 - alpha * t1 - alpha * t2 - alpha * t3
            beta = (rk20 * rk20).sum() / (rk1 * rk1).sum()
            pk = rk20 + beta * pk
            rk1 = rk20
    res = {"r": float(np.linalg.norm(r)), "rk1": float(np.linalg.norm(rk1))}
    for _ in range(wnnc_steps):
        out = G(normals, ws)
        out = out / np.maximum(np.linalg.norm(out, axis=1, keepdims=True), 1e-12)
        normals = out * np.linalg.norm(normals, axis=1, keepdims=True)
    return normals, res
