"""Winding-number grid query of the reconstruction step.

Restates the query half of CUDA-based/ext/gaussrecon_src/main_GaussRecon_cuda.cu:
* sample attributes: unit normal * area with area = 1e-5, so the point weights
  |mu| equal the area (:266-290);
* the A3 operator (the mean of the three convection kernels) with the
  continuous kernel inside the width, at the query points (:427-445) and at
  the samples themselves (:450-467);
* value = -wn at every query, smoothWidth = width / 2 (:477-480); isovalue =
  -wn at the samples, element N/2 after nth_element (:485-486).
The reference queries the corners of its adaptive PGR octree, which
``cagr_amd.pgr`` builds (and meshes); the functions here take a regular grid
(generated inside the kernel) or any caller-given points.  The points are
expected in the [-1, 1]^3 convention (the reference maps its [0,1]^3 PGR
coordinates with 2p - 1).
"""
from __future__ import annotations

from typing import Optional, Sequence

import torch

from .wn_treecode import WindingNumberTreecode


def sample_attributes(normals: torch.Tensor, area: float = 1e-5) -> torch.Tensor:
    """unit normal * area (main_GaussRecon_cuda.cu:276-286)."""
    nlen = torch.linalg.norm(normals, dim=-1, keepdim=True).clamp_min(1e-12)
    return (normals / nlen * area).contiguous()


def isovalue(wt: WindingNumberTreecode, mu: torch.Tensor, width: float, *, op: str = "A3",
             precision: Optional[str] = None) -> float:
    """-wn at the samples, element N/2 in ascending order (nth_element)."""
    n = wt.points.shape[0]
    widths = torch.full((n,), width, dtype=mu.dtype, device=mu.device)
    wn = wt.apply(op, mu, widths, continuous=True, precision=precision)[:, 0]
    kth = torch.kthvalue(wn.double().cpu(), n // 2 + 1).values
    return float(-kth)


def grid_values(wt: WindingNumberTreecode, mu: torch.Tensor, lo: Sequence[float],
                hi: Sequence[float], dims: Sequence[int], width: float, *, op: str = "A3",
                precision: Optional[str] = None) -> torch.Tensor:
    """-wn on the regular grid linspace(lo, hi, dims) per axis -> [d0, d1, d2]."""
    step = [(h - l) / (d - 1) if d > 1 else 0.0 for l, h, d in zip(lo, hi, dims)]
    return -wt.grid_query(mu, lo, step, dims, width, op=op, continuous=True, precision=precision)


def reconstruct_field(points: torch.Tensor, normals: torch.Tensor, res: int = 128,
                      bound: float = 1.2, width: float = 0.01, area: float = 1e-5,
                      leaf_size: int = 32, theta: float = 0.5, precision: Optional[str] = None):
    """Build the tree over the oriented samples, then return (grid values
    [res,res,res], isovalue, smooth width) as GaussRecon hands them to
    marching cubes."""
    wt = WindingNumberTreecode(points, leaf_size=leaf_size, theta=theta, precision=precision)
    mu = sample_attributes(normals.to(points.dtype), area)
    vals = grid_values(wt, mu, [-bound] * 3, [bound] * 3, [res] * 3, width, precision=precision)
    iso = isovalue(wt, mu, width, precision=precision)
    return vals, iso, width / 2
