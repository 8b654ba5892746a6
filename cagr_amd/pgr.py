"""GaussRecon's PGR adaptive octree and the winding-number values on its corners.

``PGROctree`` wraps the HIP builder ``cagr_pgr_build``: the reference's
``Octree::setTree`` (CUDA-based/ext/gaussrecon_src/Octree.cpp:77-400), which
maps the samples into [0,1]^3, grows the octree by PoissonRecon-style sample
splatting, clips it, and lists the corners of its leaves (``gridDataVector``).

``gauss_recon_values`` is the rest of main_GaussRecon_cuda.cu up to marching
cubes (:255-486): samples and corners mapped to [-1,1]^3 with 2p - 1, the
treecode over the samples with attributes unit normal * 1e-5, the A3 operator
(continuous kernel) at the corners and at the samples, value = -wn at every
corner, smoothWidth = width / 2 and isovalue = -wn at the samples, element
N/2 after nth_element.
"""
from __future__ import annotations

import ctypes as ct
from typing import Optional

import torch

from ._lib import check, load
from .query import sample_attributes
from .wn_treecode import WindingNumberTreecode


def _stream(device) -> ct.c_void_p:
    return ct.c_void_p(torch.cuda.current_stream(device).cuda_stream)


class PGROctree:
    """The PGR octree over raw sample coordinates ``points`` [N,3] (float32, on
    a HIP device), as ``Octree::setTree(file, max_depth, min_depth)`` builds it.

    Attributes: ``samples`` [N,3] in [0,1]^3 (``Octree::samplePoints``),
    ``corners`` [G,3] in [0,1]^3 in ``gridDataVector`` order, ``leaves`` [L,4]
    int32 (depth, x, y, z) in preorder, ``depth`` (``Octree::maxDepth`` after
    the build), ``scale`` (``maxScale``) and ``origin`` [3].
    """

    def __init__(self, points: torch.Tensor, max_depth: int = 10, min_depth: int = 1):
        if points.dim() != 2 or points.shape[1] != 3:
            raise AssertionError("points must be [N, 3]")
        if points.dtype != torch.float32:
            raise TypeError("the PGR octree reads float32 samples (fscanf %f)")
        if points.device.type != "cuda":
            raise ValueError("points must be on a HIP device")
        L = load()
        dev = points.device
        pts = points.contiguous()
        h = ct.c_void_p()
        with torch.cuda.device(dev):
            check(L.cagr_pgr_build(ct.c_void_p(pts.data_ptr()), pts.shape[0], int(max_depth),
                                   int(min_depth), _stream(dev), ct.byref(h)))
        try:
            nl, nc, dep = ct.c_int64(), ct.c_int64(), ct.c_int()
            sc = (ct.c_float * 4)()
            check(L.cagr_pgr_info(h, ct.byref(nl), ct.byref(nc), ct.byref(dep), sc))
            self.samples = torch.empty_like(pts)
            self.corners = torch.empty((nc.value, 3), dtype=torch.float32, device=dev)
            self.leaves = torch.empty((nl.value, 4), dtype=torch.int32, device=dev)
            with torch.cuda.device(dev):
                check(L.cagr_pgr_read(h, ct.c_void_p(self.samples.data_ptr()),
                                      ct.c_void_p(self.corners.data_ptr()),
                                      ct.c_void_p(self.leaves.data_ptr()), _stream(dev)))
        finally:
            L.cagr_pgr_destroy(h)
        self.depth = dep.value
        self.scale = float(sc[0])
        self.origin = [float(sc[1]), float(sc[2]), float(sc[3])]

    def query_points(self) -> torch.Tensor:
        """Corners in the [-1,1]^3 convention (main_GaussRecon_cuda.cu:294-298)."""
        return 2 * self.corners - 1

    def sample_points(self) -> torch.Tensor:
        """Samples in the [-1,1]^3 convention (main_GaussRecon_cuda.cu:272-274)."""
        return 2 * self.samples - 1


def gauss_recon_values(points: torch.Tensor, normals: torch.Tensor, *, width: float = 0.01,
                       max_depth: int = 10, min_depth: int = 1, area: float = 1e-5,
                       leaf_size: int = 1, theta: float = 0.5,
                       precision: Optional[str] = None):
    """main_GaussRecon_cuda.cu:255-486 before marching cubes.

    Returns ``(octree, values [G], isovalue, smooth_width)``: -wn at every
    corner of the PGR octree in ``gridDataVector`` order, -wn at the samples
    (element N/2 after nth_element) and width / 2.  Defaults are the
    reference's (-d 10, -m 1, -w 0.01, leaf size 1, θ = 0.5).
    """
    oct_ = PGROctree(points, max_depth, min_depth)
    pts = oct_.sample_points()
    mu = sample_attributes(normals.to(device=pts.device, dtype=pts.dtype), area)
    wt = WindingNumberTreecode(pts, leaf_size=leaf_size, theta=theta, precision=precision)
    wn_q = wt.query(oct_.query_points(), mu, width, precision=precision)[:, 0]
    n = pts.shape[0]
    widths = torch.full((n,), width, dtype=pts.dtype, device=pts.device)
    wn_s = wt.apply("A3", mu, widths, continuous=True, precision=precision)[:, 0]
    kth = torch.kthvalue(wn_s.double().cpu(), n // 2 + 1).values
    return oct_, -wn_q, float(-kth), width / 2
