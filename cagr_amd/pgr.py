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

import numpy as np
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
    the build), ``scale`` (``maxScale``), ``origin`` [3] and ``box`` [6] (the
    samples' min xyz, max xyz).  The native handle stays alive for ``mesh``.
    """

    def __init__(self, points: torch.Tensor, max_depth: int = 10, min_depth: int = 1):
        if points.dim() != 2 or points.shape[1] != 3:
            raise AssertionError("points must be [N, 3]")
        if points.dtype != torch.float32:
            raise TypeError("the PGR octree reads float32 samples (fscanf %f)")
        if points.device.type != "cuda":
            raise ValueError("points must be on a HIP device")
        self.L = load()
        self.device = points.device
        pts = points.contiguous()
        self._h = ct.c_void_p()
        with torch.cuda.device(self.device):
            check(self.L.cagr_pgr_build(ct.c_void_p(pts.data_ptr()), pts.shape[0],
                                        int(max_depth), int(min_depth), _stream(self.device),
                                        ct.byref(self._h)))
        nl, nc, dep = ct.c_int64(), ct.c_int64(), ct.c_int()
        sc = (ct.c_float * 4)()
        box = (ct.c_float * 6)()
        check(self.L.cagr_pgr_info(self._h, ct.byref(nl), ct.byref(nc), ct.byref(dep), sc))
        check(self.L.cagr_pgr_box(self._h, box))
        self.samples = torch.empty_like(pts)
        self.corners = torch.empty((nc.value, 3), dtype=torch.float32, device=self.device)
        self.leaves = torch.empty((nl.value, 4), dtype=torch.int32, device=self.device)
        with torch.cuda.device(self.device):
            check(self.L.cagr_pgr_read(self._h, ct.c_void_p(self.samples.data_ptr()),
                                       ct.c_void_p(self.corners.data_ptr()),
                                       ct.c_void_p(self.leaves.data_ptr()), _stream(self.device)))
        self.depth = dep.value
        self.scale = float(sc[0])
        self.origin = [float(sc[1]), float(sc[2]), float(sc[3])]
        self.box = [float(v) for v in box]

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self.L.cagr_pgr_destroy(h)
            self._h = None

    def query_points(self) -> torch.Tensor:
        """Corners in the [-1,1]^3 convention (main_GaussRecon_cuda.cu:294-298)."""
        return 2 * self.corners - 1

    def sample_points(self) -> torch.Tensor:
        """Samples in the [-1,1]^3 convention (main_GaussRecon_cuda.cu:272-274)."""
        return 2 * self.samples - 1

    def mesh(self, values: torch.Tensor, iso: float, smooth_width: float):
        """Marching cubes on this octree (``initLeaf`` + ``GetMCIsoTriangles(iso,
        mesh, 0, 1, 0, 0)``, Octree.cpp:852-879): ``values`` [G] float32 is the
        ``GridData::value`` of every corner.  Returns ``(vertices [V,3]
        float32 in [0,1]^3, triangles [T,3] int32)`` on the device, in the
        reference's order."""
        v = values.to(device=self.device, dtype=torch.float32).contiguous()
        if v.shape != (self.corners.shape[0],):
            raise AssertionError("values must be [n_corners]")
        nv, nt = ct.c_int64(), ct.c_int64()
        with torch.cuda.device(self.device):
            check(self.L.cagr_pgr_mesh(self._h, ct.c_void_p(v.data_ptr()), float(iso),
                                       float(smooth_width), _stream(self.device), ct.byref(nv),
                                       ct.byref(nt)))
            verts = torch.empty((nv.value, 3), dtype=torch.float32, device=self.device)
            tris = torch.empty((nt.value, 3), dtype=torch.int32, device=self.device)
            check(self.L.cagr_pgr_mesh_read(self._h, ct.c_void_p(verts.data_ptr()),
                                            ct.c_void_p(tris.data_ptr()), _stream(self.device)))
        return verts, tris

    def output_vertices(self, verts: torch.Tensor) -> np.ndarray:
        """Vertices in the input's coordinates (see ``output_transform``)."""
        return output_transform(verts.detach().cpu().numpy(), self.box, self.scale)


def output_transform(verts01, box, scale) -> np.ndarray:
    """Mesh vertices from [0,1]^3 back to the input's coordinates, as
    writePolygon2 writes them (Octree.cpp:1259-1267, ply.cpp:104-112):
    p * maxScale + translate with translate = (min + extent + min - maxScale)
    / 2 per axis; box = min xyz, max xyz of the samples; all float32."""
    b = np.asarray(box, np.float32)
    mn, ext = b[:3], b[3:] - b[:3]
    s = np.float32(scale)
    translate = ((((mn + ext) + mn) - s).astype(np.float64) / 2.0).astype(np.float32)
    return np.asarray(verts01, np.float32) * s + translate


def gauss_recon_values(points: torch.Tensor, normals: torch.Tensor, *, width: float = 0.01,
                       max_depth: int = 10, min_depth: int = 1, area: float = 1e-5,
                       leaf_size: int = 1, theta: float = 0.5,
                       precision: Optional[str] = None, op: str = "A3"):
    """main_GaussRecon_cuda.cu:255-486 before marching cubes (``op="A3"``);
    ``op="A_cpu"`` is main_GaussRecon_cpu.cpp:303-359 instead, whose field is
    the CPU-sign plain operator multiply_by_A with the continuous kernel.

    Returns ``(octree, values [G], isovalue, smooth_width)``: -wn at every
    corner of the PGR octree in ``gridDataVector`` order, -wn at the samples
    (element N/2 after nth_element) and width / 2.  Defaults are the
    reference's (-d 10, -m 1, -w 0.01, leaf size 1, θ = 0.5).
    """
    if op not in ("A3", "A_cpu"):
        raise ValueError("op must be 'A3' (GaussRecon CUDA) or 'A_cpu' (GaussRecon CPU)")
    oct_ = PGROctree(points, max_depth, min_depth)
    pts = oct_.sample_points()
    mu = sample_attributes(normals.to(device=pts.device, dtype=pts.dtype), area)
    wt = WindingNumberTreecode(pts, leaf_size=leaf_size, theta=theta, precision=precision)
    wn_q = wt.query(oct_.query_points(), mu, width, op=op, precision=precision)[:, 0]
    n = pts.shape[0]
    widths = torch.full((n,), width, dtype=pts.dtype, device=pts.device)
    wn_s = wt.apply(op, mu, widths, continuous=True, precision=precision)[:, 0]
    kth = torch.kthvalue(wn_s.double().cpu(), n // 2 + 1).values
    return oct_, -wn_q, float(-kth), width / 2


def gauss_recon(points: torch.Tensor, normals: torch.Tensor, out_path: Optional[str] = None, *,
                width: float = 0.01, max_depth: int = 10, min_depth: int = 1,
                area: float = 1e-5, leaf_size: int = 1, theta: float = 0.5,
                precision: Optional[str] = None, op: str = "A3"):
    """The whole of main_GaussRecon_cuda.cu (:255-498; ``op="A_cpu"``:
    main_GaussRecon_cpu.cpp): PGR octree, -wn at its corners and the
    isovalue, marching cubes, and the binary PLY of writePolygon2 when
    ``out_path`` is given.  Returns ``(vertices [V,3] in
    the input's coordinates (numpy float32), triangles [T,3] (numpy int32),
    isovalue)``."""
    from .pcio import save_ply_mesh
    oct_, vals, iso, sw = gauss_recon_values(points, normals, width=width, max_depth=max_depth,
                                            min_depth=min_depth, area=area,
                                            leaf_size=leaf_size, theta=theta,
                                            precision=precision, op=op)
    verts, tris = oct_.mesh(vals.float(), iso, sw)
    out_v = oct_.output_vertices(verts)
    out_t = tris.cpu().numpy()
    if out_path:
        save_ply_mesh(out_path, out_v, out_t)
    return out_v, out_t, iso
