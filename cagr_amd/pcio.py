"""Point-cloud input/output and the end-to-end orientation run.

Restates the host side of CUDA-based/main_CAGR.py around the solve:

* ``load_points`` -- the input formats of main_CAGR.py:27-39: ``.xyz`` (text,
  ``numpy.loadtxt``, the first three columns), ``.npy`` (the first three
  columns; loaded without pickle) and ``.ply`` / ``.obj`` (the reference goes
  through trimesh, which is not available here: this module parses the vertex
  element of ASCII and binary PLY files and the ``v`` lines of OBJ files
  itself, so the vertices come back in file order as trimesh's
  ``process=False`` gives them);
* ``normalize`` -- main_CAGR.py:42-45: centre of the bounding box, longest
  side, scale 2 / (1.1 * longest side);
* ``save_xyz`` -- main_CAGR.py:188-192: the *unnormalized* points and the
  normals side by side through ``numpy.savetxt`` with its default format;
* ``orient`` -- the whole run: load, normalize, build the treecode on the GPU,
  the reference's 40-iteration schedule plus 8 G steps
  (cagr_amd.solve.cagr_schedule), write ``<out_dir>/<name>.xyz``.
"""
from __future__ import annotations

import os
from typing import Optional, Tuple

import numpy as np

BBOX_SCALE = 1.1   # main_CAGR.py:42

_PLY_TYPES = {
    "char": "i1", "int8": "i1", "uchar": "u1", "uint8": "u1",
    "short": "i2", "int16": "i2", "ushort": "u2", "uint16": "u2",
    "int": "i4", "int32": "i4", "uint": "u4", "uint32": "u4",
    "float": "f4", "float32": "f4", "double": "f8", "float64": "f8",
}


def _read_ply(path: str) -> Tuple[np.ndarray, Optional[np.ndarray]]:
    with open(path, "rb") as f:
        if f.readline().strip() != b"ply":
            raise ValueError(f"{path}: not a PLY file")
        fmt = None
        elements = []          # [name, count, [(prop name, dtype or None for lists)]]
        while True:
            line = f.readline()
            if not line:
                raise ValueError(f"{path}: PLY header has no end_header")
            tok = line.decode("ascii", "replace").split()
            if not tok or tok[0] in ("comment", "obj_info"):
                continue
            if tok[0] == "format":
                fmt = tok[1]
            elif tok[0] == "element":
                elements.append([tok[1], int(tok[2]), []])
            elif tok[0] == "property":
                if not elements:
                    raise ValueError(f"{path}: PLY property before any element")
                if tok[1] == "list":
                    elements[-1][2].append((tok[-1], None))
                else:
                    if tok[1] not in _PLY_TYPES:
                        raise ValueError(f"{path}: unknown PLY type {tok[1]}")
                    elements[-1][2].append((tok[2], _PLY_TYPES[tok[1]]))
            elif tok[0] == "end_header":
                break
        if fmt not in ("ascii", "binary_little_endian", "binary_big_endian"):
            raise ValueError(f"{path}: unsupported PLY format {fmt}")
        names = [e[0] for e in elements]
        if "vertex" not in names:
            raise ValueError(f"{path}: PLY file has no vertex element")
        vi = names.index("vertex")
        for e in elements[:vi]:
            if e[1] and any(t is None for _, t in e[2]):
                raise ValueError(f"{path}: a list element before the vertices is not supported")
        _, count, props = elements[vi]
        if any(t is None for _, t in props):
            raise ValueError(f"{path}: list properties in the vertex element are not supported")
        pnames = [n for n, _ in props]
        for c in ("x", "y", "z"):
            if c not in pnames:
                raise ValueError(f"{path}: PLY vertices have no {c}")
        if fmt == "ascii":
            skip = sum(e[1] for e in elements[:vi])
            for _ in range(skip):
                f.readline()
            rows = [f.readline().split() for _ in range(count)]
            if len(rows) != count or any(len(r) < len(props) for r in rows):
                raise ValueError(f"{path}: truncated PLY vertex data")
            # each column takes its declared type first, as a binary file would
            col = {n: np.array([r[i] for r in rows], dtype=np.float64 if t[0] == "f" else np.int64)
                   .astype(t).astype(np.float64) for i, (n, t) in enumerate(props)}
        else:
            end = "<" if fmt == "binary_little_endian" else ">"
            for e in elements[:vi]:
                dt = np.dtype([(n, end + t) for n, t in e[2]])
                f.seek(dt.itemsize * e[1], os.SEEK_CUR)
            dt = np.dtype([(n, end + t) for n, t in props])
            buf = f.read(dt.itemsize * count)
            if len(buf) != dt.itemsize * count:
                raise ValueError(f"{path}: truncated PLY vertex data")
            rec = np.frombuffer(buf, dtype=dt, count=count)
            col = {n: rec[n].astype(np.float64) for n in pnames}
    pts = np.stack([col["x"], col["y"], col["z"]], 1)
    nrm = None
    if all(c in col for c in ("nx", "ny", "nz")):
        nrm = np.stack([col["nx"], col["ny"], col["nz"]], 1)
    return pts, nrm


def _read_obj(path: str) -> np.ndarray:
    vs = []
    with open(path, "r") as f:
        for line in f:
            if line.startswith("v ") or line.startswith("v\t"):
                tok = line.split()
                if len(tok) < 4:
                    raise ValueError(f"{path}: malformed vertex line: {line.strip()}")
                vs.append([float(tok[1]), float(tok[2]), float(tok[3])])
    return np.asarray(vs, dtype=np.float64).reshape(-1, 3)


def load_points(path: str) -> Tuple[np.ndarray, Optional[np.ndarray]]:
    """(points [N,3] float64, normals [N,3] or None) from .xyz/.ply/.obj/.npy
    (main_CAGR.py:27-39; the same extension checks and error)."""
    ext = os.path.splitext(path)[-1]
    if ext == ".xyz":
        d = np.loadtxt(path, ndmin=2)
        return d[:, :3].astype(np.float64), (d[:, 3:6].astype(np.float64)
                                              if d.shape[1] >= 6 else None)
    if ext == ".ply":
        return _read_ply(path)
    if ext == ".obj":
        return _read_obj(path), None
    if ext == ".npy":
        d = np.load(path, allow_pickle=False)
        return np.asarray(d[:, :3], dtype=np.float64), None
    raise NotImplementedError("The input file must be have extension xyz/ply/obj/npy")


def normalize(points: np.ndarray) -> Tuple[np.ndarray, np.ndarray, float]:
    """main_CAGR.py:42-45: (normalized points, bbox centre, scale) with
    normalized = (points - centre) * scale, scale = 2 / (longest side * 1.1)."""
    p = np.asarray(points)
    lo, hi = p.min(0), p.max(0)
    center = (lo + hi) / 2.0
    length = (hi - lo).max()
    scale = 2 / (length * BBOX_SCALE)
    return (p - center) * scale, center, scale


def save_xyz(path: str, points: np.ndarray, normals: np.ndarray) -> None:
    """main_CAGR.py:190-192: unnormalized points and normals, one row each,
    numpy.savetxt's default format."""
    np.savetxt(path, np.concatenate([np.asarray(points), np.asarray(normals)], -1))


def orient(path: str, width_config: str = "l0", *, wsmin: Optional[float] = None,
           wsmax: Optional[float] = None, iters: int = 40, out_dir: Optional[str] = "results",
           device: str = "cuda", leaf_size: int = 1, theta: float = 0.5, an: float = 1.0,
           precision: Optional[str] = None):
    """Orient the point cloud at ``path`` the way ``main_CAGR.py`` does and
    write ``<out_dir>/<name>.xyz`` (skipped when out_dir is None).  The points
    go to the GPU as float32 (main_CAGR.py:47); ``width_config`` is a preset
    l0..l5 or ``"custom"`` with ``wsmin`` / ``wsmax``.  ``leaf_size`` 1 and
    ``theta`` 0.5 are the reference's tree; ``precision`` defaults to the
    FP32-compensated operators.  Returns (normals [N,3] numpy float32,
    normalized points [N,3] numpy float64)."""
    import torch

    from .solve import PRESETS, cagr_schedule
    from .wn_treecode import WindingNumberTreecode

    if width_config == "custom":
        if wsmin is None or wsmax is None:
            raise ValueError("width_config 'custom' needs wsmin and wsmax")
        lo_w, hi_w = float(wsmin), float(wsmax)
    elif width_config in PRESETS:
        lo_w, hi_w = PRESETS[width_config]
    else:
        raise ValueError(f"width_config must be one of {sorted(PRESETS)} or 'custom'")
    raw, _ = load_points(path)
    pn, _, _ = normalize(raw)
    pts = torch.from_numpy(pn).contiguous().float().to(device)
    wt = WindingNumberTreecode(pts, leaf_size=leaf_size, theta=theta, precision=precision)
    normals, _ = cagr_schedule(wt, lo_w, hi_w, iters=iters, an=an)
    if str(pts.device).startswith("cuda"):
        torch.cuda.synchronize(pts.device)
    out = normals.detach().cpu().numpy()
    if out_dir is not None:
        os.makedirs(out_dir, exist_ok=True)
        name = os.path.basename(path)[:-4]
        save_xyz(os.path.join(out_dir, name + ".xyz"), raw, out)
    return out, pn


def save_ply_mesh(path: str, vertices, triangles) -> None:
    """Binary little-endian PLY of a triangle mesh, byte for byte as the
    reference's PlyWritePolygons writes it (gaussrecon_src/ply.cpp:72-143 with
    plyfile.cpp's header): float x, y, z per vertex; per face a uchar count
    and int indices."""
    v = np.ascontiguousarray(vertices, dtype="<f4").reshape(-1, 3)
    t = np.ascontiguousarray(triangles, dtype="<i4").reshape(-1, 3)
    header = ("ply\nformat binary_little_endian 1.0\n"
              f"element vertex {len(v)}\nproperty float x\nproperty float y\nproperty float z\n"
              f"element face {len(t)}\nproperty list uchar int vertex_indices\nend_header\n")
    faces = np.zeros(len(t), dtype=[("n", "u1"), ("i", "<i4", (3,))])
    faces["n"] = 3
    faces["i"] = t
    with open(path, "wb") as f:
        f.write(header.encode("ascii"))
        f.write(v.tobytes())
        f.write(faces.tobytes())


def reconstruct(path: str, out_path: Optional[str], *, max_depth: int = 10, min_depth: int = 1,
                width: float = 0.01, device: str = "cuda", precision: Optional[str] = None,
                op: str = "A3"):
    """GaussRecon on an oriented point file, as ``main_GaussReconCUDA -i path -o
    out_path -d max_depth -m min_depth -w width`` (main_GaussRecon_cuda.cu:
    225-498): the samples and normals of ``path`` (x y z nx ny nz, read as
    float32 like the reference's fscanf), the PGR octree, -wn at its corners,
    the isovalue, marching cubes and the binary PLY at ``out_path`` (skipped
    when None); ``op="A_cpu"`` runs main_GaussReconCPU's field instead.
    Returns (vertices [V,3] float32 in the input's coordinates, triangles
    [T,3] int32, isovalue)."""
    import torch

    from .pgr import gauss_recon

    raw, nrm = load_points(path)
    if nrm is None:
        raise ValueError(f"{path} holds no normals (GaussRecon reads x y z nx ny nz)")
    pts = torch.from_numpy(np.ascontiguousarray(raw, dtype=np.float32)).to(device)
    nml = torch.from_numpy(np.ascontiguousarray(nrm, dtype=np.float32)).to(device)
    return gauss_recon(pts, nml, out_path, width=width, max_depth=max_depth,
                       min_depth=min_depth, precision=precision, op=op)
