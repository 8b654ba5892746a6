"""CAGR treecode oracle -- TEST INFRASTRUCTURE ONLY.

Python (numpy + ctypes) front end to ``oracle/libcagr_oracle.so`` (the FP64
C++ restatement in ``oracle/cagr_oracle.cpp``) and, where it was built, to
``oracle/_ref/libcagr_ref.so`` (the reference's own torch-free CPU core driven
by ``oracle/ref_harness.cpp``).  Only ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s cpu_baseline leg import this module, and only as the
checker; the ``cagr_amd`` product path never does.

The tree arrays use the reference's serialised layout
(CUDA-based/ext/wn_treecode/wn_treecode_cpu/wn_treecode_cpu_treeutils.cpp:202-258):
parent [M] i64, children [M,8] i64 (-1 = none), is_leaf [M] u8, half_w [M] f64,
num_points [M] i64, node2point_index [sum] i64, node2point_indexstart [M] i64.
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libcagr_oracle.so")
_REF_PATH = os.path.join(_HERE, "_ref", "libcagr_ref.so")
_PGR_REF_PATH = os.path.join(_HERE, "_ref", "libpgr_ref.so")

# op codes shared with cagr_oracle.cpp
OPS = {
    "A": 0, "A_cpu": 1, "AT": 2, "G": 3,
    "A21": 4, "A22": 5, "A23": 6,
    "A2T1": 7, "A2T2": 8, "A2T3": 9,
    "A3": 10,
}
IN_DIM = {k: (1 if k in ("AT", "A2T1", "A2T2", "A2T3") else 3) for k in OPS}
OUT_DIM = {k: (3 if k in ("AT", "G", "A2T1", "A2T2", "A2T3") else 1) for k in OPS}

_P = ct.c_void_p
_I = ct.c_int64
_D = ct.c_double


def build(force: bool = False) -> None:
    """Compile the oracle (and oracle/_ref when the reference is present)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "all"], check=True)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ct.CDLL(_LIB_PATH)
        L.orc_tree_build.restype = _P
        L.orc_tree_build.argtypes = [_P, _I, _I, _I]
        L.orc_tree_info.argtypes = [_P, _P, _P, _P, _P]
        L.orc_tree_export.argtypes = [_P] * 8
        L.orc_tree_free.argtypes = [_P]
        L.orc_upward.argtypes = [_I, _P, _P, _P, _P, _P, _P, _P, _P, _I, _P, _P, _P]
        L.orc_apply.argtypes = [_I, _I, _P, _P, _P, _P, _I, _P, _P, _P, _P, _P, _P, _P, _P,
                                _D, _I, _D, _P, _P, _P, _P]
        L.orc_direct.argtypes = [_I, _I, _P, _P, _I, _P, _P, _I, _D, _P]
        L.orc_adjoint.argtypes = [_I, _I, _P, _P, _I, _P, _I, _P, _P, _P, _P, _P, _P, _P,
                                  _D, _I, _D, _P, _P]
        L.orc_num_threads.restype = _I
        L.orc_set_num_threads.argtypes = [_I]
        L.pgr_oracle_build.argtypes = [_P, _I, ct.c_int, ct.c_int, _P, _P, _P, _P]
        L.pgr_oracle_fetch.argtypes = [_P, _P]
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(_REF_PATH)


def ref_lib():
    global _ref
    if _ref is None:
        R = ct.CDLL(_REF_PATH)
        R.ref_tree_build_f64.restype = _P
        R.ref_tree_build_f64.argtypes = [_P, _I, _I, _I]
        R.ref_tree_build_f32.restype = _P
        R.ref_tree_build_f32.argtypes = [_P, _I, _I, _I]
        R.ref_tree_info.argtypes = [_P, _I, _P, _P, _P, _P]
        R.ref_tree_export.argtypes = [_P, _I] + [_P] * 7
        R.ref_tree_free.argtypes = [_P, _I]
        R.ref_upward_f64.argtypes = [_I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I, _I, _P, _P, _P]
        R.ref_apply_f64.argtypes = [_I, _I, _P, _P, _P, _P, _I, _P, _P, _P, _P, _P, _P, _P,
                                    _P, _I, _P]
        _ref = R
    return _ref


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_P)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


@dataclass
class Tree:
    """A tree in the reference's serialised layout (numpy, host)."""
    parent: np.ndarray
    children: np.ndarray
    is_leaf: np.ndarray
    half_w: np.ndarray
    num_points: np.ndarray
    node2point_index: np.ndarray
    node2point_indexstart: np.ndarray
    num_leaves: int = 0
    depth: int = 0

    @property
    def num_nodes(self) -> int:
        return int(self.parent.shape[0])

    def arrays(self):
        return (self.parent, self.children, self.is_leaf, self.half_w, self.num_points,
                self.node2point_index, self.node2point_indexstart)


def _alloc_tree(m: int, ll: int) -> Tree:
    return Tree(parent=np.empty(m, np.int64), children=np.empty((m, 8), np.int64),
                is_leaf=np.empty(m, np.uint8), half_w=np.empty(m, np.float64),
                num_points=np.empty(m, np.int64), node2point_index=np.empty(ll, np.int64),
                node2point_indexstart=np.empty(m, np.int64))


def build_tree(points, max_depth: int = 15, leaf_size: int = 1) -> Tree:
    """Oracle octree over points [N,3] (compared in double)."""
    p = _c(points, np.float64)
    L = lib()
    h = L.orc_tree_build(_ptr(p), p.shape[0], max_depth, leaf_size)
    m, nl, dp, ll = (np.zeros(1, np.int64) for _ in range(4))
    L.orc_tree_info(h, _ptr(m), _ptr(nl), _ptr(dp), _ptr(ll))
    t = _alloc_tree(int(m[0]), int(ll[0]))
    L.orc_tree_export(h, *(_ptr(a) for a in t.arrays()))
    L.orc_tree_free(h)
    t.num_leaves, t.depth = int(nl[0]), int(dp[0])
    return t


def ref_build_tree(points, max_depth: int = 15, leaf_size: int = 1, f32: bool = False) -> Tree:
    """The reference's own recursive builder (oracle/_ref)."""
    R = ref_lib()
    p = _c(points, np.float32 if f32 else np.float64)
    fn = R.ref_tree_build_f32 if f32 else R.ref_tree_build_f64
    h = fn(_ptr(p), p.shape[0], max_depth, leaf_size)
    m, nl, dp, ll = (np.zeros(1, np.int64) for _ in range(4))
    R.ref_tree_info(h, int(f32), _ptr(m), _ptr(nl), _ptr(dp), _ptr(ll))
    t = _alloc_tree(int(m[0]), int(ll[0]))
    R.ref_tree_export(h, int(f32), *(_ptr(a) for a in t.arrays()))
    R.ref_tree_free(h, int(f32))
    t.num_leaves, t.depth = int(nl[0]), int(dp[0])
    return t


def point_weights(attrs) -> np.ndarray:
    """||attr||_2 per point, as wn_treecode_func.py:131 computes it."""
    a = _c(attrs, np.float64)
    if a.ndim == 1:
        a = a[:, None]
    s = np.zeros(a.shape[0])
    for c in range(a.shape[1]):
        s = s + a[:, c] * a[:, c]
    return np.sqrt(s)


def upward(tree: Tree, points, attrs, weights=None):
    """Node moments: (node_attrs [M,C], node_reppoints [M,3], node_weights [M])."""
    p = _c(points, np.float64)
    a = _c(attrs, np.float64)
    if a.ndim == 1:
        a = a[:, None]
    w = point_weights(a) if weights is None else _c(weights, np.float64)
    m = tree.num_nodes
    na = np.zeros((m, a.shape[1]))
    nr = np.zeros((m, 3))
    nw = np.zeros(m)
    lib().orc_upward(m, _ptr(tree.children), _ptr(tree.is_leaf), _ptr(tree.num_points),
                     _ptr(tree.node2point_index), _ptr(tree.node2point_indexstart), _ptr(p),
                     _ptr(w), _ptr(a), a.shape[1], _ptr(na), _ptr(nr), _ptr(nw))
    return na, nr, nw


def ref_upward(tree: Tree, points, attrs, weights, tree_depth: int = 15):
    p = _c(points, np.float64)
    a = _c(attrs, np.float64)
    if a.ndim == 1:
        a = a[:, None]
    w = _c(weights, np.float64)
    m = tree.num_nodes
    na = np.zeros((m, a.shape[1]))
    nr = np.zeros((m, 3))
    nw = np.zeros(m)
    ref_lib().ref_upward_f64(m, _ptr(tree.parent), _ptr(tree.children), _ptr(tree.is_leaf),
                             _ptr(tree.num_points), _ptr(tree.node2point_index),
                             _ptr(tree.node2point_indexstart), _ptr(p), _ptr(w), _ptr(a),
                             a.shape[1], tree_depth, _ptr(na), _ptr(nr), _ptr(nw))
    return na, nr, nw


def apply(op: str, tree: Tree, points, attrs, widths, queries=None, node_moments=None,
          theta: float = 0.5, continuous: bool = False, an: float = 1.0,
          weights=None, counts: bool = False):
    """Treecode op of the reference (per-query DFS, FP64).

    Returns out [Nq, out_dim] (and, with counts=True, (n_far, n_near, n_visit)).
    Node moments default to the upward pass with weights ||attr|| -- exactly
    what every forward_* of wn_treecode_func.py does before its traversal.
    """
    p = _c(points, np.float64)
    a = _c(attrs, np.float64)
    if a.ndim == 1:
        a = a[:, None]
    assert a.shape[1] == IN_DIM[op]
    q = p if queries is None else _c(queries, np.float64)
    qw = _c(np.broadcast_to(np.asarray(widths, np.float64), (q.shape[0],)), np.float64)
    if node_moments is None:
        node_moments = upward(tree, p, a, weights)
    na, nr = _c(node_moments[0], np.float64), _c(node_moments[1], np.float64)
    nq = q.shape[0]
    out = np.zeros((nq, OUT_DIM[op]))
    cf, cn, cv = (np.zeros(nq, np.int64) for _ in range(3))
    lib().orc_apply(OPS[op], nq, _ptr(q), _ptr(qw), _ptr(p), _ptr(a), tree.num_nodes,
                    _ptr(tree.children), _ptr(tree.is_leaf), _ptr(tree.half_w),
                    _ptr(tree.num_points), _ptr(tree.node2point_index),
                    _ptr(tree.node2point_indexstart), _ptr(na), _ptr(nr), 2.0 / theta,
                    int(continuous), an, _ptr(out), _ptr(cf), _ptr(cn), _ptr(cv))
    if counts:
        return out, (cf, cn, cv)
    return out


def ref_apply(op: str, tree: Tree, points, attrs, widths, node_moments, queries=None,
              continuous: bool = False):
    """The reference's own CPU A (CPU sign) / AT / G launchers (oracle/_ref)."""
    code = {"A_cpu": 0, "AT": 1, "G": 2}[op]
    p = _c(points, np.float64)
    a = _c(attrs, np.float64)
    if a.ndim == 1:
        a = a[:, None]
    q = p if queries is None else _c(queries, np.float64)
    qw = _c(np.broadcast_to(np.asarray(widths, np.float64), (q.shape[0],)), np.float64)
    na, nr = _c(node_moments[0], np.float64), _c(node_moments[1], np.float64)
    out = np.zeros((q.shape[0], OUT_DIM[op]))
    ref_lib().ref_apply_f64(code, q.shape[0], _ptr(q), _ptr(qw), _ptr(p), _ptr(a),
                            tree.num_nodes, _ptr(tree.children), _ptr(tree.is_leaf),
                            _ptr(tree.half_w), _ptr(tree.num_points),
                            _ptr(tree.node2point_index), _ptr(tree.node2point_indexstart),
                            _ptr(na), _ptr(nr), int(continuous), _ptr(out))
    return out


def direct(op: str, points, attrs, widths, queries=None, continuous: bool = False,
           an: float = 1.0):
    """Brute-force direct sum over all sources (theta -> 0)."""
    p = _c(points, np.float64)
    a = _c(attrs, np.float64)
    if a.ndim == 1:
        a = a[:, None]
    q = p if queries is None else _c(queries, np.float64)
    qw = _c(np.broadcast_to(np.asarray(widths, np.float64), (q.shape[0],)), np.float64)
    out = np.zeros((q.shape[0], OUT_DIM[op]))
    lib().orc_direct(OPS[op], q.shape[0], _ptr(q), _ptr(qw), p.shape[0], _ptr(p), _ptr(a),
                     int(continuous), an, _ptr(out))
    return out


def adjoint(op: str, tree: Tree, points, g, widths, node_reppoints, queries=None,
            theta: float = 0.5, continuous: bool = False, an: float = 1.0):
    """Exact adjoint of the linear treecode op with fixed centres."""
    p = _c(points, np.float64)
    q = p if queries is None else _c(queries, np.float64)
    qw = _c(np.broadcast_to(np.asarray(widths, np.float64), (q.shape[0],)), np.float64)
    gg = _c(g, np.float64).reshape(q.shape[0], OUT_DIM[op])
    nr = _c(node_reppoints, np.float64)
    res = np.zeros((p.shape[0], IN_DIM[op]))
    lib().orc_adjoint(OPS[op], q.shape[0], _ptr(q), _ptr(qw), p.shape[0], _ptr(p),
                      tree.num_nodes, _ptr(tree.children), _ptr(tree.is_leaf),
                      _ptr(tree.half_w), _ptr(tree.num_points), _ptr(tree.node2point_index),
                      _ptr(tree.node2point_indexstart), _ptr(nr), 2.0 / theta,
                      int(continuous), an, _ptr(gg), _ptr(res))
    return res


# ---------------------------------------------------------------- PGR octree
@dataclass
class PGR:
    """GaussRecon's PGR octree query set: samples in [0,1]^3 [N,3] f32, leaves
    [L,4] i32 (depth, x, y, z) in preorder, corners [G,3] f32 in gridDataVector
    order, depth = Octree::maxDepth after the build (tree height + 1)."""
    samples: np.ndarray
    leaves: np.ndarray
    corners: np.ndarray
    depth: int


def pgr_build(points, max_depth: int = 10, min_depth: int = 1) -> PGR:
    """The oracle's restatement (oracle/pgr_oracle.cpp) on float32 samples."""
    L = lib()
    p = np.ascontiguousarray(points, dtype=np.float32)
    s = np.zeros_like(p)
    nl, nc, md = ct.c_int64(), ct.c_int64(), ct.c_int()
    rc = L.pgr_oracle_build(_ptr(p), p.shape[0], int(max_depth), int(min_depth), _ptr(s),
                            ct.byref(nl), ct.byref(nc), ct.byref(md))
    if rc:
        raise ValueError("pgr_oracle_build rejected its arguments")
    lv = np.zeros((nl.value, 4), np.int32)
    cs = np.zeros((nc.value, 3), np.float32)
    L.pgr_oracle_fetch(_ptr(lv), _ptr(cs))
    return PGR(s, lv, cs, md.value)


def pgr_ref_available() -> bool:
    return os.path.exists(_PGR_REF_PATH)


def pgr_ref_build(points, max_depth: int = 10, min_depth: int = 1, workdir: str = None) -> PGR:
    """The reference's own Octree::setTree (oracle/_ref/libpgr_ref.so), fed a
    .xyz file whose decimal text round-trips the float32 samples exactly."""
    import tempfile
    R = ct.CDLL(_PGR_REF_PATH)
    R.pgr_ref_build.argtypes = [ct.c_char_p, ct.c_int, ct.c_int, _P, _P, _P, _P]
    R.pgr_ref_fetch.argtypes = [_P, _P, _P]
    R.pgr_ref_leaves.argtypes = [_P]
    p = np.ascontiguousarray(points, dtype=np.float32)
    with tempfile.TemporaryDirectory(dir=workdir) as d:
        path = os.path.join(d, "samples.xyz")
        np.savetxt(path, np.hstack([p, np.zeros_like(p)]), fmt="%.9g")
        ns, ng, nl = ct.c_int64(), ct.c_int64(), ct.c_int64()
        md = ct.c_int()
        if R.pgr_ref_build(path.encode(), int(max_depth), int(min_depth), ct.byref(ns),
                           ct.byref(ng), ct.byref(md), ct.byref(nl)):
            raise RuntimeError("reference Octree::setTree failed")
    s = np.zeros((ns.value, 3), np.float32)
    g = np.zeros((ng.value, 3), np.float32)
    sc = np.zeros(4, np.float32)
    R.pgr_ref_fetch(_ptr(s), _ptr(g), _ptr(sc))
    lv = np.zeros((nl.value, 4), np.int32)
    R.pgr_ref_leaves(_ptr(lv))
    return PGR(s, lv, g, md.value)


def pgr_ref_mesh(values, iso: float, smooth_width: float, ply_path: str = ""):
    """The reference's own marching cubes on the octree of the last
    pgr_ref_build call (oracle/_ref/libpgr_ref.so): vertices [V,3] f32 in
    [0,1]^3, polygons as (offsets [P+1], indices); writes its PLY when a path
    is given."""
    R = ct.CDLL(_PGR_REF_PATH)
    R.pgr_ref_mesh.argtypes = [_P, ct.c_float, ct.c_float, ct.c_char_p, _P, _P, _P]
    R.pgr_ref_mesh_fetch.argtypes = [_P, _P, _P]
    v = np.ascontiguousarray(values, dtype=np.float32)
    nv, npo, ni = ct.c_int64(), ct.c_int64(), ct.c_int64()
    if R.pgr_ref_mesh(_ptr(v), float(iso), float(smooth_width), ply_path.encode(), ct.byref(nv),
                      ct.byref(npo), ct.byref(ni)):
        raise RuntimeError("reference mesh failed (no octree built)")
    pts = np.zeros((nv.value, 3), np.float32)
    off = np.zeros(npo.value + 1, np.int64)
    idx = np.zeros(ni.value, np.int32)
    R.pgr_ref_mesh_fetch(_ptr(pts), _ptr(off), _ptr(idx))
    return pts, off, idx


def sphere_field(corners, centre=0.5, radius=0.3):
    """A reproducible float32 test field: |c - centre|^2 - r^2, every step a
    float32 operation in a fixed order."""
    c = np.asarray(corners, np.float32)
    d = c - np.float32(centre)
    return ((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]) - np.float32(radius * radius)


def winding_field(samples01, normals, corners01, width=0.01, area=1e-5):
    """GaussRecon's field from the FP64 oracle: -wn (A3, continuous) at the
    corners and the isovalue (-wn at the samples, element N/2), as float32."""
    p64 = (2 * np.asarray(samples01, np.float32) - 1).astype(np.float64)
    nml = np.asarray(normals, np.float64)
    mu = nml / np.maximum(np.linalg.norm(nml, axis=1, keepdims=True), 1e-12) * area
    t = build_tree(p64, 15, 1)
    q = (2 * np.asarray(corners01, np.float32) - 1).astype(np.float64)
    vals = -apply("A3", t, p64, mu, width, queries=q, continuous=True)[:, 0]
    ws = apply("A3", t, p64, mu, width, continuous=True)[:, 0]
    iso = -np.partition(ws, len(ws) // 2)[len(ws) // 2]
    return vals.astype(np.float32), np.float32(iso)
