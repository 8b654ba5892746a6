"""The reference's native treecode modules, ``wn_treecode._cuda`` and
``wn_treecode._cpu``, with their own function names, argument lists, shapes
and dtypes, running on the cagr_amd C ABI.

Reference interfaces (CUDA-based/ext/wn_treecode/):
  * wn_treecode_cpu/wn_treecode_cpu_torch_interface.cpp:104-118   build_tree
  * wn_treecode_cuda/wn_treecode_cuda_torch_interface.cu:41-134   scatter_point_attrs_to_nodes
    (CPU twin wn_treecode_cpu_torch_interface.cpp:120-210)
  * wn_treecode_cuda/wn_treecode_cuda_torch_interface.cu:137-777  multiply_by_{A,A3,AT,G,A21,A22,
    A23,A2T1,A2T2,A2T3} (CPU twins multiply_by_{A,AT,G}, cpu_torch_interface.cpp:213-378)
  * module contents: cuda_torch_interface.cu:779-791, cpu_torch_interface.cpp:385-391

Both modules compute on the GPU.  ``_cpu`` takes and returns host tensors, as
the reference's CPU module does, and works on copies on the current ROCm
device; its ``multiply_by_A`` has the CPU module's sign (cpu_kernels.cpp:
91-95).  ``_cuda`` takes device tensors and raises ``RuntimeError`` for host
tensors (CHECK_INPUT_FOR_CUDA); its ``multiply_by_A2`` (the reference kernel
with the comma-operator race, cuda_kernels.cu:835-837) raises
``NotImplementedError``.

The tree arrays are the reference's 7-array layout.  cagr_amd keeps its own
octree per point set (Morton-sorted point ranges, not per-node index lists),
so a call looks up the tree handle whose layout equals the arrays it is
given: the one ``build_tree`` made, or one built from the given points at the
reference's leaf size 1 and max depth 15, or at the leaf size and depth the
arrays imply.  The match is checked element for element, once per set of
array tensors, and ``RuntimeError`` is raised when no handle matches.
``multiply_by_*`` evaluate with the node moments they are handed
(``node_attrs``, ``node_reppoints``), as the reference's kernels do.

Keyword-only extras: ``leaf_size`` on ``build_tree``; ``continuous`` (the
kernels' smoothed inner branch, which the reference's pybind layer leaves
off), ``an`` (convection strength, reference 1) and ``precision`` ("fp64" /
"fp32c"; default by the points' dtype) on ``multiply_by_*``.
"""
from __future__ import annotations

import collections
import sys
import types
from typing import Dict, Optional

import numpy as np
import torch

from ._lib import IN_DIM, OP, OUT_DIM, check
from .wn_treecode import _PREC, WindingNumberTreecode, _dt, _ptr, _stream

_MAX_TREES = 8        # handles kept (least recently used dropped first)
_trees: "collections.OrderedDict[int, _Entry]" = collections.OrderedDict()
_verified: Dict[tuple, "_Entry"] = {}


class _Entry:
    """One tree handle and its reference layout, cached per device."""

    def __init__(self, wt: WindingNumberTreecode):
        self.wt = wt
        self.info = wt.info()
        self.host_points = None
        self._layout = {}

    def layout(self, device, dtype) -> Dict[str, torch.Tensor]:
        key = (str(device), dtype)
        if key not in self._layout:
            a = self.wt.reference_layout()
            names = ("parent", "children", "is_leaf", "half_w", "num_points",
                     "node2point_index", "node2point_indexstart")
            ts = dict(zip(names, (torch.from_numpy(x) for x in a)))
            ts["is_leaf"] = ts["is_leaf"].to(torch.bool)
            ts["half_w"] = ts["half_w"].to(dtype)
            self._layout[key] = {k: v.to(device) for k, v in ts.items()}
        return self._layout[key]

    def same_points(self, points: torch.Tensor) -> bool:
        mine = self.wt.points
        if mine.shape != points.shape or mine.dtype != points.dtype:
            return False
        if points.device.type == "cpu":
            if self.host_points is None:
                self.host_points = mine.cpu()
            return torch.equal(self.host_points, points)
        return torch.equal(mine, points)


def _register(wt: WindingNumberTreecode) -> _Entry:
    e = _Entry(wt)
    _trees[id(e)] = e
    while len(_trees) > _MAX_TREES:
        _, old = _trees.popitem(last=False)
        for k in [k for k, v in _verified.items() if v is old]:
            del _verified[k]
    return e


def _device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("cagr_amd runs on the GPU only: no ROCm device is visible")
    return torch.device("cuda", torch.cuda.current_device())


def _check_input(t: torch.Tensor, name: str, cpu: bool) -> None:
    if cpu and t.device.type != "cpu":
        raise RuntimeError(f"{name} must be a CPU tensor")
    if not cpu and not t.is_cuda:
        raise RuntimeError(f"{name} must be a CUDA tensor")
    if not t.is_contiguous():
        raise RuntimeError(f"{name} must be contiguous")


def _depths_from_parents(parent: np.ndarray) -> np.ndarray:
    """Node depths from the preorder parent list (parent < child, root -1)."""
    d = np.zeros(parent.shape[0], np.int64)
    has = parent >= 0
    for _ in range(17):
        nd = np.where(has, d[np.maximum(parent, 0)] + 1, 0)
        if np.array_equal(nd, d):
            break
        d = nd
    return d


def _hypotheses(is_leaf: np.ndarray, counts: np.ndarray, depth: np.ndarray):
    """(leaf_size, max_depth) pairs whose tree may equal the given one: the
    reference's (1, 15); the largest leaf (capped at 64) with depth 15; and
    the largest leaf above the deepest level with that level as the depth
    limit (leaves there may hold more points)."""
    out = [(1, 15)]
    leaves = is_leaf.astype(bool)
    if leaves.any():
        big = int(min(max(counts[leaves].max(), 1), 64))
        out.append((big, 15))
        dmax = int(depth.max())
        above = leaves & (depth < dmax)
        if above.any():
            out.append((int(min(max(counts[above].max(), 1), 64)), dmax))
        else:
            out.append((1, dmax))
    seen, res = set(), []
    for h in out:
        if h not in seen:
            seen.add(h)
            res.append(h)
    return res


def _tree_for(points: torch.Tensor, arrays: Dict[str, torch.Tensor], depth_of) -> _Entry:
    """The handle whose reference layout equals `arrays` (any subset of the
    7 names) for these points; depth_of() gives the node depths (numpy) for
    the leaf-size / depth hypotheses."""
    key = (points.data_ptr(), points._version, tuple(points.shape), points.dtype,
           str(points.device)) + tuple((k, arrays[k].data_ptr(), arrays[k]._version,
                                        tuple(arrays[k].shape)) for k in sorted(arrays))
    e = _verified.get(key)
    if e is not None:
        _trees.move_to_end(id(e))
        return e

    def matches(e: _Entry) -> bool:
        if not e.same_points(points):
            return False
        mine = e.layout(points.device, points.dtype)
        for k, v in arrays.items():
            ref = mine[k]
            if ref.shape != v.shape or not torch.equal(ref, v.to(ref.dtype)):
                return False
        return True

    m = arrays["children"].shape[0] if "children" in arrays else arrays["parent"].shape[0]
    for e in reversed(list(_trees.values())):
        if e.info["num_nodes"] == m and matches(e):
            _verified[key] = e
            _trees.move_to_end(id(e))
            return e
    dev = _device()
    is_leaf = arrays["is_leaf"].cpu().numpy()
    counts = arrays["num_points"].cpu().numpy()
    for leaf, md in _hypotheses(is_leaf, counts, depth_of()):
        with torch.cuda.device(dev):
            wt = WindingNumberTreecode(points.to(dev), md, leaf_size=leaf)
        e = _register(wt)
        if e.info["num_nodes"] != m:
            del _trees[id(e)]
            continue
        if matches(e):
            _verified[key] = e
            return e
        del _trees[id(e)]
    raise RuntimeError("the tree arrays do not describe an octree cagr_amd builds for these "
                       "points (tried the reference's leaf size 1 / max depth 15 and the leaf "
                       "size and depth the arrays imply): pass the arrays of build_tree")


# ---------------------------------------------------------------- the functions
def build_tree(points_cpu: torch.Tensor, max_depth: int, *, leaf_size: int = 1):
    """cpu_torch_interface.cpp:41-118: the 7 tree tensors on the host, in the
    reference's order and dtypes (int64, int64 [M,8], bool, points' dtype,
    int64, int64, int64).  The tree is built on the GPU."""
    _check_input(points_cpu, "points_tensor", cpu=True)
    if points_cpu.dtype not in (torch.float32, torch.float64):
        raise RuntimeError(f"build_tree not implemented for dtype '{points_cpu.dtype}'")
    dev = _device()
    with torch.cuda.device(dev):
        wt = WindingNumberTreecode(points_cpu.to(dev), int(max_depth), leaf_size=leaf_size)
    e = _register(wt)
    e.host_points = points_cpu.clone()
    lay = e.layout("cpu", points_cpu.dtype)
    return [lay[k].clone() for k in ("parent", "children", "is_leaf", "half_w", "num_points",
                                     "node2point_index", "node2point_indexstart")]


def _scatter(cpu: bool):
    def scatter_point_attrs_to_nodes(node_parent_list, node_children_list, points, point_weights,
                                     point_attrs, node2point_index, node2point_indexstart,
                                     num_points_in_node, node_is_leaf_list, tree_depth):
        """scatter_point_attrs_to_nodes (cuda_torch_interface.cu:41-134): per
        node, the attribute sum, the weighted centre and the weight sum, in
        the points' dtype (computed in FP64)."""
        args = dict(node_parent_list=node_parent_list, node_children_list=node_children_list,
                    points=points, point_weights=point_weights, point_attrs=point_attrs,
                    node2point_index=node2point_index, node2point_indexstart=node2point_indexstart,
                    num_points_in_node=num_points_in_node, node_is_leaf_list=node_is_leaf_list)
        for k, v in args.items():
            _check_input(v, k, cpu)
        attr_dim = point_attrs.shape[1]
        assert attr_dim in (1, 3)
        arrays = {"parent": node_parent_list, "children": node_children_list,
                  "is_leaf": node_is_leaf_list, "num_points": num_points_in_node,
                  "node2point_index": node2point_index,
                  "node2point_indexstart": node2point_indexstart}
        e = _tree_for(points, arrays,
                      lambda: _depths_from_parents(node_parent_list.cpu().numpy()))
        if int(tree_depth) < e.info["depth"]:
            raise ValueError(f"tree_depth {int(tree_depth)} is below the tree's depth "
                             f"{e.info['depth']}: the upper levels would stay unscattered")
        dev = e.wt.device
        na, nr, nw = e.wt.upward(point_attrs.to(dev).contiguous(),
                                 point_weights.to(dev).contiguous())
        out = [x.to(points.dtype) for x in (na, nr, nw)]
        return [x.cpu() for x in out] if cpu else out
    return scatter_point_attrs_to_nodes


def _multiply(op: str, cpu: bool):
    def multiply_by(query_points, query_width, points, point_attrs, node2point_index,
                    node2point_indexstart, node_children_list, node_attrs, node_is_leaf_list,
                    node_half_w_list, node_reppoints, num_points_in_node, *, continuous=False,
                    an: float = 1.0, precision: Optional[str] = None):
        args = dict(query_points=query_points, query_width=query_width, points=points,
                    point_attrs=point_attrs, node2point_index=node2point_index,
                    node2point_indexstart=node2point_indexstart,
                    node_children_list=node_children_list, node_attrs=node_attrs,
                    node_is_leaf_list=node_is_leaf_list, node_half_w_list=node_half_w_list,
                    node_reppoints=node_reppoints, num_points_in_node=num_points_in_node)
        for k, v in args.items():
            _check_input(v, k, cpu)
        dt = points.dtype
        for k in ("query_points", "query_width", "point_attrs", "node_attrs", "node_half_w_list",
                  "node_reppoints"):
            if args[k].dtype != dt:
                raise RuntimeError(f"{k} must have the points' dtype {dt}")
        n, m, nq = points.shape[0], node_children_list.shape[0], query_points.shape[0]
        assert point_attrs.shape == (n, IN_DIM[op]), (op, tuple(point_attrs.shape))
        assert node_attrs.shape == (m, IN_DIM[op]) and node_reppoints.shape == (m, 3)
        assert query_points.shape == (nq, 3) and query_width.shape == (nq,)
        arrays = {"children": node_children_list, "is_leaf": node_is_leaf_list,
                  "half_w": node_half_w_list, "num_points": num_points_in_node,
                  "node2point_index": node2point_index,
                  "node2point_indexstart": node2point_indexstart}

        def depth_of():
            hw = node_half_w_list.double().cpu().numpy()
            return np.rint(-np.log2(np.where(hw > 0, hw, 1.0))).astype(np.int64)

        e = _tree_for(points, arrays, depth_of)
        wt = e.wt
        dev = wt.device
        own = query_points.data_ptr() == points.data_ptr() and query_points.shape == points.shape
        q = None if own else query_points.to(dev).contiguous()
        pa = point_attrs.to(dev).contiguous()
        nat = node_attrs.to(dev).contiguous()
        nrp = node_reppoints.to(dev).contiguous()
        qw = query_width.to(dev).contiguous()
        out = torch.empty((nq, OUT_DIM[op]), dtype=dt, device=dev)
        prec = _PREC[precision or ("fp64" if dt == torch.float64 else "fp32c")]
        with torch.cuda.device(dev):
            check(wt.L.cagr_apply_moments(wt._h, OP[op], prec, _ptr(pa), _dt(pa), _ptr(nat),
                                          _ptr(nrp), _ptr(q), _dt(points), nq, _ptr(qw),
                                          int(continuous), float(an), _ptr(out), _stream(dev)))
        return out.cpu() if cpu else out
    multiply_by.__name__ = f"multiply_by_{op if op != 'A_cpu' else 'A'}"
    multiply_by.__doc__ = (f"multiply_by_{op} of the reference's native module "
                           "(cuda_torch_interface.cu:137-777): the treecode operator at "
                           "query_points with the given node moments.")
    return multiply_by


def _no_A2(*args, **kwargs):
    raise NotImplementedError(
        "multiply_by_A2 is not provided: the reference kernel races on its output "
        "(cuda_kernels.cu:835-837); use multiply_by_A21/A22/A23")


def _module(name: str, cpu: bool) -> types.ModuleType:
    mod = types.ModuleType(name, f"cagr_amd stand-in for the reference's wn_treecode.{name[-4:]}")
    mod.scatter_point_attrs_to_nodes = _scatter(cpu)
    if cpu:
        mod.build_tree = build_tree
        mod.multiply_by_A = _multiply("A_cpu", True)
        for op in ("AT", "G"):
            setattr(mod, f"multiply_by_{op}", _multiply(op, True))
    else:
        for op in ("A", "A3", "AT", "G", "A21", "A22", "A23", "A2T1", "A2T2", "A2T3"):
            setattr(mod, f"multiply_by_{op}", _multiply(op, False))
        mod.multiply_by_A2 = _no_A2
    sys.modules[name] = mod
    return mod


_cuda = _module("cagr_amd._cuda", cpu=False)
_cpu = _module("cagr_amd._cpu", cpu=True)
