"""``WindingNumberTreecode`` -- drop-in for the reference's plugin class.

Mirrors CUDA-based/ext/wn_treecode/wn_treecode_func.py:3-392: the same class
name, constructor, ``forward_*`` methods, argument shapes, dtypes, output
shapes and ``AssertionError`` on bad shapes, so ``main_CAGR.py``'s solve loop
runs unchanged on ROCm tensors.  Every call goes to the cagr_amd HIP library
through its C ABI (cagr_amd/_lib.py); the tree is built on the GPU.

Differences from the reference, all opt-in or documented:
  * keyword-only ``theta`` (opening parameter, reference 0.5), ``leaf_size``
    (reference 1) and ``precision`` ("fp64" / "fp32c"; default: "fp64" for
    float64 points, "fp32c" for float32 points);
  * fused extras ``forward_A2`` (A21, A22, A23 in one pass) and
    ``forward_A2T_sum`` (A2T1(s0) + A2T2(s1) + A2T3(s2) in one pass), plus
    ``forward_A3``, ``query`` and ``grid_query`` for the winding-number query;
  * CPU tensors are rejected: this implementation has no CPU path (the
    reference's ``_cpu`` module cannot run CAGR either: it lacks A2k/A2Tk).
  * ``forward_A`` returns the CUDA sign convention of the reference's GPU
    module (+(x-y).mu/r^3); ``forward_A_cpu_sign`` gives the CPU module's.
"""
from __future__ import annotations

import ctypes as ct
from typing import Optional

import numpy as np
import torch

from . import _lib
from ._lib import F32, F64, OP, IN_DIM, OUT_DIM, check

_PREC = {"fp64": _lib.PREC_FP64, "fp32c": _lib.PREC_FP32C}


def _dt(t: torch.Tensor) -> int:
    if t.dtype == torch.float32:
        return F32
    if t.dtype == torch.float64:
        return F64
    raise TypeError(f"cagr_amd supports float32 and float64 tensors, got {t.dtype}")


def _ptr(t: Optional[torch.Tensor]):
    return None if t is None else ct.c_void_p(t.data_ptr())


def _stream(device) -> ct.c_void_p:
    return ct.c_void_p(torch.cuda.current_stream(device).cuda_stream)


class WindingNumberTreecode:
    def __init__(self, points: torch.Tensor, max_tree_depth: int = 15, *, theta: float = 0.5,
                 leaf_size: int = 1, precision: Optional[str] = None):
        """points: [N, 3] float32/float64 on a ROCm device."""
        assert len(points.shape) == 2
        assert points.shape[1] == 3
        if not points.is_cuda:
            raise RuntimeError("cagr_amd runs on the GPU only: move the points to a ROCm device")
        self.L = _lib.load()
        self.is_cuda = True
        self.device = points.device
        self.points = points.contiguous()
        self.tree_depth = max_tree_depth
        self.theta = float(theta)
        self.leaf_size = int(leaf_size)
        if precision is None:
            precision = "fp64" if points.dtype == torch.float64 else "fp32c"
        if precision not in _PREC:
            raise ValueError(f"precision must be one of {sorted(_PREC)}")
        self.precision = precision
        self._h = ct.c_void_p()
        with torch.cuda.device(self.device):
            check(self.L.cagr_tree_build(_ptr(self.points), self.points.shape[0],
                                         _dt(self.points), max_tree_depth, self.leaf_size,
                                         self.theta, _stream(self.device), ct.byref(self._h)))
        self._ref_layout = None
        self.walk_pieces = (0, 0)   # (pieces, waves cut) of the self-query walks (balance)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self.L.cagr_tree_destroy(h)
            self._h = None

    # ------------------------------------------------------------ tree info
    def info(self):
        vals = [ct.c_int64() for _ in range(4)]
        check(self.L.cagr_tree_info(self._h, *(ct.byref(v) for v in vals)))
        return dict(num_nodes=vals[0].value, num_leaves=vals[1].value, depth=vals[2].value,
                    list_len=vals[3].value)

    def reference_layout(self):
        """The reference's 7 serialised tree tensors (wn_treecode_cpu_treeutils.cpp:202-258),
        in the reference's order: parent, children, is_leaf, half_w, num_points,
        node2point_index, node2point_indexstart (host numpy arrays)."""
        if self._ref_layout is None:
            inf = self.info()
            m, ll = inf["num_nodes"], inf["list_len"]
            arrs = [np.empty(m, np.int64), np.empty((m, 8), np.int64), np.empty(m, np.uint8),
                    np.empty(m, np.float64), np.empty(m, np.int64), np.empty(ll, np.int64),
                    np.empty(m, np.int64)]
            check(self.L.cagr_tree_export(self._h, *(a.ctypes.data_as(ct.c_void_p) for a in arrs)))
            self._ref_layout = arrs
        return self._ref_layout

    def _layout_tensor(self, i, dtype=None):
        a = torch.from_numpy(self.reference_layout()[i])
        if dtype is not None:
            a = a.to(dtype)
        return a.to(self.device)

    # the reference's tree attributes (wn_treecode_func.py:34-43), on demand
    node_parent_list = property(lambda self: self._layout_tensor(0))
    node_children_list = property(lambda self: self._layout_tensor(1))
    node_is_leaf_list = property(lambda self: self._layout_tensor(2, torch.bool))
    node_half_w_list = property(lambda self: self._layout_tensor(3, self.points.dtype))
    num_points_in_node = property(lambda self: self._layout_tensor(4))
    node2point_index = property(lambda self: self._layout_tensor(5))
    node2point_indexstart = property(lambda self: self._layout_tensor(6))

    def profile(self, enable: bool) -> None:
        """Start (clear) or stop HIP-event timing of this tree's launches."""
        check(self.L.cagr_profile(self._h, int(enable)))

    def profile_read(self):
        """{'traverse': [ms per launch], 'upward': [ms per upward pass]} in launch
        order since profile(True)."""
        res = {}
        for kind, name in ((0, "traverse"), (1, "upward")):
            n, ms = ct.c_int64(), ct.c_double()
            check(self.L.cagr_profile_read(self._h, kind, ct.byref(n), ct.byref(ms), None, 0))
            each = np.zeros(max(n.value, 1))
            check(self.L.cagr_profile_read(self._h, kind, ct.byref(n), ct.byref(ms),
                                           each.ctypes.data_as(ct.c_void_p), n.value))
            res[name] = each[: n.value].tolist()
        return res

    def permutation(self) -> np.ndarray:
        """Sorted position -> original index (the leaf-preorder point order)."""
        perm = np.empty(self.points.shape[0], np.int32)
        check(self.L.cagr_tree_permutation(self._h, perm.ctypes.data_as(ct.c_void_p)))
        return perm

    # ------------------------------------------------------------ core call
    def _prec(self, precision):
        return _PREC[precision or self.precision]

    def apply(self, op: str, attrs: torch.Tensor, widths: torch.Tensor, *, continuous=False,
              an: float = 1.0, precision: Optional[str] = None,
              weights: Optional[torch.Tensor] = None) -> torch.Tensor:
        """Any operator with the tree's points as queries; attrs [N,in], widths [N]."""
        n = self.points.shape[0]
        attrs = attrs.contiguous()
        widths = widths.contiguous()
        if widths.dtype != attrs.dtype:
            widths = widths.to(attrs.dtype)
        if attrs.device != self.device or widths.device != self.device:
            raise RuntimeError("attrs and widths must be on the tree's device")
        assert attrs.shape == (n, IN_DIM[op]), (op, tuple(attrs.shape))
        out = torch.empty((n, OUT_DIM[op]), dtype=attrs.dtype, device=self.device)
        if weights is not None:
            weights = weights.contiguous().to(attrs.dtype)
            assert weights.shape == (n,)
        with torch.cuda.device(self.device):
            if weights is None:
                rc = self.L.cagr_apply(self._h, OP[op], self._prec(precision), _ptr(attrs),
                                       _dt(attrs), _ptr(widths), int(continuous), float(an),
                                       _ptr(out), _stream(self.device))
            else:
                rc = self.L.cagr_apply_weighted(self._h, OP[op], self._prec(precision),
                                                _ptr(attrs), _dt(attrs), _ptr(weights),
                                                _ptr(widths), int(continuous), float(an),
                                                _ptr(out), _stream(self.device))
        check(rc)
        return out

    def apply_adjoint(self, op: str, g: torch.Tensor, widths: torch.Tensor, weights: torch.Tensor,
                      *, continuous=False, an: float = 1.0,
                      precision: Optional[str] = None) -> torch.Tensor:
        """Exact adjoint of ``apply(op, ., widths, weights=weights)`` (node centres
        fixed by ``weights``): g [N, out_dim] -> [N, 3]."""
        n = self.points.shape[0]
        g = g.contiguous()
        assert g.shape == (n, OUT_DIM[op]), (op, tuple(g.shape))
        widths = widths.contiguous().to(g.dtype)
        weights = weights.contiguous().to(g.dtype)
        assert widths.shape == (n,) and weights.shape == (n,)
        out = torch.empty((n, 3), dtype=g.dtype, device=self.device)
        with torch.cuda.device(self.device):
            check(self.L.cagr_apply_adjoint(self._h, OP[op], self._prec(precision), _ptr(g), _dt(g),
                                            _ptr(weights), _ptr(widths), int(continuous), float(an),
                                            _ptr(out), _stream(self.device)))
        return out

    def balance(self, op: str = "A2", attrs: Optional[torch.Tensor] = None, *,
                piece_cost: float = 0.25, max_pieces: int = 16) -> np.ndarray:
        """Balance the self-query walks: cut every walk that costs more than
        `piece_cost` times the mean into node-range pieces of about that cost
        (at most `max_pieces` a wave), dispatched in Morton order, so that
        the longest walks do not set the launch's length; with
        `piece_cost` = 0 (no pieces) the whole waves run longest first within
        each XCD's part of the tree; the batched form's work items get
        per-XCD chunks of equal cost.  Costed by the exact interaction counts
        of `op` on `attrs` (default: unit attributes).  Returns the per-query
        cost used."""
        if attrs is None:
            attrs = torch.ones((self.points.shape[0], IN_DIM[op]), dtype=self.points.dtype,
                               device=self.device)
        far, near, _ = self.counts(op, attrs)
        cost = (far + near).cpu().numpy().astype(np.int64)
        check(self.L.cagr_order_items(self._h, cost.ctypes.data_as(ct.c_void_p)))
        # the pieces are costed on the centres of the counting pass's upward pass
        cnt = np.zeros(2, np.int64)
        with torch.cuda.device(self.device):
            check(self.L.cagr_split_walks(self._h, float(piece_cost), int(max_pieces),
                                          cnt.ctypes.data_as(ct.c_void_p), _stream(self.device)))
        self.walk_pieces = (int(cnt[0]), int(cnt[1]))   # (pieces, waves cut)
        self._balance_args = (op, float(piece_cost), int(max_pieces))
        return cost

    def in_tree_order(self):
        """This cloud as a new treecode whose input order is this tree's point
        order, and that order: row s of the clone is row perm[s] of this cloud.
        The clone's tree is this one (the same Morton codes, the same order
        inside every leaf), so its permutation is the identity and its
        operators read and write [N, .] vectors without gathers or scatters
        through it; a solver that keeps its vectors in this order permutes
        only its right-hand side and its solution.  Walks cut into pieces
        here are cut again on the clone (costed with unit attributes)."""
        perm = torch.from_numpy(self.permutation().astype(np.int64)).to(self.device)
        t = WindingNumberTreecode(self.points[perm], self.tree_depth, theta=self.theta,
                                  leaf_size=self.leaf_size, precision=self.precision)
        if self.walk_pieces[0] > 0:
            op, piece_cost, max_pieces = self._balance_args
            t.balance(op, piece_cost=piece_cost, max_pieces=max_pieces)
        return t, perm

    def unbalance(self) -> None:
        """Back to Morton dispatch order, uncut walks."""
        check(self.L.cagr_order_items(self._h, None))
        self.walk_pieces = (0, 0)

    def adjoint_plan(self, weights: torch.Tensor) -> None:
        """Record the far-field step sequence of the FP32 adjoint for the node
        centres of ``weights`` [N] (an iterative solve applies the adjoint many
        times with the same centres); ``apply_adjoint_planned`` then runs
        without atomics and deterministically."""
        n = self.points.shape[0]
        weights = weights.contiguous()
        assert weights.shape == (n,)
        with torch.cuda.device(self.device):
            check(self.L.cagr_adjoint_plan(self._h, _ptr(weights), _dt(weights),
                                           _stream(self.device)))

    def apply_adjoint_planned(self, op: str, g: torch.Tensor, widths: torch.Tensor, *,
                              continuous=False, an: float = 1.0,
                              precision: Optional[str] = None) -> torch.Tensor:
        """``apply_adjoint`` with the weights of the adjoint plan."""
        n = self.points.shape[0]
        g = g.contiguous()
        assert g.shape == (n, OUT_DIM[op]), (op, tuple(g.shape))
        widths = widths.contiguous().to(g.dtype)
        assert widths.shape == (n,)
        out = torch.empty((n, 3), dtype=g.dtype, device=self.device)
        with torch.cuda.device(self.device):
            check(self.L.cagr_apply_adjoint_planned(self._h, OP[op], self._prec(precision), _ptr(g),
                                                    _dt(g), _ptr(widths), int(continuous),
                                                    float(an), _ptr(out), _stream(self.device)))
        return out

    def upward(self, attrs: torch.Tensor, weights: Optional[torch.Tensor] = None):
        """FP64 node moments (node_attrs [M,C], node_reppoints [M,3], node_weights [M])
        as scatter_point_attrs_to_nodes returns them (preorder node ids)."""
        attrs = attrs.contiguous()
        c = attrs.shape[1]
        m = self.info()["num_nodes"]
        na = torch.empty((m, c), dtype=torch.float64, device=self.device)
        nr = torch.empty((m, 3), dtype=torch.float64, device=self.device)
        nw = torch.empty((m,), dtype=torch.float64, device=self.device)
        if weights is not None:
            weights = weights.contiguous().to(attrs.dtype)
        with torch.cuda.device(self.device):
            check(self.L.cagr_upward(self._h, _ptr(attrs), _dt(attrs), c, _ptr(weights), _ptr(na),
                                     _ptr(nr), _ptr(nw), _stream(self.device)))
        return na, nr, nw

    def counts(self, op: str, attrs: torch.Tensor, weights: Optional[torch.Tensor] = None):
        """Exact (n_far, n_near, n_visit) per query of `op`'s traversal."""
        n = self.points.shape[0]
        attrs = attrs.contiguous()
        outs = [torch.empty(n, dtype=torch.int64, device=self.device) for _ in range(3)]
        if weights is not None:
            weights = weights.contiguous().to(attrs.dtype)
        with torch.cuda.device(self.device):
            check(self.L.cagr_count(self._h, OP[op], _ptr(attrs), _dt(attrs), _ptr(weights),
                                    *(_ptr(o) for o in outs), _stream(self.device)))
        return tuple(outs)

    # ------------------------------------------------------------ reference API
    def _check_vec(self, normals, widths):
        assert self.points.shape == normals.shape
        assert len(widths.shape) == 1
        assert self.points.shape[0] == widths.shape[0]

    def _check_val(self, values, widths):
        assert len(values.shape) == 2
        assert values.shape[0] == self.points.shape[0]
        assert values.shape[1] == 1
        assert len(widths.shape) == 1
        assert self.points.shape[0] == widths.shape[0]

    def forward_A(self, normals, widths):
        """normals: [N, 3], widths: [N,] -> [N, 1] (wn_treecode_func.py:45-80)."""
        self._check_vec(normals, widths)
        return self.apply("A", normals, widths)

    def forward_A_cpu_sign(self, normals, widths):
        self._check_vec(normals, widths)
        return self.apply("A_cpu", normals, widths)

    def forward_AT(self, values, widths):
        """values: [N, 1], widths: [N,] -> [N, 3] (wn_treecode_func.py:82-120)."""
        self._check_val(values, widths)
        return self.apply("AT", values, widths)

    def forward_A21(self, normals, widths):
        self._check_vec(normals, widths)
        return self.apply("A21", normals, widths)

    def forward_A22(self, normals, widths):
        self._check_vec(normals, widths)
        return self.apply("A22", normals, widths)

    def forward_A23(self, normals, widths):
        self._check_vec(normals, widths)
        return self.apply("A23", normals, widths)

    def forward_A2T1(self, values, widths):
        self._check_val(values, widths)
        return self.apply("A2T1", values, widths)

    def forward_A2T2(self, values, widths):
        self._check_val(values, widths)
        return self.apply("A2T2", values, widths)

    def forward_A2T3(self, values, widths):
        self._check_val(values, widths)
        return self.apply("A2T3", values, widths)

    def forward_G(self, normals, widths):
        """normals: [N, 3], widths: [N,] -> [N, 3] (wn_treecode_func.py:356-391)."""
        self._check_vec(normals, widths)
        return self.apply("G", normals, widths)

    # ------------------------------------------------------------ fused extras
    def forward_A2(self, normals, widths, precision=None):
        """[A21, A22, A23] as one [N, 3] tensor from one upward pass and one traversal."""
        self._check_vec(normals, widths)
        return self.apply("A2", normals, widths, precision=precision)

    def forward_A2T_sum(self, values3, widths, precision=None):
        """A2T1(values3[:, :1]) + A2T2(values3[:, 1:2]) + A2T3(values3[:, 2:]) -> [N, 3]."""
        assert values3.shape == self.points.shape
        assert len(widths.shape) == 1 and widths.shape[0] == self.points.shape[0]
        return self.apply("A2T_sum", values3, widths, precision=precision)

    def forward_A3(self, normals, widths, continuous=True, precision=None):
        """(A21 + A22 + A23) / 3 at the points (multiply_by_A3, cuda_kernels.cu:694-763)."""
        self._check_vec(normals, widths)
        return self.apply("A3", normals, widths, continuous=continuous, precision=precision)

    def query(self, queries: torch.Tensor, normals: torch.Tensor, width: float, *, op="A3",
              continuous=True, precision=None, widths: Optional[torch.Tensor] = None,
              an: float = 1.0):
        """Operator at external query points [Nq, 3] (e.g. GaussRecon's sample pass)."""
        assert normals.shape == (self.points.shape[0], IN_DIM[op])
        queries = queries.contiguous()
        normals = normals.contiguous()
        nq = queries.shape[0]
        out = torch.empty((nq, OUT_DIM[op]), dtype=normals.dtype, device=self.device)
        if widths is not None:
            widths = widths.contiguous().to(normals.dtype)
        with torch.cuda.device(self.device):
            check(self.L.cagr_apply_queries(self._h, OP[op], self._prec(precision), _ptr(normals),
                                            _dt(normals), _ptr(queries), _dt(queries), nq,
                                            _ptr(widths), float(width), int(continuous), float(an),
                                            _ptr(out), _stream(self.device)))
        return out

    def grid_query(self, normals: torch.Tensor, lo, step, dims, width: float, *, op="A3",
                   continuous=True, precision=None, an: float = 1.0) -> torch.Tensor:
        """Operator on the regular grid lo + i*step (i < dims), returned as [d0, d1, d2]."""
        assert normals.shape == self.points.shape
        normals = normals.contiguous()
        d = [int(v) for v in dims]
        out = torch.empty(d, dtype=normals.dtype, device=self.device)
        lo_a = (ct.c_double * 3)(*[float(v) for v in lo])
        st_a = (ct.c_double * 3)(*[float(v) for v in step])
        dm_a = (ct.c_int64 * 3)(*d)
        with torch.cuda.device(self.device):
            check(self.L.cagr_grid_query(self._h, OP[op], self._prec(precision), _ptr(normals),
                                         _dt(normals), lo_a, st_a, dm_a, float(width),
                                         int(continuous), float(an), _ptr(out),
                                         _stream(self.device)))
        return out
