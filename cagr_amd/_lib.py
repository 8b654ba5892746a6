"""ctypes binding of the cagr_amd C ABI (include/cagr_amd.h).

The library is ``cagr_amd/libcagr_amd.so``, built in-tree by
``cagr_amd._build.build()``.  There is no fallback: if the library is
missing or fails to load, every entry point raises.
"""
from __future__ import annotations

import ctypes as ct
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcagr_amd.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "cagr_amd.h")

F32, F64 = 0, 1
PREC_FP64, PREC_FP32C = 0, 1
OP = {
    "A": 0, "A_cpu": 1, "AT": 2, "G": 3,
    "A21": 4, "A22": 5, "A23": 6,
    "A2T1": 7, "A2T2": 8, "A2T3": 9,
    "A3": 10, "A2": 11, "A2T_sum": 12,
}
IN_DIM = {"A": 3, "A_cpu": 3, "AT": 1, "G": 3, "A21": 3, "A22": 3, "A23": 3, "A2T1": 1,
          "A2T2": 1, "A2T3": 1, "A3": 3, "A2": 3, "A2T_sum": 3}
OUT_DIM = {"A": 1, "A_cpu": 1, "AT": 3, "G": 3, "A21": 1, "A22": 1, "A23": 1, "A2T1": 3,
           "A2T2": 3, "A2T3": 3, "A3": 1, "A2": 3, "A2T_sum": 3}

# every symbol include/cagr_amd.h declares
EXPORTS = (
    "cagr_last_error", "cagr_version", "cagr_tree_build", "cagr_tree_destroy", "cagr_tree_info",
    "cagr_tree_export", "cagr_tree_permutation", "cagr_upward", "cagr_apply",
    "cagr_apply_weighted", "cagr_apply_adjoint", "cagr_apply_queries",
    "cagr_apply_queries_noalloc", "cagr_apply_moments", "cagr_grid_query",
    "cagr_count", "cagr_forest_build", "cagr_forest_roots", "cagr_set_launch",
    "cagr_order_items", "cagr_split_walks", "cagr_profile", "cagr_profile_read",
    "cagr_cg_step", "cagr_cg_work_len", "cagr_adjoint_plan",
    "cagr_apply_adjoint_planned", "cagr_pgr_build", "cagr_pgr_destroy", "cagr_pgr_info",
    "cagr_pgr_read", "cagr_pgr_box", "cagr_pgr_mesh", "cagr_pgr_mesh_read",
)

_P, _I, _D, _C = ct.c_void_p, ct.c_int64, ct.c_double, ct.c_int
_SIGS = {
    "cagr_last_error": (ct.c_char_p, []),
    "cagr_version": (_C, []),
    "cagr_tree_build": (_C, [_P, _I, _C, _C, _C, _D, _P, ct.POINTER(_P)]),
    "cagr_tree_destroy": (_C, [_P]),
    "cagr_tree_info": (_C, [_P, _P, _P, _P, _P]),
    "cagr_tree_export": (_C, [_P] * 8),
    "cagr_tree_permutation": (_C, [_P, _P]),
    "cagr_upward": (_C, [_P, _P, _C, _C, _P, _P, _P, _P, _P]),
    "cagr_apply": (_C, [_P, _C, _C, _P, _C, _P, _C, _D, _P, _P]),
    "cagr_apply_weighted": (_C, [_P, _C, _C, _P, _C, _P, _P, _C, _D, _P, _P]),
    "cagr_apply_adjoint": (_C, [_P, _C, _C, _P, _C, _P, _P, _C, _D, _P, _P]),
    "cagr_apply_queries": (_C, [_P, _C, _C, _P, _C, _P, _C, _I, _P, _D, _C, _D, _P, _P]),
    "cagr_apply_queries_noalloc": (_C, [_P, _C, _C, _P, _C, _P, _C, _I, _P, _D, _C, _D, _P, _P]),
    "cagr_apply_moments": (_C, [_P, _C, _C, _P, _C, _P, _P, _P, _C, _I, _P, _C, _D, _P, _P]),
    "cagr_grid_query": (_C, [_P, _C, _C, _P, _C, _P, _P, _P, _D, _C, _D, _P, _P]),
    "cagr_count": (_C, [_P, _C, _P, _C, _P, _P, _P, _P, _P]),
    "cagr_forest_build": (_C, [_P, _P, _I, _C, _C, _C, _D, _P, ct.POINTER(_P)]),
    "cagr_forest_roots": (_C, [_P, _P]),
    "cagr_set_launch": (_C, [_P, _C, _C, _I]),
    "cagr_order_items": (_C, [_P, _P]),
    "cagr_split_walks": (_C, [_P, _D, _C, _P, _P]),
    "cagr_profile": (_C, [_P, _C]),
    "cagr_profile_read": (_C, [_P, _C, _P, _P, _P, _I]),
    "cagr_cg_step": (_C, [_I, _P, _P, _P, _P, _P, _P, _P, _I, _P]),
    "cagr_cg_work_len": (_I, []),
    "cagr_adjoint_plan": (_C, [_P, _P, _C, _P]),
    "cagr_apply_adjoint_planned": (_C, [_P, _C, _C, _P, _C, _P, _C, _D, _P, _P]),
    "cagr_pgr_build": (_C, [_P, _I, _C, _C, _P, ct.POINTER(_P)]),
    "cagr_pgr_destroy": (None, [_P]),
    "cagr_pgr_info": (_C, [_P, _P, _P, _P, _P]),
    "cagr_pgr_read": (_C, [_P, _P, _P, _P, _P]),
    "cagr_pgr_box": (_C, [_P, _P]),
    "cagr_pgr_mesh": (_C, [_P, _P, ct.c_float, ct.c_float, _P, _P, _P]),
    "cagr_pgr_mesh_read": (_C, [_P, _P, _P, _P]),
}

_lib = None


class CagrError(RuntimeError):
    pass


def load():
    """Load libcagr_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise CagrError(
                f"{LIB_PATH} is missing: build the HIP library first "
                "(python -m cagr_amd._build, or __graft_entry__.build())")
        L = ct.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != 0:
        msg = load().cagr_last_error().decode(errors="replace")
        raise CagrError(f"cagr_amd error {rc}: {msg}")
