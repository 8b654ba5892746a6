"""Batched form: many clouds of uneven size in one forest.

``BatchedTreecode`` holds one octree per cloud (each exactly the tree a
single-cloud ``WindingNumberTreecode`` would build) in one device forest; one
upward pass covers every tree and one persistent traversal launch serves every
cloud's queries, pulling 64-query work items from an atomic queue (longest
first once ``balance`` has costed them).  Inputs and outputs are the clouds'
arrays stored back to back.
"""
from __future__ import annotations

import ctypes as ct
from typing import Optional, Sequence

import numpy as np
import torch

from . import _lib
from ._lib import check
from .wn_treecode import WindingNumberTreecode, _dt, _ptr, _stream


class BatchedTreecode(WindingNumberTreecode):
    def __init__(self, clouds: Sequence[torch.Tensor], max_tree_depth: int = 15, *,
                 theta: float = 0.5, leaf_size: int = 32, precision: Optional[str] = None):
        """clouds: list of [n_c, 3] tensors on one ROCm device (same dtype)."""
        assert len(clouds) >= 1
        dev = clouds[0].device
        if not clouds[0].is_cuda:
            raise RuntimeError("cagr_amd runs on the GPU only: move the points to a ROCm device")
        for c in clouds:
            assert c.dim() == 2 and c.shape[1] == 3 and c.device == dev and c.dtype == clouds[0].dtype
        self.L = _lib.load()
        self.is_cuda = True
        self.device = dev
        self.points = torch.cat([c.contiguous() for c in clouds]).contiguous()
        self.sizes = np.array([c.shape[0] for c in clouds], np.int64)
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.int64)
        self.tree_depth = max_tree_depth
        self.theta = float(theta)
        self.leaf_size = int(leaf_size)
        if precision is None:
            precision = "fp64" if self.points.dtype == torch.float64 else "fp32c"
        self.precision = precision
        self._h = ct.c_void_p()
        with torch.cuda.device(dev):
            check(self.L.cagr_forest_build(_ptr(self.points), self.offsets.ctypes.data_as(ct.c_void_p),
                                           len(clouds), _dt(self.points), max_tree_depth,
                                           self.leaf_size, self.theta, _stream(dev),
                                           ct.byref(self._h)))
        self._ref_layout = None

    @property
    def num_clouds(self) -> int:
        return len(self.sizes)

    def node_offsets(self) -> np.ndarray:
        out = np.zeros(self.num_clouds + 1, np.int64)
        check(self.L.cagr_forest_roots(self._h, out.ctypes.data_as(ct.c_void_p)))
        return out

    def split(self, t: torch.Tensor):
        """Per-cloud views of a concatenated [N_total, ...] tensor."""
        return [t[a:b] for a, b in zip(self.offsets[:-1], self.offsets[1:])]

    def set_launch(self, waves_per_block: int = 4, grid_blocks: int = 0) -> None:
        check(self.L.cagr_set_launch(self._h, 1, int(waves_per_block), int(grid_blocks)))

