"""Autotuner of the batched treecode: leaf size, opening parameter theta and
the launch shape of the persistent traversal.

For each (leaf size, theta) the forest is built once; the treecode error of
that configuration is measured against FP64 direct sums at a random sample of
queries in every cloud (the direct sum is the theta -> 0 limit, evaluated by
the same library on a tree with theta = 1e-6), and every block shape is timed
on whole matvec pairs (A2 fused, then the fused A2T sum).  The chosen
configuration is the fastest whose sampled error does not exceed the error of
the reference setting (theta = 0.5, leaf 32) by more than `slack`: the tuner
trades speed only at equal or better accuracy.
"""
from __future__ import annotations

import itertools
import time
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np
import torch

from .batched import BatchedTreecode
from .wn_treecode import WindingNumberTreecode


@dataclass
class TuneResult:
    leaf_size: int
    theta: float
    waves_per_block: int
    ms_per_pair: float
    err: float


@dataclass
class TuneReport:
    best: TuneResult
    table: List[TuneResult] = field(default_factory=list)
    budget: float = 0.0


def _direct_sample(clouds, mu_list, width, samples_per_cloud, seed):
    """FP64 direct-sum A2 (three columns) at sampled points of every cloud."""
    rng = np.random.default_rng(seed)
    idx, ref = [], []
    for c, (pts, mu) in enumerate(zip(clouds, mu_list)):
        n = pts.shape[0]
        k = min(samples_per_cloud, n)
        sel = np.sort(rng.choice(n, k, replace=False))
        wt = WindingNumberTreecode(pts.double(), leaf_size=64, theta=1e-6, precision="fp64")
        q = pts.double()[torch.from_numpy(sel).to(pts.device)]
        out = wt.query(q, mu.double(), width, op="A2", continuous=False, precision="fp64")
        idx.append(sel)
        ref.append(out)
    return idx, ref


def _time_pairs(bt, mu, w, precision, steps):
    for _ in range(2):
        a = bt.forward_A2(mu, w, precision=precision)
        bt.forward_A2T_sum(-a, w, precision=precision)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        a = bt.forward_A2(mu, w, precision=precision)
        bt.forward_A2T_sum(-a, w, precision=precision)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps * 1e3


def autotune(clouds: Sequence[torch.Tensor], mu_list: Sequence[torch.Tensor], width: float,
             leaf_sizes=(16, 32, 64), thetas=(0.4, 0.5, 0.6), waves_per_block=(1, 2, 4, 8),
             precision: str = "fp32c", steps: int = 5, samples_per_cloud: int = 64,
             slack: float = 1.0, seed: int = 0, log=None) -> TuneReport:
    idx, ref = _direct_sample(clouds, mu_list, width, samples_per_cloud, seed)
    mu = torch.cat([m.contiguous() for m in mu_list]).to(clouds[0].dtype).contiguous()
    table = []
    errs = {}
    for leaf, theta in itertools.product(leaf_sizes, thetas):
        bt = BatchedTreecode(list(clouds), leaf_size=leaf, theta=theta, precision=precision)
        w = torch.full((bt.points.shape[0],), width, dtype=bt.points.dtype, device=bt.device)
        out = bt.forward_A2(mu, w, precision="fp64").double()
        e_num, e_den = 0.0, 0.0
        for c, part in enumerate(bt.split(out)):
            sel = torch.from_numpy(idx[c]).to(out.device)
            e_num = max(e_num, float((part[sel] - ref[c]).abs().max()))
            e_den = max(e_den, float(ref[c].abs().max()))
        err = e_num / e_den
        errs[(leaf, theta)] = err
        bt.balance("A2", mu)
        for wpb in waves_per_block:
            bt.set_launch(wpb)
            ms = _time_pairs(bt, mu, w, precision, steps)
            r = TuneResult(leaf, theta, wpb, ms, err)
            table.append(r)
            if log:
                log(f"[autotune] leaf {leaf} theta {theta} waves/block {wpb}: {ms:.3f} ms/pair, "
                    f"sampled error {err:.3e}")
        del bt
    budget = errs.get((32, 0.5), min(errs.values())) * slack
    ok = [r for r in table if r.err <= budget * (1 + 1e-12)]
    best = min(ok, key=lambda r: r.ms_per_pair)
    return TuneReport(best=best, table=table, budget=budget)
