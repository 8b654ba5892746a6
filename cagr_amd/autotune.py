"""Autotuner of the batched treecode: leaf size, opening parameter theta and
the launch shape of the persistent traversal, under a stated error budget.

Accuracy.  For each (leaf size, theta) the forest is built once and its FP64
A2 is compared with FP64 direct sums at a random sample of queries in every
cloud (the direct sum is the theta -> 0 limit, evaluated by the same library
on a tree with theta = 1e-6).  A cloud's error is max|treecode - direct| /
max|direct| over its sampled rows; a configuration's error is the largest
over the clouds.  A configuration is admissible when that error is at most
``budget`` -- a tolerance the caller states (default 5e-2).  The treecode's
own error against direct sums at width 0.002 spans 3e-4 .. 1.6e-2 on single
benchmark clouds and, as the worst cloud of the 64-cloud batch, 2.0e-2
(leaf 64, theta 0.4) .. 5.9e-2 (leaf 16, theta 0.6), 3.1e-2 at the reference
setting leaf 32 / theta 0.5.

Speed.  Every admissible (leaf, theta) is timed at every block shape on whole
matvec pairs (A2 fused, then the fused A2T sum), in ``rounds`` interleaved
rounds (each round visits every configuration once, so slow drift of the
clock affects all alike); a configuration's time is its median over the
rounds.  The report says whether the winner was also the fastest in every
round (``stable``) and by how much it leads the next configuration.

Size classes.  ``autotune_classes`` splits the clouds by size and tunes every
class on its own (small clouds may prefer other leaves than large ones);
``ClassBatches`` then runs one forest, with its own tuned setting, per class.
"""
from __future__ import annotations

import itertools
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from .batched import BatchedTreecode
from .wn_treecode import WindingNumberTreecode

DEFAULT_BUDGET = 5e-2


@dataclass
class TuneResult:
    leaf_size: int
    theta: float
    waves_per_block: int
    ms_per_pair: float
    err: float
    rounds_ms: List[float] = field(default_factory=list)


@dataclass
class TuneReport:
    best: TuneResult
    table: List[TuneResult] = field(default_factory=list)
    budget: float = 0.0
    errors: Dict[Tuple[int, float], float] = field(default_factory=dict)
    stable: bool = False      # the winner was the fastest in every round
    lead: float = 0.0         # next-best median time / winner's median time - 1


def _direct_sample(clouds, mu_list, width, samples_per_cloud, seed):
    """FP64 direct-sum A2 (three columns) at sampled points of every cloud."""
    rng = np.random.default_rng(seed)
    idx, ref = [], []
    for pts, mu in zip(clouds, mu_list):
        n = pts.shape[0]
        k = min(samples_per_cloud, n)
        sel = np.sort(rng.choice(n, k, replace=False)) if n else np.zeros(0, np.int64)
        if k == 0:
            idx.append(sel)
            ref.append(None)
            continue
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


def _sampled_error(bt, mu, width, idx, ref) -> float:
    w = torch.full((bt.points.shape[0],), width, dtype=bt.points.dtype, device=bt.device)
    out = bt.forward_A2(mu, w, precision="fp64").double()
    worst = 0.0
    for c, part in enumerate(bt.split(out)):
        if ref[c] is None:
            continue
        sel = torch.from_numpy(idx[c]).to(out.device)
        den = float(ref[c].abs().max())
        if den > 0:
            worst = max(worst, float((part[sel] - ref[c]).abs().max()) / den)
    return worst


def autotune(clouds: Sequence[torch.Tensor], mu_list: Sequence[torch.Tensor], width: float,
             leaf_sizes=(16, 32, 64), thetas=(0.4, 0.5, 0.6), waves_per_block=(1, 2, 4, 8),
             precision: str = "fp32c", steps: int = 5, rounds: int = 3,
             samples_per_cloud: int = 64, budget: float = DEFAULT_BUDGET, seed: int = 0,
             log=None) -> TuneReport:
    """Fastest (leaf, theta, waves per block) whose sampled error against
    direct sums is within ``budget`` (see the module docstring).  If no
    configuration meets the budget, the most accurate one is returned."""
    idx, ref = _direct_sample(clouds, mu_list, width, samples_per_cloud, seed)
    mu = torch.cat([m.contiguous() for m in mu_list]).to(clouds[0].dtype).contiguous()
    errs: Dict[Tuple[int, float], float] = {}
    forests = {}
    for leaf, theta in itertools.product(leaf_sizes, thetas):
        bt = BatchedTreecode(list(clouds), leaf_size=leaf, theta=theta, precision=precision)
        err = _sampled_error(bt, mu, width, idx, ref)
        errs[(leaf, theta)] = err
        if log:
            log(f"[autotune] leaf {leaf} theta {theta}: sampled error {err:.3e} "
                f"({'within' if err <= budget else 'over'} budget {budget:.1e})")
        if err <= budget:
            bt.balance("A2", mu)
            forests[(leaf, theta)] = bt
        else:
            del bt
    if not forests:   # nothing meets the budget: the most accurate configuration
        leaf, theta = min(errs, key=errs.get)
        bt = BatchedTreecode(list(clouds), leaf_size=leaf, theta=theta, precision=precision)
        bt.balance("A2", mu)
        forests[(leaf, theta)] = bt
    w = torch.full((mu.shape[0],), width, dtype=clouds[0].dtype, device=clouds[0].device)
    configs = [(lt, wpb) for lt in forests for wpb in waves_per_block]
    times = {c: [] for c in configs}
    for _ in range(rounds):
        for lt, wpb in configs:
            bt = forests[lt]
            bt.set_launch(wpb)
            times[(lt, wpb)].append(_time_pairs(bt, mu, w, precision, steps))
    table = []
    for (lt, wpb), ts in times.items():
        r = TuneResult(lt[0], lt[1], wpb, float(np.median(ts)), errs[lt], ts)
        table.append(r)
        if log:
            log(f"[autotune] leaf {lt[0]} theta {lt[1]} waves/block {wpb}: median "
                f"{r.ms_per_pair:.3f} ms/pair over {rounds} rounds {['%.3f' % x for x in ts]}")
    table.sort(key=lambda r: r.ms_per_pair)
    best = table[0]
    stable = all(min(table, key=lambda r: r.rounds_ms[i]) is best for i in range(rounds))
    lead = table[1].ms_per_pair / best.ms_per_pair - 1.0 if len(table) > 1 else float("inf")
    return TuneReport(best=best, table=table, budget=budget, errors=errs, stable=stable, lead=lead)


# ---------------------------------------------------------------- size classes
def size_classes(sizes: Sequence[int], bounds=(10_000, 60_000)) -> List[np.ndarray]:
    """Cloud indices per size class: [0, b0), [b0, b1), ..., [b_last, inf)."""
    sizes = np.asarray(sizes)
    edges = [0, *bounds, np.inf]
    out = []
    for lo, hi in zip(edges[:-1], edges[1:]):
        sel = np.nonzero((sizes >= lo) & (sizes < hi))[0]
        if sel.size:
            out.append(sel)
    return out


def autotune_classes(clouds, mu_list, width, bounds=(10_000, 60_000), log=None, **kw):
    """autotune() per size class; returns [(cloud indices, TuneReport)]."""
    res = []
    for sel in size_classes([c.shape[0] for c in clouds], bounds):
        if log:
            log(f"[autotune] class of {len(sel)} clouds, sizes "
                f"{min(clouds[i].shape[0] for i in sel)}..{max(clouds[i].shape[0] for i in sel)}")
        rep = autotune([clouds[i] for i in sel], [mu_list[i] for i in sel], width, log=log, **kw)
        res.append((sel, rep))
    return res


class ClassBatches:
    """One tuned forest per size class over the same clouds: forward_A2 /
    forward_A2T_sum take and return the clouds' arrays back to back in the
    original cloud order, like a single BatchedTreecode."""

    def __init__(self, clouds, classes, precision: str = "fp32c"):
        self.sizes = np.array([c.shape[0] for c in clouds], np.int64)
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)])
        self.parts = []
        for sel, rep in classes:
            b = rep.best
            bt = BatchedTreecode([clouds[i] for i in sel], leaf_size=b.leaf_size, theta=b.theta,
                                 precision=precision)
            bt.set_launch(b.waves_per_block)
            rows = np.concatenate([np.arange(self.offsets[i], self.offsets[i + 1]) for i in sel])
            self.parts.append((bt, torch.from_numpy(rows).to(clouds[0].device)))

    def balance(self, mu):
        for bt, rows in self.parts:
            bt.balance("A2", mu[rows].contiguous())

    def _apply(self, name, x, w, precision):
        out = torch.empty((x.shape[0], 3), dtype=x.dtype, device=x.device)
        for bt, rows in self.parts:
            out[rows] = getattr(bt, name)(x[rows].contiguous(), w[rows].contiguous(),
                                          precision=precision)
        return out

    def forward_A2(self, mu, w, precision=None):
        return self._apply("forward_A2", mu, w, precision)

    def forward_A2T_sum(self, s3, w, precision=None):
        return self._apply("forward_A2T_sum", s3, w, precision)
