"""CAGR orientation solves on the GPU treecode.

* ``cagr_schedule`` -- the reference's own 40-iteration solve
  (CUDA-based/main_CAGR.py:90-174: 31 steepest-descent steps on the normal
  equations with the reference "transpose" A2Tk, a residual reset, 9 CG steps,
  then 8 WNNC G-steps with rescaling), quirks included (the [N,1] sum in A_r,
  the width that keeps shrinking in the CG phase).  A21/A22/A23 run as one
  fused operator and the three A2Tk as one fused sum -- the same decisions and
  moments as the separate calls.
* ``cgnr_solve`` -- the tolerance solve of the benchmark: CG on the normal
  equations A^T A mu = A^T b of the linear operator A mu = -[A21 mu; A22 mu;
  A23 mu] (node centres fixed, unit point weights, one width), with the exact
  adjoint kernel for A^T, FP64 vectors and scalars, stopping when
  ||A^T(b - A mu)|| / ||A^T b|| <= tol in FP64; CG restarts from the true
  residual when it stalls.
* ``cgnr_solve_mixed`` -- the same solve with the FP32-compensated operator
  doing the iterations (all CG vectors, dot products and updates FP64), the
  stopping test and every restart computed with the FP64 operator.
"""
from __future__ import annotations

import time
from typing import Optional

import torch
import torch.nn.functional as F

from ._lib import check
from .wn_treecode import WindingNumberTreecode, _ptr, _stream

PRESETS = {"l0": (0.002, 0.016), "l1": (0.01, 0.04), "l2": (0.02, 0.08), "l3": (0.03, 0.12),
           "l4": (0.04, 0.16), "l5": (0.05, 0.2)}


def cagr_schedule(wt: WindingNumberTreecode, wsmin: float, wsmax: float, iters: int = 40,
                  n_sd: int = 30, wnnc_steps: int = 8, an: float = 1.0,
                  precision: Optional[str] = None, dtype: Optional[torch.dtype] = None):
    """main_CAGR.py:90-174.  Returns (normals [N,3], {'r': ||r||, 'rk1': ||rk1||})."""
    dev = wt.device
    dtype = dtype or wt.points.dtype
    n = wt.points.shape[0]
    widths = torch.ones(n, dtype=dtype, device=dev)
    b3 = torch.full((n, 3), 0.5, dtype=dtype, device=dev)

    def A2(mu, ws):
        return wt.apply("A2", mu, widths * ws, an=an, precision=precision)

    def A2T(s3, ws):
        return wt.apply("A2T_sum", s3, widths * ws, an=an, precision=precision)

    normals = torch.zeros((n, 3), dtype=dtype, device=dev)
    ws = wsmin + 0.5 * (wsmax - wsmin)
    rk1 = A2T(b3, ws)     # computed and overwritten before use, as the reference does
    pk = rk1
    r = rk1
    for i in range(iters):
        ws = wsmin + ((iters - 1 - i) / (iters - 1)) * (wsmax - wsmin)
        if i <= n_sd:
            a_mu = -A2(normals, ws)
            r = A2T(b3, ws) - A2T(a_mu, ws) - 0.0 * normals
            a_r = -(A2(r, ws).sum(1, keepdim=True))
            alpha = (r * r).sum() / (a_r * a_r).sum()
            normals = normals + alpha * r
            if i == n_sd:
                a_mu = -A2(normals, ws)
                rk1 = A2T(b3, ws) - A2T(a_mu, ws) - 0.0 * normals
                pk = rk1
        else:
            at_a_pk = A2T(-A2(pk, ws), ws)
            alpha = (rk1 * rk1).sum() / (at_a_pk * pk).sum()
            normals = normals + alpha * pk
            rk20 = rk1 - alpha * at_a_pk
            beta = (rk20 * rk20).sum() / (rk1 * rk1).sum()
            pk = rk20 + beta * pk
            rk1 = rk20
    res = {"r": float(torch.linalg.norm(r)), "rk1": float(torch.linalg.norm(rk1))}
    for _ in range(wnnc_steps):
        out = wt.apply("G", normals, widths * ws, precision=precision)
        out = F.normalize(out, dim=-1).contiguous()
        normals = out * torch.linalg.norm(normals, dim=-1, keepdim=True)
    return normals, res


class NormalOperator:
    """A mu = -[A21 mu; A22 mu; A23 mu] (as [N,3]) with node centres fixed by
    unit weights, and its exact adjoint; counts operator applications.  With
    ``plan`` the FP32c adjoint replays the walk recorded by the tree's adjoint
    plan (built here, once: the centres never change during a solve)."""

    def __init__(self, wt: WindingNumberTreecode, width: float, an: float = 1.0,
                 plan: bool = False):
        n = wt.points.shape[0]
        self.wt, self.an = wt, an
        self.w64 = torch.full((n,), width, dtype=torch.float64, device=wt.device)
        self.ones64 = torch.ones(n, dtype=torch.float64, device=wt.device)
        self.calls = {"fp64": 0, "fp32c": 0}
        self.plan = plan
        if plan:
            wt.adjoint_plan(self.ones64)

    def A(self, mu, precision):
        self.calls[precision] += 1
        return -self.wt.apply("A2", mu, self.w64, weights=self.ones64, an=self.an,
                              precision=precision)

    def AT(self, g, precision):
        self.calls[precision] += 1
        if self.plan and precision == "fp32c":
            return -self.wt.apply_adjoint_planned("A2", g, self.w64, an=self.an,
                                                  precision=precision)
        return -self.wt.apply_adjoint("A2", g, self.w64, self.ones64, an=self.an,
                                      precision=precision)

    def normal(self, p, precision):
        # the two negations of A and A^T cancel (exactly: IEEE arithmetic is
        # odd-symmetric), so the normal operator applies neither
        self.calls[precision] += 2
        w = self.wt
        y = w.apply("A2", p, self.w64, weights=self.ones64, an=self.an, precision=precision)
        if self.plan and precision == "fp32c":
            return w.apply_adjoint_planned("A2", y, self.w64, an=self.an, precision=precision)
        return w.apply_adjoint("A2", y, self.w64, self.ones64, an=self.an, precision=precision)


class LibCGStep:
    """One CG iteration's vector work by the library (``cagr_cg_step``) on
    the vectors' device: step(x, r, p, q, rr, rr_new) with rr / rr_new 0-dim
    FP64 device tensors (r.r before and after)."""

    def __init__(self, lib, device):
        self.L = lib
        self.device = device
        self.work = torch.empty(int(lib.cagr_cg_work_len()), dtype=torch.float64, device=device)

    def __call__(self, x, r, p, q, rr, rr_new):
        with torch.cuda.device(self.device):
            check(self.L.cagr_cg_step(r.numel(), _ptr(x), _ptr(r), _ptr(p), _ptr(q),
                                      rr.data_ptr(), rr_new.data_ptr(), _ptr(self.work),
                                      self.work.numel(), _stream(self.device)))


def _cg(op, rhs, x0, tol_abs, max_iter, precision, log=None, log_every=500, stall: int = 0,
        check_every: int = 10, step=None):
    """CG on A^T A x = rhs (FP64 vectors); stops on the recurrence residual.
    An iteration's vector work after q = A^T A p is ``step`` (default: the
    library's ``cagr_cg_step`` for ``op.wt``, three launches, reductions in a
    fixed order).  The
    scalars alpha and beta stay on the device and the host reads the
    residual history only every ``check_every`` iterations, so the host never
    waits for the GPU inside an iteration (a run may take up to
    ``check_every`` - 1 iterations more than it needs; they only lower the
    residual).  ``stall`` > 0: also stops once the residual has not fallen
    below 0.9 of its best value for that many iterations (an operator whose
    forward and adjoint are not exact transposes of each other -- the FP32c
    pair -- makes CG stagnate near its precision floor; the caller restarts
    from the true residual).  ``log(msg)`` (optional) gets the recurrence
    residual every ``log_every`` iterations."""
    x = x0.clone().contiguous()
    r = (rhs - op.normal(x, precision) if x0.abs().max() > 0 else rhs.clone()).contiguous()
    p = r.clone()
    rv = r.view(-1)
    # hist[0] = r.r before the block, hist[j + 1] = r.r after its iteration j
    hist = torch.empty(check_every + 1, dtype=torch.float64, device=r.device)
    hist[0] = torch.dot(rv, rv)
    if step is None:
        step = LibCGStep(op.wt.L, r.device)
    it = 0
    best, since = float(hist[0]), 0
    if best ** 0.5 <= tol_abs:
        return x, 0
    while it < max_iter:
        n = min(check_every, max_iter - it)
        for j in range(n):
            q = op.normal(p, precision).contiguous()
            step(x, r, p, q, hist[j], hist[j + 1])
        rrs = hist[1:n + 1].tolist()   # the one host read of the block
        hist[0] = hist[n]
        done = False
        for j, v in enumerate(rrs):
            k = it + j + 1
            if log is not None and k % log_every == 0:
                log(f"cg ({precision}) iteration {k}: residual {v ** 0.5:.3e} (target {tol_abs:.3e})")
            if v ** 0.5 <= tol_abs:
                done = True
            elif v < 0.81 * best:
                best, since = v, 0
            else:
                since += 1
                if stall and since >= stall and not done:
                    if log is not None:
                        log(f"cg ({precision}) stalled at iteration {k}: residual {v ** 0.5:.3e}")
                    done = True
        it += n
        if done:
            break
    return x, it


def _cgnr(wt: WindingNumberTreecode, width: float, tol: float, b: float, max_iter: int,
          max_restarts: int, an: float, inner: str, plan: bool, log, stall: int,
          tree_order: bool = True):
    """CGNR to ||A^T(b - A mu)|| / ||A^T b|| <= tol, verified with the FP64
    operator: CG on the correction equation with the ``inner`` operator until
    it converges or stalls (``stall`` iterations without a 10 % gain; 0 =
    never), then the true residual in FP64 and a restart from it.  With
    ``tree_order`` the solve runs on the cloud in its tree's point order
    (``in_tree_order``, built inside the timed region: the same tree, every
    operator's vectors read and written without the permutation) and the
    solution is returned in the input order."""
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    perm = None
    if tree_order:
        wt, perm = wt.in_tree_order()
    op = NormalOperator(wt, width, an, plan=plan)
    n = wt.points.shape[0]
    bt = torch.full((n, 3), b, dtype=torch.float64, device=wt.device)
    rhs = op.AT(bt, "fp64")
    nrhs = float(torch.linalg.norm(rhs))
    mu = torch.zeros_like(rhs)
    r = rhs.clone()
    iters, restarts, history = 0, 0, []
    while True:
        res = float(torch.linalg.norm(r)) / nrhs
        history.append(res)
        if res <= tol or restarts > max_restarts or iters >= max_iter:
            break
        d, it = _cg(op, r, torch.zeros_like(r), tol * nrhs, max_iter - iters, inner, log,
                    stall=stall)
        iters += it
        mu += d
        r = op.AT(bt - op.A(mu, "fp64"), "fp64")
        restarts += 1
    if perm is not None:
        out = torch.empty_like(mu)
        out[perm] = mu
        mu = out
    torch.cuda.synchronize()
    return mu, {"iters": iters, "restarts": restarts - 1, "rel_residual": res, "history": history,
                "seconds": time.perf_counter() - t0, "calls": dict(op.calls)}


def cgnr_solve(wt: WindingNumberTreecode, width: float, tol: float = 1e-6, b: float = 0.5,
               max_iter: int = 5000, an: float = 1.0, log=None, stall: int = 1500,
               max_restarts: int = 10, tree_order: bool = True):
    """Plain FP64 CGNR to ||A^T(b - A mu)|| / ||A^T b|| <= tol, with the same
    restart rule as the mixed solve (``stall`` = 0: one CG run).
    ``tree_order``: solve on the cloud in its tree's point order (the
    solution comes back in the input order)."""
    return _cgnr(wt, width, tol, b, max_iter, max_restarts, an, "fp64", False, log, stall,
                 tree_order)


def cgnr_solve_mixed(wt: WindingNumberTreecode, width: float, tol: float = 1e-6, b: float = 0.5,
                     max_iter: int = 5000, max_restarts: int = 10, an: float = 1.0,
                     plan: bool = True, log=None, stall: int = 1500, tree_order: bool = True):
    """CGNR whose operator applications run in FP32c while every vector, dot
    product and update stays FP64; when the recurrence converges -- or
    stagnates (``stall`` iterations without a 10 % gain: the FP32c forward and
    adjoint are transposes of each other only to FP32 rounding, which puts a
    floor near 2e-6 under CG at 1M points) -- the residual is recomputed with
    the FP64 operator and exact adjoint and CG restarts on the correction from
    that true residual (iterative refinement; the correction is small, so its
    floor is lower) until the FP64 test passes.  ``plan``: the FP32c adjoint
    runs through an adjoint plan (its build is inside the timed region).
    ``tree_order``: as ``cgnr_solve``."""
    return _cgnr(wt, width, tol, b, max_iter, max_restarts, an, "fp32c", plan, log, stall,
                 tree_order)
