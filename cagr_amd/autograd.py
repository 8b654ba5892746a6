"""Differentiable treecode operator.

``treecode_op(wt, mu, widths, op="A2")`` applies the treecode with node
centres fixed by ``weights`` (default: every point weight 1, i.e. the
centroid of each node), which makes the operator linear in ``mu``; its
backward is the exact adjoint kernel (cagr_apply_adjoint), so gradients of any
loss on the output -- the normal-equation residual, a loss on the winding
field -- are exact gradients of the discrete operator, including its far-field
approximations.  (The reference's own A2Tk is the adjoint of the
direction-reversed operator, not of A2k, and its centres move with |mu|, which
makes its A nonlinear in mu.)  Gradients flow to ``mu`` only.
"""
from __future__ import annotations

from typing import Optional

import torch

from .wn_treecode import WindingNumberTreecode


class _TreecodeOp(torch.autograd.Function):
    @staticmethod
    def forward(ctx, mu, wt, op, widths, weights, continuous, an, precision):
        ctx.wt, ctx.op, ctx.continuous, ctx.an, ctx.precision = wt, op, continuous, an, precision
        ctx.save_for_backward(widths, weights)
        return wt.apply(op, mu.detach(), widths, weights=weights, continuous=continuous, an=an,
                        precision=precision)

    @staticmethod
    def backward(ctx, gout):
        widths, weights = ctx.saved_tensors
        gmu = None
        if ctx.needs_input_grad[0]:
            gmu = ctx.wt.apply_adjoint(ctx.op, gout.contiguous(), widths, weights,
                                       continuous=ctx.continuous, an=ctx.an,
                                       precision=ctx.precision)
        return gmu, None, None, None, None, None, None, None


def fixed_weights(wt: WindingNumberTreecode, dtype=None) -> torch.Tensor:
    """Unit point weights: node centres at the node centroids."""
    return torch.ones(wt.points.shape[0], dtype=dtype or wt.points.dtype, device=wt.device)


def treecode_op(wt: WindingNumberTreecode, mu: torch.Tensor, widths: torch.Tensor,
                op: str = "A2", weights: Optional[torch.Tensor] = None, *,
                continuous: bool = False, an: float = 1.0,
                precision: Optional[str] = None) -> torch.Tensor:
    """Linear treecode operator `op` applied to mu [N,3], differentiable in mu."""
    if op not in ("A", "A_cpu", "A21", "A22", "A23", "A3", "A2"):
        raise ValueError(f"no exact adjoint for operator {op}")
    if weights is None:
        weights = fixed_weights(wt, mu.dtype)
    widths = widths.to(mu.dtype).contiguous()
    weights = weights.to(mu.dtype).contiguous()
    return _TreecodeOp.apply(mu.contiguous(), wt, op, widths, weights, continuous, an, precision)
