"""torch.autograd wrapper of the device L^p objective and its backward kernel.

J_p(a) = 2n ||a*a||_p / (sum a)^2 for every row of a [B x n]; the backward
pass calls sidon_lp_grad, which recomputes a*a on the device instead of saving
it.  No reference counterpart: the reference forms only the scaled descent
direction G/R - M inside its fused kernel (src/kernels.cuh:165-172, :244).
"""
from __future__ import annotations

import torch


class LpObjective(torch.autograd.Function):
    @staticmethod
    def forward(ctx, a: torch.Tensor, p: float) -> torch.Tensor:
        from . import lp_grad
        a = a.contiguous()
        J, _ = lp_grad(a, p, want_J=True, want_grad=False)
        ctx.save_for_backward(a)
        ctx.p = p
        return J

    @staticmethod
    def backward(ctx, gJ: torch.Tensor):
        from . import lp_grad
        (a,) = ctx.saved_tensors
        _, g = lp_grad(a, ctx.p, want_J=False, want_grad=True)
        return g * gJ.to(g.dtype).unsqueeze(1), None


def lp_objective(a: torch.Tensor, p: float) -> torch.Tensor:
    return LpObjective.apply(a, p)
