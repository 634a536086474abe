"""torch.autograd wrapper: forward = an evaluation kernel, backward = the adjoint kernel."""

import torch


class SmolyakEval(torch.autograd.Function):
    """y = I^Lambda[f](x); dL/dx = vjp(x, dL/dy) (SmolyakBarycentricInterpolator.vjp)."""

    @staticmethod
    def forward(ctx, x, interpolant, precision):
        ctx.interpolant = interpolant
        ctx.precision = precision
        ctx.save_for_backward(x)
        return interpolant.evaluate(x.detach(), precision=precision)

    @staticmethod
    def backward(ctx, grad_y):
        (x,) = ctx.saved_tensors
        gx = ctx.interpolant.vjp(x.detach(), grad_y.contiguous(), precision=ctx.precision)
        return gx.to(x.dtype), None, None
