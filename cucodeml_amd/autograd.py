"""torch.autograd binding of the engine's likelihood and adjoint gradient,
for the gradcheck harness: ``LnL.apply(x, eng)`` is lnL at codeml's
parameter vector x (float64, host tensor); its backward returns the adjoint
gradient from the same engine call (bsa_lnl_grad), scaled by the incoming
gradient.  The engine does the work on the GPU; torch only carries x and
the result."""
import numpy as np
import torch


class LnL(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, eng):
        xv = x.detach().cpu().numpy().astype(np.float64)
        neg, g = eng.lnl_grad(xv)
        ctx.save_for_backward(torch.from_numpy(-g))   # d lnL / dx
        return x.new_tensor(-neg)

    @staticmethod
    def backward(ctx, grad_out):
        (g,) = ctx.saved_tensors
        return grad_out * g.to(grad_out.device, grad_out.dtype), None


def lnl(x, eng):
    """lnL(x) with autograd support through the engine's adjoint."""
    return LnL.apply(x, eng)
