"""torch.autograd wrapper of recompose: the forward pass is the MGARD
recompose kernel, the backward pass its transpose (prodm_recompose_adjoint),
so gradients of a quantity of interest flow from the reconstructed field back
to the stored coefficients."""
import torch

from . import mdr


class Recompose(torch.autograd.Function):
    """u = recompose(coefficients); coefficients in the decompose layout,
    dense, dims[0] slowest."""

    @staticmethod
    def forward(ctx, coef, dims, L):
        ctx.dims, ctx.L = list(dims), L
        return mdr.recompose(coef.contiguous(), dims, L)

    @staticmethod
    def backward(ctx, grad_u):
        return mdr.recompose_adjoint(grad_u.contiguous(), ctx.dims, ctx.L), None, None


def recompose(coef, dims, L):
    return Recompose.apply(coef, dims, L)
