"""Differentiable CBM product: ``y = M x`` with the hand-written adjoint
``dx = M^T dy`` (reverse tree gather, then the Delta^T product).  The
reference has no backward pass; its closest piece is the CPU op
s_spmm_csr_t_int32 (cbm/cbm_extensions.cpp:742-782)."""
import torch


class CbmProduct(torch.autograd.Function):
    """Forward ``M x`` on the GPU; backward ``M^T g``.

    If ``x`` is BF16 the incoming gradient is stored in BF16 too (the product
    accumulates in FP32 either way) and the returned gradient is cast to BF16.
    """

    @staticmethod
    def forward(ctx, x, handle):
        y = torch.empty((handle.num_rows, x.size(1)), dtype=torch.float32, device=x.device)
        handle.forward(x, y)
        ctx.handle = handle
        ctx.in_dtype = x.dtype
        return y

    @staticmethod
    def backward(ctx, gy):
        handle = ctx.handle
        g = gy.to(ctx.in_dtype) if ctx.in_dtype == torch.bfloat16 else gy.to(torch.float32)
        g = g.contiguous()
        dx = torch.empty((handle.num_rows, g.size(1)), dtype=torch.float32, device=g.device)
        handle.matmul_t(g, dx)
        return dx.to(ctx.in_dtype), None


def cbm_apply(handle, x):
    """``M @ x`` with autograd support; ``handle`` is a :class:`cusparseCBM`."""
    return CbmProduct.apply(x, handle)
