"""A two-layer GCN over the normalised adjacency D^-1/2 A D^-1/2 whose sparse
products run on the CBM kernels (forward) and the hand-written adjoint
(backward).  The reference's harness imports ``GCNConv`` but never builds a
model (benchmark/utilities.py:11); this is the training step its products
serve.

Layer order: features are multiplied by the weight first and then propagated,
``H = relu(Â (X W1))``, ``out = Â (H W2)``, so both sparse products run at the
hidden width.
"""
import torch

from .autograd import cbm_apply
from .cbm_cusparse import cusparseCBM


class CsrProduct(torch.autograd.Function):
    """``y = A x`` with rocSPARSE, backward ``A^T g`` (the CSR-transpose gradient)."""

    @staticmethod
    def forward(ctx, x, csr):
        y = torch.empty((csr.num_rows, x.size(1)), dtype=torch.float32, device=x.device)
        csr.matmul(x.contiguous(), y)
        ctx.csr = csr
        return y

    @staticmethod
    def backward(ctx, gy):
        dx = torch.empty((ctx.csr.num_cols, gy.size(1)), dtype=torch.float32, device=gy.device)
        ctx.csr.matmul_t(gy.contiguous(), dx)
        return dx, None


def csr_apply(csr, x):
    return CsrProduct.apply(x, csr)


class GCN(torch.nn.Module):
    """``adjacency`` is a ``cbm4dad`` (CBM kernels) or ``csr4dad`` (rocSPARSE)."""

    def __init__(self, adjacency, in_dim, hidden, out_dim, *, seed=0, device="cuda"):
        super().__init__()
        gen = torch.Generator().manual_seed(seed)
        self.w1 = torch.nn.Parameter((torch.randn(in_dim, hidden, generator=gen) / in_dim ** 0.5).to(device))
        self.w2 = torch.nn.Parameter((torch.randn(hidden, out_dim, generator=gen) / hidden ** 0.5).to(device))
        self.adjacency = adjacency

    def propagate(self, h):
        ext = self.adjacency.cusparse_extension
        if isinstance(ext, cusparseCBM):
            return cbm_apply(ext, h)
        return csr_apply(ext, h)

    def forward(self, x):
        h = torch.relu(self.propagate(x @ self.w1))
        return self.propagate(h @ self.w2)
