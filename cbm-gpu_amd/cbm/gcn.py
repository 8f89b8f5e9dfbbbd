"""GNN layers whose sparse products run on the CBM kernels (forward) and the
hand-written adjoint (backward), and a two-layer GCN built from them.  The
reference's harness imports ``GCNConv`` and ``GINConv`` but never builds a
model (benchmark/utilities.py:11); its operator surface holds the GIN-style
product ``A x + (1 + eps) x`` (s_spmm_add_csr_int32 and
s_spmm_update_add_csr_int32, cbm/cbm_extensions.cpp:446-591).

* :class:`GCNConv`  ``out = Â (x W) + b`` with Â the adjacency object's product
  (``cbm4dad``: D^-1/2 A D^-1/2), weight first so the product runs at the
  output width.
* :class:`GINConv`  ``out = mlp(A x + (1 + eps) x)`` with A the adjacency
  object's product (``cbm4mm``); eps fixed or trained.
* :class:`GCN`      two GCNConv layers, ``out = Â relu(Â (X W1)) W2``.
"""
import torch

from . import _lib
from .autograd import cbm_apply
from .cbm_cusparse import current_stream_handle, cusparseCBM


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


def propagate(adjacency, h):
    """``M h`` with autograd, on the CBM kernels (``cbm4*``) or rocSPARSE (``csr4*``)."""
    ext = adjacency.cusparse_extension
    if isinstance(ext, cusparseCBM):
        return cbm_apply(ext, h)
    return csr_apply(ext, h)


# The weight gradient x^T g of a layer sums over the N rows; a single product
# gives a 256 x 256 output a handful of workgroups, so for tall inputs the
# rows are cut into WGRAD_SPLITS chunks whose products run as one batched
# product and are then added (N = 169,343, 128 x 256 output: 1.03 -> 0.28 ms;
# 256 x 256: 1.04 -> 0.49 ms; the sum is also closer to FP64, 5e-7 against
# 3e-6 relative).  Inputs with fewer rows than WGRAD_SPLITS * 256 use one product.
WGRAD_SPLITS = 64


def weight_grad(x, g):
    """``x^T g`` for tall ``x`` (N x k) and ``g`` (N x m), FP32."""
    n = x.size(0)
    c = n // WGRAD_SPLITS
    if c < 256:
        return x.t() @ g
    m = c * WGRAD_SPLITS
    xt = x[:m].reshape(WGRAD_SPLITS, c, x.size(1)).transpose(1, 2).contiguous()
    gb = g[:m].reshape(WGRAD_SPLITS, c, g.size(1)).contiguous()
    out = torch.bmm(xt, gb).sum(0)
    if m < n:
        out = out + x[m:].t() @ g[m:]
    return out


class Linear(torch.autograd.Function):
    """``y = x W`` with the weight gradient of :func:`weight_grad`."""

    @staticmethod
    def forward(ctx, x, w):
        ctx.save_for_backward(x, w)
        return x @ w

    @staticmethod
    def backward(ctx, gy):
        x, w = ctx.saved_tensors
        gy = gy.contiguous()
        dx = gy @ w.t() if ctx.needs_input_grad[0] else None
        dw = weight_grad(x, gy) if ctx.needs_input_grad[1] else None
        return dx, dw


def linear(x, w):
    return Linear.apply(x.contiguous(), w)


def _axpy(coef, x, y):
    """y += coef * x (FP32, the library's row kernel)."""
    rc = _lib.hip().cbm_op_axpy(int(y.size(0)), int(y.size(1)), float(coef), x.data_ptr(), x.stride(0),
                                y.data_ptr(), y.stride(0), current_stream_handle(y.device))
    if rc != _lib.CBM_OK:
        _lib._raise(rc, _lib.hip().cbm_op_last_error().decode(errors="replace"), "axpy")


class CbmAddProduct(torch.autograd.Function):
    """``y = M x + (1 + eps) x`` (the GIN aggregation) and its adjoint
    ``dx = M^T g + (1 + eps) g``, ``d eps = <x, g>``.  Computed in FP32 (a
    BF16 ``x`` is widened; its gradient is returned in BF16)."""

    @staticmethod
    def forward(ctx, x, eps, handle):
        ctx.in_dtype = x.dtype
        x = x.contiguous().float()
        y = torch.empty((handle.num_rows, x.size(1)), dtype=torch.float32, device=x.device)
        handle.forward(x, y)
        coef = 1.0 + float(eps)
        _axpy(coef, x, y)
        ctx.handle = handle
        ctx.coef = coef
        ctx.save_for_backward(x, eps)
        return y

    @staticmethod
    def backward(ctx, gy):
        x, eps = ctx.saved_tensors
        g = gy.contiguous().float()
        dx = torch.empty((ctx.handle.num_rows, g.size(1)), dtype=torch.float32, device=g.device)
        ctx.handle.matmul_t(g, dx)
        _axpy(ctx.coef, g, dx)
        deps = (x * g).sum().reshape(eps.shape).to(eps.dtype) if ctx.needs_input_grad[1] else None
        return dx.to(ctx.in_dtype), deps, None


class GCNConv(torch.nn.Module):
    """``out = Â (x W) + b``; ``adjacency`` is a ``cbm4dad`` (CBM kernels) or a
    ``csr4dad`` (rocSPARSE), or any ``cbm4*`` / ``csr4*`` for another Â."""

    def __init__(self, adjacency, in_dim, out_dim, *, bias=True, seed=0, device="cuda"):
        super().__init__()
        gen = torch.Generator().manual_seed(seed)
        self.weight = torch.nn.Parameter((torch.randn(in_dim, out_dim, generator=gen) / in_dim ** 0.5).to(device))
        self.bias = torch.nn.Parameter(torch.zeros(out_dim, device=device)) if bias else None
        self.adjacency = adjacency

    def forward(self, x):
        out = propagate(self.adjacency, linear(x, self.weight))
        return out + self.bias if self.bias is not None else out


class GINConv(torch.nn.Module):
    """``out = mlp(A x + (1 + eps) x)``; ``adjacency`` is a ``cbm4mm`` (or any
    ``cbm4*`` for another product).  ``mlp`` defaults to the identity."""

    def __init__(self, adjacency, mlp=None, *, eps=0.0, train_eps=False, device="cuda"):
        super().__init__()
        ext = adjacency.cusparse_extension
        if not isinstance(ext, cusparseCBM):
            raise TypeError("GINConv runs on the CBM kernels: pass a cbm4mm / cbm4ad / cbm4dad adjacency")
        self.adjacency = adjacency
        self.mlp = mlp if mlp is not None else torch.nn.Identity()
        e = torch.tensor([float(eps)], device=device)
        self.eps = torch.nn.Parameter(e) if train_eps else e

    def forward(self, x):
        return self.mlp(CbmAddProduct.apply(x, self.eps, self.adjacency.cusparse_extension))


class GCN(torch.nn.Module):
    """Two GCNConv layers without bias: ``Â relu(Â (X W1)) W2``, the second
    product at the output width.  ``adjacency`` is a ``cbm4dad`` (CBM kernels)
    or ``csr4dad`` (rocSPARSE)."""

    def __init__(self, adjacency, in_dim, hidden, out_dim, *, seed=0, device="cuda"):
        super().__init__()
        gen = torch.Generator().manual_seed(seed)
        self.w1 = torch.nn.Parameter((torch.randn(in_dim, hidden, generator=gen) / in_dim ** 0.5).to(device))
        self.w2 = torch.nn.Parameter((torch.randn(hidden, out_dim, generator=gen) / hidden ** 0.5).to(device))
        self.adjacency = adjacency

    def propagate(self, h):
        return propagate(self.adjacency, h)

    def forward(self, x):
        h = torch.relu(self.propagate(linear(x, self.w1)))
        return self.propagate(linear(h, self.w2))
