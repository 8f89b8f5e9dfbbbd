"""``A @ X`` in CSR form on the GPU (rocSPARSE): drop-in for cbm/csr4mm.py:6-39."""
import torch

from . import csr_cusparse as cusparse


def csr_from_edge_index(edge_index, num_nodes, device, values=None):
    """CSR (int32 indices) of the matrix with ones (or ``values``) at ``edge_index``.
    Input that is already sorted by (row, col) without duplicates (what the
    copying-model generator returns) skips the COO coalesce of cbm/csr4mm.py:14-19."""
    ei = edge_index.detach().to("cpu", torch.int64)
    n = int(num_nodes)
    key = ei[0] * n + ei[1]
    if values is None and key.numel() > 0 and bool((key[1:] > key[:-1]).all()):
        crow = torch.zeros(n + 1, dtype=torch.int64)
        crow[1:] = torch.bincount(ei[0], minlength=n).cumsum(0)
        return torch.sparse_csr_tensor(crow.to(torch.int32), ei[1].to(torch.int32),
                                       torch.ones(ei.size(1), dtype=torch.float32), (n, n)).to(device=device)
    vals = torch.ones(ei.size(1), dtype=torch.float32) if values is None else values.detach().to("cpu", torch.float32)
    a = torch.sparse_coo_tensor(ei, vals, (n, n)).coalesce().to_sparse_csr()
    return torch.sparse_csr_tensor(a.crow_indices().to(torch.int32), a.col_indices().to(torch.int32),
                                   a.values().to(torch.float32), a.shape).to(device=device)


class csr4mm:
    """``csr4mm(edge_index)``: ``.a`` is the CUDA CSR adjacency, ``.matmul(x, y)``
    writes ``y = A x`` with rocSPARSE."""

    def __init__(self, edge_index, *, num_nodes=None, device="cuda"):
        # N = max index + 1 (cbm/csr4mm.py:11).
        self.num_nodes = int(num_nodes) if num_nodes is not None else int(edge_index.max()) + 1
        self.a = csr_from_edge_index(edge_index, self.num_nodes, device)
        self.cusparse_extension = cusparse.cusparseCSR(self.a)

    def matmul(self, x, y):
        """y = A x (cbm/csr4mm.py:31-39)."""
        self.cusparse_extension.matmul(x, y)

    def matmul_t(self, x, y):
        """y = A^T x."""
        self.cusparse_extension.matmul_t(x, y)

    def _degree_scale(self):
        """D^-1/2 from d = A 1 computed by the CSR product itself (cbm/csr4ad.py:12-19),
        power in FP64 rounded to FP32, 0 for degree-0 rows."""
        n = self.a.size(0)
        d = torch.zeros(n, 1, dtype=torch.float32, device=self.a.device)
        x = torch.ones(n, 1, dtype=torch.float32, device=self.a.device)
        csr4mm.matmul(self, x, d)
        d = d.view(-1).double()
        return torch.where(d > 0, d.clamp_min(1.0).pow(-0.5), torch.zeros_like(d)).float()
