"""``D^-1/2 A D^-1/2 @ X`` in CSR form on the GPU: drop-in for cbm/csr4dad.py:5-36."""
import torch

from . import csr_cusparse as cusparse
from .csr4mm import csr4mm


class csr4dad(csr4mm):
    def __init__(self, edge_index, **kwargs):
        super().__init__(edge_index, **kwargs)
        self.D = self._degree_scale()
        n = self.a.size(0)
        rows = torch.repeat_interleave(torch.arange(n, device=self.a.device), self.a.crow_indices().diff().long())
        new_values = self.a.values() * self.D[rows] * self.D[self.a.col_indices()]
        self.a = torch.sparse_csr_tensor(self.a.crow_indices(), self.a.col_indices(), new_values.to(torch.float32),
                                         self.a.shape)
        self.cusparse_extension = cusparse.cusparseCSR(self.a)
