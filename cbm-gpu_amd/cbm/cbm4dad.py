"""``D^-1/2 A D^-1/2 @ X`` in CBM form: drop-in for cbm/cbm4dad.py:5-54."""
import torch

from .cbm4mm import cbm4mm
from ._hip_sigs import MODE_DADX


class cbm4dad(cbm4mm):
    """Every Delta row and column is scaled by D^-1/2 and the propagation
    uses the per-edge factor D_v / D_parent(v) (cbm/cbm4dad.py:21-51,
    cbm/cbm.cu:88).  The reference's per-row Python loop (:28-34) ends up
    scaling every row; this builds the same values in one vectorised step.
    ``matmul(x, y)`` writes ``y = D^-1/2 A D^-1/2 x``."""

    def _scale(self):
        self.D = self._degree_scale()
        crow = self.deltas.crow_indices()
        column_indices = self.deltas.col_indices()
        rows = torch.repeat_interleave(torch.arange(self.num_nodes, device=crow.device), crow.diff().long())
        new_values = self.deltas.values() * self.D[column_indices] * self.D[rows]
        self.deltas = torch.sparse_csr_tensor(crow, column_indices, new_values.to(torch.float32), self.deltas.shape)
        self.cusparse_extension = self.cusparse_extension.with_scaling(MODE_DADX, self.D)

    def matmul(self, x, y):
        self.cusparse_extension.matmul_DAX_DADX(x, y)
