"""``A D^-1/2 @ X`` in CBM form: drop-in for cbm/cbm4ad.py:5-42."""
import torch

from .cbm4mm import cbm4mm
from ._hip_sigs import MODE_ADX


class cbm4ad(cbm4mm):
    """Delta's columns are scaled by D^-1/2 (cbm/cbm4ad.py:21-36); the GPU
    kernels apply the scale to each gathered row of X instead of storing
    scaled values.  ``matmul(x, y)`` writes ``y = A D^-1/2 x``."""

    def _scale(self):
        self.D = self._degree_scale()
        column_indices = self.deltas.col_indices()
        new_values = self.deltas.values() * self.D[column_indices]
        self.deltas = torch.sparse_csr_tensor(self.deltas.crow_indices(), column_indices,
                                              new_values.to(torch.float32), self.deltas.shape)
        self.cusparse_extension = self.cusparse_extension.with_scaling(MODE_ADX, self.D)
