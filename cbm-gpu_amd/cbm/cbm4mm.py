"""``A @ X`` in CBM form: drop-in for cbm/cbm4mm.py:7-54."""
import torch

from . import cbm_cusparse as cusparse
from . import cbm_mkl_cpp as cbm_mkl
from ._hip_sigs import MODE_AX
from .autograd import cbm_apply


def _num_rows(edge_index, num_nodes):
    if num_nodes is not None:
        return int(num_nodes)
    # N = max index + 1 (cbm/cbm4mm.py:12): trailing isolated nodes are dropped.
    return int(max(int(edge_index[0].max()), int(edge_index[1].max()))) + 1


class cbm4mm:
    """CBM form of the binary matrix given by ``edge_index`` (``[2, nnz]``),
    ``edge_values`` (all ones) and the pruning parameter ``alpha``.

    Attributes as in the reference: ``deltas`` (CUDA sparse CSR, int32/float32),
    ``num_nodes``, ``mca_offset``, ``mca_src_idx``, ``mca_dst_idx`` and
    ``cusparse_extension``.  ``matmul(x, y)`` writes ``y = A x`` in place.
    Additions: ``x @``-style autograd product via :meth:`apply`, BF16 ``x``,
    ``num_nodes=``, the candidate-edge scan on the GPU (``candidates_on``) and
    the build statistics in ``build_stats``.
    """

    def __init__(self, edge_index, edge_values, alpha=0, *, num_nodes=None, device="cuda", n_threads=0,
                 candidates_on="device"):
        num_rows = _num_rows(edge_index, num_nodes)
        # The candidate-edge scan of the build runs on the target GPU
        # ("device"), or on the host (None); the structure is the same.
        if candidates_on == "device":
            candidates_on = device if torch.device(device).type == "cuda" else None
        built = cbm_mkl.build(edge_index[0], edge_index[1], edge_values, num_rows, alpha, n_threads=n_threads,
                              candidates_on=candidates_on)
        self.build_stats = built.stats
        # Delta as CSR sorted by (row, col): what coalesce().to_sparse_csr()
        # gives in the reference (cbm/cbm4mm.py:26-42), built directly.
        self.deltas = torch.sparse_csr_tensor(built.csr_ptr, built.csr_col, built.csr_val,
                                              (num_rows, num_rows)).to(device=device)
        self.num_nodes = num_rows
        self.mca_offset = built.branches.to(dtype=torch.int32, device=device)
        self.mca_src_idx = built.tree_src.to(dtype=torch.int32, device=device)
        self.mca_dst_idx = built.tree_dst.to(dtype=torch.int32, device=device)
        dummy = torch.empty(0, device=device)
        self.cusparse_extension = cusparse.cusparseCBM(self.deltas, self.mca_offset, self.mca_src_idx,
                                                       self.mca_dst_idx, dummy, mode=MODE_AX, n_threads=n_threads)

    def matmul(self, x, y):
        self.cusparse_extension.matmul_AX_ADX(x, y)

    def apply(self, x):
        """``M @ x`` as a new FP32 tensor, differentiable w.r.t. ``x``."""
        return cbm_apply(self.cusparse_extension, x)

    __matmul__ = apply

    def matmul_t(self, g, out):
        """``out = M^T g`` (the adjoint)."""
        self.cusparse_extension.matmul_t(g, out)

    def _degree_scale(self):
        """D^-1/2 from the row degrees d = A 1, computed with the CBM product
        itself at F = 1 as in cbm/cbm4ad.py:12-19 (exact integers in FP32);
        the power is taken in FP64 and rounded once to FP32.  Degree-0 rows get
        0 instead of the reference's inf, so fused scalings never form 0 * inf."""
        n = self.num_nodes
        dev = self.deltas.device
        d = torch.zeros(n, 1, dtype=torch.float32, device=dev)
        x = torch.ones(n, 1, dtype=torch.float32, device=dev)
        cbm4mm.matmul(self, x, d)
        d = d.view(-1).double()
        return torch.where(d > 0, d.clamp_min(1.0).pow(-0.5), torch.zeros_like(d)).float()
