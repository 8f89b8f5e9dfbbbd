"""MI355X counterpart of the reference extension module ``cbm_mkl_cpp``
(cbm/cbm_extensions.cpp:909-919).

``init`` keeps the reference signature and return value.  It runs on the host
(libcbm_host.so): MKL's syrk and the arbok arborescence are replaced by a
streaming overlap scan and our own minimum-arborescence code, with the
candidate edges created in the reference order.  Optionally the candidate
scan runs on the GPU (HIP kernels of cand_api.hip), which gives the same list.
"""
import torch

from . import structure


def _as_i32(t):
    t = torch.as_tensor(t)
    if t.dtype not in (torch.int32, torch.int64, torch.int16, torch.uint8, torch.int8):
        raise TypeError(f'index tensor must be integral, got "{t.dtype}"')
    return t.detach().to("cpu", torch.int32).contiguous()


def build(src_row_idx, src_col_idx, src_values, src_n_rows, alpha, n_threads=0, candidates_on=None):
    """Full build: the ``init`` arrays plus the parent array, Delta as CSR
    and statistics (see :class:`cbm.structure.Build`).  ``candidates_on``: a
    CUDA device runs the candidate-edge scan there (same result)."""
    return structure.build(_as_i32(src_row_idx), _as_i32(src_col_idx),
                           torch.as_tensor(src_values).detach().to("cpu", torch.float32).contiguous(),
                           int(src_n_rows), int(alpha), n_threads=n_threads, candidates_on=candidates_on)


def init(src_row_idx, src_col_idx, src_values, src_n_rows, alpha, *, device=None):
    """``cbm_mkl_cpp.init(row, col, val, n_rows, alpha)`` (cbm/cbm_extensions.cpp:60-345).

    Returns ``[delta_row, delta_col, delta_val, mca_branches, mca_row_idx,
    mca_col_idx]`` as CPU tensors (int32, int32, float32, int32, int32, int32):
    Delta in COO construction order, the prefix offsets of the non-trivial
    branches, and the tree edges (parent, child) in BFS order per branch.
    Like the reference, the index and value tensors must be int32 / float32
    (``CHECK_DTYPE``, :65-67).  ``device`` (an addition): a CUDA device on
    which the candidate edges are scanned; the arrays are the same.
    """
    for t, dt, name in ((src_row_idx, torch.int32, "src_row_idx"), (src_col_idx, torch.int32, "src_col_idx"),
                        (src_values, torch.float32, "src_values")):
        if not isinstance(t, torch.Tensor) or t.dtype != dt:
            raise RuntimeError(f'"{name}" is not a tensor of type "{dt}"')
    return structure.build(src_row_idx, src_col_idx, src_values, int(src_n_rows), int(alpha),
                           candidates_on=device).init_tuple()


# The eight operators registered next to init (cbm/cbm_extensions.cpp:911-918).
from .ops import (  # noqa: E402,F401
    s_fused_update_csr_int32,
    s_spmm_add_csr_int32,
    s_spmm_csr_int32,
    s_spmm_csr_t_int32,
    s_spmm_fused_update_csr_int32,
    s_spmm_update_add_csr_int32,
    s_spmm_update_csr_int32,
    s_update_csr_int32,
)
