"""MI355X counterpart of the reference extension module ``csr_cusparse``
(cbm/csr.cu:128-133): class :class:`cusparseCSR` over rocSPARSE SpMM.  It is
the CSR comparator on the same GPU and the CSR reference output of the parity
tests (benchmark/compare_matmul_cuda.py:50)."""
import ctypes

import torch

from . import _lib
from .cbm_cusparse import _require_cuda, current_stream_handle

# rocsparse_spmm_alg values (rocsparse-types.h).
ALG_DEFAULT, ALG_CSR, ALG_CSR_ROW_SPLIT, ALG_CSR_MERGE, ALG_CSR_MERGE_PATH = 0, 1, 4, 5, 9


class cusparseCSR:
    """``cusparseCSR(adjacency)``: adjacency is a CUDA sparse CSR tensor with
    int32 indices and float32 values (cbm/csr.cu:26-48).  The tensor is kept
    alive by the handle (the reference keeps only a dangling pointer,
    cbm/csr.cu:29)."""

    def __init__(self, adjacency, algorithm=ALG_DEFAULT):
        _require_cuda(adjacency)
        if adjacency.layout != torch.sparse_csr:
            raise TypeError("adjacency must be a sparse CSR tensor")
        self.a = adjacency
        self._crow = adjacency.crow_indices()
        self._col = adjacency.col_indices()
        self._val = adjacency.values()
        for t, dt, name in ((self._crow, torch.int32, "crow_indices"), (self._col, torch.int32, "col_indices"),
                            (self._val, torch.float32, "values")):
            if t.dtype != dt:
                raise TypeError(f'"{name}" is not a tensor of type "{dt}"')
        self.num_rows, self.num_cols = int(adjacency.size(0)), int(adjacency.size(1))
        self.device = adjacency.device
        self._lib = _lib.hip()
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            rc = self._lib.csr_create(self.num_rows, self.num_cols, int(self._col.numel()), self._crow.data_ptr(),
                                      self._col.data_ptr() or None, self._val.data_ptr() or None, ctypes.byref(h))
        self._check(rc, "cusparseCSR")
        self._h = h
        self._alg = algorithm
        self._t = None
        self.set_algorithm(algorithm)

    def _check(self, rc, what):
        if rc != _lib.CBM_OK:
            msg = self._lib.csr_last_error().decode(errors="replace")
            _lib._raise(rc, msg, what)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.csr_destroy(h)
            self._h = None

    def set_algorithm(self, algorithm):
        self._alg = algorithm
        self._check(self._lib.csr_set_algorithm(self._h, int(algorithm)), "set_algorithm")
        if getattr(self, "_t", None) is not None:
            self._t.set_algorithm(algorithm)

    def transposed(self):
        """A^T as its own CSR handle, built once on the device (a sort of the
        entries by column); matmul_t runs it as a plain SpMM, which is far
        faster than rocSPARSE's transposed-operand SpMM."""
        if self._t is None:
            dev = self.device
            crow, col = self._crow.long(), self._col.long()
            rows = torch.repeat_interleave(torch.arange(self.num_rows, device=dev), crow.diff())
            order = torch.argsort(col * self.num_rows + rows)
            t_crow = torch.zeros(self.num_cols + 1, dtype=torch.int64, device=dev)
            t_crow[1:] = torch.cumsum(torch.bincount(col, minlength=self.num_cols), 0)
            at = torch.sparse_csr_tensor(t_crow.to(torch.int32), rows[order].to(torch.int32), self._val[order],
                                         (self.num_cols, self.num_rows))
            self._t = cusparseCSR(at, algorithm=self._alg)
        return self._t

    def _run(self, trans, rhs, res):
        _require_cuda(rhs)
        _require_cuda(res)
        if rhs.dtype != torch.float32 or res.dtype != torch.float32:
            raise TypeError("cusparseCSR operands must be float32")
        rows_in, rows_out = (self.num_rows, self.num_cols) if trans else (self.num_cols, self.num_rows)
        if rhs.dim() != 2 or int(rhs.size(0)) != rows_in:
            raise ValueError(f"rhs must be {rows_in} x F, got {tuple(rhs.shape)}")
        f = int(rhs.size(1))
        if tuple(res.shape) != (rows_out, f) or res.stride(1) != 1:
            raise ValueError(f"res must be a row-major {rows_out} x {f} tensor")
        if rhs.stride(1) != 1:
            rhs = rhs.contiguous()
        if f == 0:
            return 0
        rc = self._lib.csr_matmul(self._h, int(trans), rhs.data_ptr(), rhs.stride(0), res.data_ptr(), res.stride(0),
                                  f, current_stream_handle(self.device))
        self._check(rc, "matmul")
        return 0

    def matmul(self, rhs, res):
        """res = A rhs (cbm/csr.cu:56-103)."""
        return self._run(0, rhs, res)

    def matmul_t(self, rhs, res):
        """res = A^T rhs (the comparator of the adjoint), as a plain SpMM over
        the explicitly transposed matrix."""
        return self.transposed().matmul(rhs, res)

    def matmul_t_rocsparse(self, rhs, res):
        """res = A^T rhs with rocSPARSE's transposed-operand SpMM."""
        return self._run(1, rhs, res)
