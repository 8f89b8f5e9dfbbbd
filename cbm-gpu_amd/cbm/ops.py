"""The eight CPU operators of the reference extension ``cbm_mkl_cpp``
(cbm/cbm_extensions.cpp:351-907, registered at :909-919), with the same
names and argument order, running on device tensors with the HIP kernels.

A sparse operand is the MKL 4-array CSR of the reference: ``row_ptr_b``,
``row_ptr_e``, ``col_idx`` (int32) and ``values`` (float32).  ``dst`` is the
output argument, written in place.  The reference's ops run on host tensors
with MKL and OpenMP; these require CUDA tensors and raise otherwise (there
is no host path).  As in the reference, dtypes are checked (``CHECK_DTYPE``,
cbm/cbm_extensions.cpp:14-16) and a mismatch raises ``RuntimeError``.

Differences, all deliberate: the sparse operand's column count is N (the
reference passes ``rhs.size(1)`` in four ops, :374,464,518,618), and the
propagation launches exactly one group per branch.
"""
import ctypes

import torch

from . import _lib
from .cbm_cusparse import current_stream_handle


def _check(t, dtype, name):
    if not isinstance(t, torch.Tensor) or t.dtype != dtype:
        raise RuntimeError(f'"{name}" is not a tensor of type "{dtype}"')
    if not t.is_cuda:
        raise RuntimeError(f'"{name}" is not a CUDA tensor: the HIP operators have no host path')
    return t


def _dense(t, name, dtype=torch.float32):
    _check(t, dtype, name)
    if t.dim() != 2 or t.stride(1) != 1:
        raise RuntimeError(f'"{name}" must be a 2-D tensor with unit column stride')
    return t


def _ok(rc, what):
    if rc != _lib.CBM_OK:
        msg = _lib.hip().cbm_op_last_error().decode(errors="replace")
        _lib._raise(rc, msg, what)


def _csr_args(row_b, row_e, col, val):
    _check(row_b, torch.int32, "lhs_row_ptr_b")
    _check(row_e, torch.int32, "lhs_row_ptr_e")
    _check(col, torch.int32, "lhs_col_idx")
    _check(val, torch.float32, "lhs_values")
    return row_b.contiguous(), row_e.contiguous(), col.contiguous(), val.contiguous()


def _spmm(row_b, row_e, col, val, rhs, dst, self_coef=None):
    row_b, row_e, col, val = _csr_args(row_b, row_e, col, val)
    _dense(rhs, "rhs")
    _dense(dst, "dst")
    n_rows = int(dst.size(0))
    f = int(rhs.size(1))
    if row_b.numel() < n_rows or row_e.numel() < n_rows or int(dst.size(1)) != f:
        raise ValueError("row pointer arrays or dst do not match the operand shapes")
    rc = _lib.hip().cbm_op_spmm(n_rows, row_b.data_ptr(), row_e.data_ptr(), col.data_ptr() or None,
                                val.data_ptr() or None, rhs.data_ptr(), rhs.stride(0), dst.data_ptr(),
                                dst.stride(0), f, float(self_coef or 0.0), int(self_coef is not None),
                                current_stream_handle(dst.device))
    _ok(rc, "spmm")


def _update(mca_branches, mca_row_idx, mca_col_idx, dst, multiplier=None):
    _check(mca_branches, torch.int32, "mca_branches")
    _check(mca_row_idx, torch.int32, "mca_row_idx")
    _check(mca_col_idx, torch.int32, "mca_col_idx")
    _dense(dst, "dst")
    if multiplier is not None:
        _check(multiplier, torch.float32, "multiplier")
        multiplier = multiplier.contiguous()
    br = mca_branches.contiguous()
    rc = _lib.hip().cbm_op_update(max(0, int(br.numel()) - 1), br.data_ptr(), mca_row_idx.contiguous().data_ptr(),
                                  mca_col_idx.contiguous().data_ptr(),
                                  multiplier.data_ptr() if multiplier is not None else None, dst.data_ptr(),
                                  dst.stride(0), int(dst.size(1)), current_stream_handle(dst.device))
    _ok(rc, "update")


def _axpy(coef, x, dst):
    rc = _lib.hip().cbm_op_axpy(int(dst.size(0)), int(dst.size(1)), float(coef), x.data_ptr(), x.stride(0),
                                dst.data_ptr(), dst.stride(0), current_stream_handle(dst.device))
    _ok(rc, "axpy")


def s_spmm_update_csr_int32(lhs_row_ptr_b, lhs_row_ptr_e, lhs_col_idx, lhs_values, rhs, mca_branches, mca_row_idx,
                            mca_col_idx, dst):
    """dst = Delta rhs, then the tree propagation (cbm/cbm_extensions.cpp:351-444)."""
    _spmm(lhs_row_ptr_b, lhs_row_ptr_e, lhs_col_idx, lhs_values, rhs, dst)
    _update(mca_branches, mca_row_idx, mca_col_idx, dst)


def s_spmm_fused_update_csr_int32(lhs_row_ptr_b, lhs_row_ptr_e, lhs_col_idx, lhs_values, rhs, mca_branches,
                                  mca_row_idx, mca_col_idx, multiplier, dst):
    """dst = Delta' rhs, then the scaled propagation y_dst += y_src m_dst / m_src (:593-698)."""
    _spmm(lhs_row_ptr_b, lhs_row_ptr_e, lhs_col_idx, lhs_values, rhs, dst)
    _update(mca_branches, mca_row_idx, mca_col_idx, dst, multiplier)


def s_spmm_add_csr_int32(lhs_row_ptr_b, lhs_row_ptr_e, lhs_col_idx, lhs_values, rhs, epsilon, dst):
    """dst = A rhs + (1 + epsilon) rhs, fused in one pass (:446-492)."""
    _spmm(lhs_row_ptr_b, lhs_row_ptr_e, lhs_col_idx, lhs_values, rhs, dst, self_coef=1.0 + float(epsilon))


def s_spmm_update_add_csr_int32(lhs_row_ptr_b, lhs_row_ptr_e, lhs_col_idx, lhs_values, rhs, epsilon, mca_branches,
                                mca_row_idx, mca_col_idx, dst):
    """dst = (I - P)^-1 Delta rhs + (1 + epsilon) rhs (:494-591)."""
    _spmm(lhs_row_ptr_b, lhs_row_ptr_e, lhs_col_idx, lhs_values, rhs, dst)
    _update(mca_branches, mca_row_idx, mca_col_idx, dst)
    _axpy(1.0 + float(epsilon), rhs, dst)


def s_spmm_csr_int32(lhs_row_ptr_b, lhs_row_ptr_e, lhs_col_idx, lhs_values, rhs, dst):
    """dst = A rhs (:700-740)."""
    _spmm(lhs_row_ptr_b, lhs_row_ptr_e, lhs_col_idx, lhs_values, rhs, dst)


def s_spmm_csr_t_int32(lhs_row_ptr_b, lhs_row_ptr_e, lhs_col_idx, lhs_values, rhs, dst):
    """dst = A^T rhs (:742-782), with A given by rows; A has dst.size(0)
    columns and rhs.size(0) rows.  A^T is sorted out on the device (stable
    radix sort by column) in a workspace from torch's allocator, then summed
    by the row-pass kernel (cbm_op_spmm_t)."""
    row_b, row_e, col, val = _csr_args(lhs_row_ptr_b, lhs_row_ptr_e, lhs_col_idx, lhs_values)
    _dense(rhs, "rhs")
    _dense(dst, "dst")
    n_rows, n_cols, f = int(rhs.size(0)), int(dst.size(0)), int(rhs.size(1))
    if int(dst.size(1)) != f or row_b.numel() < n_rows or row_e.numel() < n_rows:
        raise ValueError("row pointer arrays or dst do not match the operand shapes")
    nnz = int((row_e[:n_rows].long() - row_b[:n_rows].long()).clamp_min(0).sum()) if n_rows else 0
    lib = _lib.hip()
    ws_bytes = int(lib.cbm_op_spmm_t_workspace_bytes(n_rows, n_cols, nnz))
    if ws_bytes < 0:
        raise ValueError(f"no transpose workspace for {n_rows} x {n_cols} with {nnz} entries")
    ws = torch.empty(ws_bytes + 256, dtype=torch.uint8, device=dst.device)
    ws_ptr = (ws.data_ptr() + 255) // 256 * 256
    rc = lib.cbm_op_spmm_t(n_rows, n_cols, row_b.data_ptr() or None, row_e.data_ptr() or None,
                           col.data_ptr() or None, val.data_ptr() or None, nnz, rhs.data_ptr(), rhs.stride(0),
                           dst.data_ptr(), dst.stride(0), f, ws_ptr, ws_bytes, current_stream_handle(dst.device))
    _ok(rc, "spmm_t")


def s_update_csr_int32(mca_branches, mca_row_idx, mca_col_idx, dst):
    """In place: for each branch, in order, dst[col] += dst[row] (:784-838)."""
    _update(mca_branches, mca_row_idx, mca_col_idx, dst)


def s_fused_update_csr_int32(mca_branches, mca_row_idx, mca_col_idx, multiplier, dst):
    """In place: dst[col] += dst[row] * m[col] / m[row] branch by branch (:841-907)."""
    _update(mca_branches, mca_row_idx, mca_col_idx, dst, multiplier)


OPS = ("s_spmm_update_csr_int32", "s_spmm_fused_update_csr_int32", "s_spmm_add_csr_int32",
       "s_spmm_update_add_csr_int32", "s_spmm_csr_int32", "s_spmm_csr_t_int32", "s_update_csr_int32",
       "s_fused_update_csr_int32")
