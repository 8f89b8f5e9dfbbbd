"""The allocation-free forward entry point, cbm_matmul_noalloc, called
through the C ABI: it equals cbm_matmul bit for bit (same kernels and order;
the XCD ranges are off at these sizes, and the one-launch forward, which
needs the handle's flags, is off) on two calls in a row, the second
with a smaller input, on graphs whose product uses the static scratch: the
hub graph (long rows: chunk partial sums) and a chain of depth 80
(compensated levels: compensation rows).  It also matches the FP64 oracle,
and refuses a width whose scratch would not fit."""
import ctypes

import numpy as np
import pytest
import torch

from cbm import _lib
from cbm.cbm4dad import cbm4dad
from cbm.cbm4mm import cbm4mm
from hubgraph import HUB_N, hub_graph
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CHAIN = 80


def _chain_graph(m=CHAIN):
    """Rows i < m nested (row i+1 = row i minus one column): a chain of depth m."""
    pairs = set()
    for i in range(m):
        for j in range(m - i):
            pairs.add((i, m + j))
            pairs.add((m + j, i))
    return torch.tensor(sorted(pairs), dtype=torch.int32).t().contiguous(), 2 * m


def _call(lib, fn, h, x, y):
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    code = 1 if x.dtype == torch.bfloat16 else 0
    return getattr(lib, fn)(h._h, code, x.data_ptr(), x.stride(0), y.data_ptr(), y.stride(0), int(x.size(1)), st)


def _check_pair(lib, h, a, mode, x):
    """cbm_matmul_noalloc vs cbm_matmul (bitwise) and vs FP64."""
    n, f = x.shape
    h.reserve(f)
    y_ref = torch.full((n, f), float("nan"), device=x.device)
    y = torch.full((n, f), float("nan"), device=x.device)
    assert _call(lib, "cbm_matmul", h, x, y_ref) == 0, lib.cbm_last_error()
    rc = _call(lib, "cbm_matmul_noalloc", h, x, y)
    assert rc == 0, lib.cbm_last_error()
    torch.cuda.synchronize()
    assert torch.equal(y, y_ref)
    assert O.max_rel_error(y.cpu().numpy(), O.forward(a, x.float().cpu().numpy(), mode)) <= 1e-5


def _rand(n, f, seed):
    return np.random.default_rng(seed).standard_normal((n, f)).astype(np.float32)


def test_noalloc_equals_cbm_matmul_on_long_rows(cuda):
    ei = hub_graph()
    a = O.csr_of(ei.numpy(), HUB_N)
    m = cbm4dad(ei, torch.ones(ei.shape[1]), alpha=0, device=cuda)
    h = m.cusparse_extension
    h.set_fused(False)  # cbm_matmul's multi-launch form, which the static scratch serves
    assert h.plan_stats["fwd_gslots"] > 0 and h.plan_stats["fwd_slots"] > 0  # long rows use the scratch
    lib = _lib.hip()
    for f, seed in ((64, 1), (16, 2)):  # the second call smaller
        x = torch.from_numpy(_rand(HUB_N, f, seed)).to(cuda)
        _check_pair(lib, h, a, "dadx", x)
    h.set_forward_split(False)  # the fused form: per-level slots
    _check_pair(lib, h, a, "dadx", torch.from_numpy(_rand(HUB_N, 32, 3)).to(cuda))
    h.set_forward_split(None)
    _check_pair(lib, h, a, "dadx", torch.from_numpy(_rand(HUB_N, 128, 4)).to(cuda).to(torch.bfloat16))


def test_noalloc_equals_cbm_matmul_on_compensated_levels(cuda):
    ei, n = _chain_graph()
    a = O.csr_of(ei.numpy(), n)
    m = cbm4mm(ei, torch.ones(ei.shape[1]), alpha=0, device=cuda)
    h = m.cusparse_extension
    assert h.depth >= CHAIN > 32  # levels from 32 on carry compensation rows
    lib = _lib.hip()
    for f, seed in ((40, 5), (8, 6)):
        _check_pair(lib, h, a, "ax", torch.from_numpy(_rand(n, f, seed)).to(cuda))


def test_noalloc_refuses_what_does_not_fit(cuda):
    ei, n = _chain_graph()
    m = cbm4mm(ei, torch.ones(ei.shape[1]), alpha=0, device=cuda)
    h = m.cusparse_extension
    h.set_kahan_depth(0)  # every level past the first carries a compensation row
    lib = _lib.hip()
    # (n - first compensated position) x f floats > 64 MiB
    f = (64 << 20) // 4 // (n - h.level_sizes[0]) + 8
    f = (f + 3) // 4 * 4
    x = torch.zeros(n, f, device=cuda)
    y = torch.zeros(n, f, device=cuda)
    assert _call(lib, "cbm_matmul_noalloc", h, x, y) == _lib.CBM_ERR_ARG
    assert b"static scratch" in lib.cbm_last_error()
