"""Compensated propagation on a deep tree.  Rows i < M of the matrix are
nested (row i+1 = row i minus one column), so the arborescence is a chain of
depth M and the forward product is a running sum along it: Y_i = x + Y_(i+1).
With x = 0.1 (not representable) plain FP32 propagation drifts by about
M * u relative; compensated levels stay at the FP32 rounding of the result."""
import numpy as np
import pytest
import torch

from cbm.cbm4mm import cbm4mm
from oracle import oracle as O

pytestmark = pytest.mark.gpu

M = 1000


@pytest.fixture(scope="module")
def chain(cuda):
    rows, cols = [], []
    for i in range(M):
        for j in range(M - i):
            rows += [i, M + j]
            cols += [M + j, i]
    ei = torch.tensor(sorted(set(zip(rows, cols))), dtype=torch.int32).t().contiguous()
    m = cbm4mm(ei, torch.ones(ei.shape[1]), alpha=0, device=cuda)
    assert m.cusparse_extension.depth >= M
    return O.csr_of(ei.numpy(), 2 * M), m


def test_compensation_removes_depth_drift(chain, cuda):
    a, m = chain
    h = m.cusparse_extension
    x = np.full((2 * M, 8), 0.1, np.float32)
    x[:, 1] = 1.0 / 3.0
    x[:, 2] = np.float32(0.7)
    ref = O.forward(a, x, "ax")
    y = torch.empty(2 * M, 8, device=cuda)
    h.set_kahan_depth(None)
    h.forward(torch.from_numpy(x).to(cuda), y)
    err_plain = O.max_rel_error(y.cpu().numpy(), ref)
    h.set_kahan_depth(32)
    h.forward(torch.from_numpy(x).to(cuda), y)
    err_comp = O.max_rel_error(y.cpu().numpy(), ref)
    h.set_kahan_depth(0)
    h.forward(torch.from_numpy(x).to(cuda), y)
    err_all = O.max_rel_error(y.cpu().numpy(), ref)
    assert err_plain > 2e-6, err_plain  # the drift is real at this depth
    assert err_comp < 2e-7 and err_all < 2e-7, (err_comp, err_all)
    assert err_comp < err_plain / 10


@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16])
def test_compensated_random_inputs_and_bf16(chain, cuda, dtype):
    a, m = chain
    h = m.cusparse_extension
    h.set_kahan_depth(32)
    xt = torch.from_numpy(np.random.default_rng(1).standard_normal((2 * M, 16)).astype(np.float32)).to(cuda)
    xt = xt.to(dtype)
    y = torch.empty(2 * M, 16, device=cuda)
    h.forward(xt, y)
    assert O.max_rel_error(y.cpu().numpy(), O.forward(a, xt.float().cpu().numpy(), "ax")) <= 1e-6
