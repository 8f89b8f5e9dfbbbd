"""GNN layers over the CBM product (cbm/gcn.py): GINConv's aggregation
A x + (1 + eps) x and GCNConv's D^-1/2 A D^-1/2 (x W) + b, forward and every
gradient (input, eps, weight, bias) against FP64 dense products on the host,
and the GIN aggregation against the reference-named operator
s_spmm_update_add_csr_int32 on the same CBM pieces."""
import numpy as np
import pytest
import torch

from cbm import cbm_mkl_cpp, structure
from cbm.cbm4dad import cbm4dad
from cbm.cbm4mm import cbm4mm
from cbm.gcn import GCNConv, GINConv
from oracle import oracle as O

pytestmark = pytest.mark.gpu

N = 2708


@pytest.fixture(scope="module")
def graph(cuda):
    ei = structure.copying_model(N, 7, 0.8, 0.58, 0, 1001)
    a = O.csr_of(ei.numpy(), N)
    return ei, a


def _rand(shape, seed):
    return np.random.default_rng(seed).standard_normal(shape).astype(np.float32)


@pytest.mark.parametrize("alpha", [0, 3])
def test_gin_aggregation_and_gradients(graph, cuda, alpha):
    ei, a = graph
    m = cbm4mm(ei, torch.ones(ei.shape[1]), alpha=alpha, device=cuda)
    layer = GINConv(m, eps=0.3, train_eps=True, device=cuda)
    x, g = _rand((N, 48), 1), _rand((N, 48), 2)
    xt = torch.from_numpy(x).to(cuda).requires_grad_(True)
    out = layer(xt)
    (out * torch.from_numpy(g).to(cuda)).sum().backward()
    x64, g64 = x.astype(np.float64), g.astype(np.float64)
    eps = float(layer.eps.detach().cpu()[0])
    assert eps == pytest.approx(0.3, abs=1e-7)
    ref = a @ x64 + (1.0 + eps) * x64
    assert O.max_rel_error(out.detach().cpu().numpy(), ref) <= 1e-6
    # A is symmetric: A^T g = A g
    assert O.max_rel_error(xt.grad.cpu().numpy(), a @ g64 + (1.0 + eps) * g64) <= 1e-6
    assert float(layer.eps.grad.cpu()[0]) == pytest.approx(float((x64 * g64).sum()), rel=1e-5)


def test_gin_with_mlp(graph, cuda):
    ei, a = graph
    m = cbm4mm(ei, torch.ones(ei.shape[1]), alpha=0, device=cuda)
    torch.manual_seed(0)
    mlp = torch.nn.Sequential(torch.nn.Linear(16, 32), torch.nn.ReLU(), torch.nn.Linear(32, 8)).to(cuda)
    layer = GINConv(m, mlp, eps=-0.25, device=cuda)
    x = _rand((N, 16), 3)
    out = layer(torch.from_numpy(x).to(cuda)).detach().cpu().double()
    w1, b1, w2, b2 = (p.detach().cpu().double() for p in mlp.parameters())
    h = torch.from_numpy(a @ x.astype(np.float64) + 0.75 * x.astype(np.float64))
    ref = torch.relu(h @ w1.T + b1) @ w2.T + b2
    assert O.max_rel_error(out.numpy(), ref.numpy()) <= 1e-5


def test_gin_aggregation_equals_the_reference_operator(graph, cuda):
    """s_spmm_update_add_csr_int32 (cbm/cbm_extensions.cpp:494-591) on the
    init arrays gives A x + (1 + eps) x too."""
    ei, a = graph
    m = cbm4mm(ei, torch.ones(ei.shape[1]), alpha=2, device=cuda)
    x = torch.from_numpy(_rand((N, 32), 4)).to(cuda)
    via_layer = GINConv(m, eps=0.5, device=cuda)(x)
    d = m.deltas
    crow, col, val = d.crow_indices(), d.col_indices(), d.values()
    out = torch.empty(N, 32, device=cuda)
    cbm_mkl_cpp.s_spmm_update_add_csr_int32(crow[:-1].contiguous(), crow[1:].contiguous(), col, val, x, 0.5,
                                            m.mca_offset, m.mca_src_idx, m.mca_dst_idx, out)
    assert O.max_rel_error(via_layer.cpu().numpy(), out.cpu().numpy()) <= 1e-6
    ref = a @ x.cpu().double().numpy() + 1.5 * x.cpu().double().numpy()
    assert O.max_rel_error(out.cpu().numpy(), ref) <= 1e-6


def test_gcnconv_forward_and_gradients(graph, cuda):
    ei, a = graph
    adj = cbm4dad(ei, torch.ones(ei.shape[1]), alpha=1, device=cuda)
    layer = GCNConv(adj, 24, 40, seed=7, device=cuda)
    with torch.no_grad():
        layer.bias.copy_(torch.linspace(-1, 1, 40, device=cuda))
    x, g = _rand((N, 24), 5), _rand((N, 40), 6)
    out = layer(torch.from_numpy(x).to(cuda))
    (out * torch.from_numpy(g).to(cuda)).sum().backward()
    m = O.product_matrix(a, "dadx")
    w = layer.weight.detach().cpu().double().numpy()
    b = layer.bias.detach().cpu().double().numpy()
    x64, g64 = x.astype(np.float64), g.astype(np.float64)
    assert O.max_rel_error(out.detach().cpu().numpy(), m @ (x64 @ w) + b) <= 1e-5
    assert O.max_rel_error(layer.weight.grad.cpu().numpy(), x64.T @ (m.T @ g64)) <= 1e-5
    assert O.max_rel_error(layer.bias.grad.cpu().numpy(), g64.sum(axis=0)) <= 1e-5


def test_gin_rejects_a_csr_adjacency(graph, cuda):
    from cbm.csr4mm import csr4mm
    ei, _ = graph
    with pytest.raises(TypeError, match="CBM kernels"):
        GINConv(csr4mm(ei, num_nodes=N, device=cuda), device=cuda)


@pytest.mark.parametrize("n", [1000, 64 * 256 * 2 + 37])
def test_weight_grad_against_fp64(cuda, n):
    """x^T g by chunks of rows in one batched product (with the rows past the
    last whole chunk added by themselves), or by one product for short x."""
    from cbm.gcn import weight_grad
    gen = torch.Generator(device=cuda).manual_seed(n)
    x = torch.randn(n, 128, device=cuda, generator=gen)
    g = torch.randn(n, 64, device=cuda, generator=gen)
    ref = x.double().t() @ g.double()
    got = weight_grad(x, g)
    assert got.shape == (128, 64) and got.dtype == torch.float32
    assert float((got.double() - ref).abs().max() / ref.abs().max()) <= 1e-5


def test_linear_gradients_match_autograd_of_a_plain_product(cuda):
    from cbm.gcn import linear
    gen = torch.Generator(device=cuda).manual_seed(3)
    x = torch.randn(40000, 32, device=cuda, generator=gen, dtype=torch.float64).float().requires_grad_(True)
    w = torch.randn(32, 16, device=cuda, generator=gen).requires_grad_(True)
    gy = torch.randn(40000, 16, device=cuda, generator=gen)
    (linear(x, w) * gy).sum().backward()
    x2 = x.detach().double().requires_grad_(True)
    w2 = w.detach().double().requires_grad_(True)
    (x2 @ w2 * gy.double()).sum().backward()
    for a, b in ((x.grad, x2.grad), (w.grad, w2.grad)):
        assert float((a.double() - b).abs().max() / b.abs().max()) <= 1e-5
