"""The uncompressed form: the cbm4* classes keep the same product over A
itself beside the CBM structure (small matrices) and the handle runs it
where its cost model predicts it faster.  Checked: which form the model
picks for the small and the large benchmark shapes, the routed products
against the FP64 oracle (forward and adjoint, FP32 and BF16), forced
routing both ways, and graph capture of the routed step."""
import numpy as np
import pytest
import torch

import bench
from cbm import structure
from cbm.cbm4ad import cbm4ad
from cbm.cbm4dad import cbm4dad
from cbm.cbm4mm import cbm4mm
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CLASSES = {"ax": cbm4mm, "adx": cbm4ad, "dadx": cbm4dad}


def _graph(name, alpha=None):
    w = bench.WORKLOADS[name]
    ei = structure.copying_model(w["n"], w["blocks"], w["p"], w["lam"], w["seed_deg"], w["seed"])
    return ei, w["n"], w["alpha"] if alpha is None else alpha


def _rand(shape, seed):
    return np.random.default_rng(seed).standard_normal(shape).astype(np.float32)


@pytest.mark.parametrize("mode", ["ax", "adx", "dadx"])
def test_small_graph_runs_over_a_and_matches_the_oracle(cuda, mode):
    ei, n, alpha = _graph("c1")
    a = O.csr_of(ei.numpy(), n)
    m = CLASSES[mode](ei, torch.ones(ei.shape[1]), alpha=alpha, device=cuda, uncompressed=True)
    h = m.cusparse_extension
    h.reserve(64)
    for f, dtype in ((64, torch.float32), (64, torch.bfloat16), (5, torch.float32)):
        assert h.uncompressed_active(dtype, f)  # depth 9: the tree's launches cost more
        x = torch.from_numpy(_rand((n, f), 3 + f)).to(cuda).to(dtype)
        y = torch.empty(n, f, device=cuda)
        dx = torch.empty(n, f, device=cuda)
        m.matmul(x, y)
        m.matmul_t(x, dx)
        xr = x.float().cpu().numpy()
        assert O.max_rel_error(y.cpu().numpy(), O.forward(a, xr, mode)) <= 1e-5
        assert O.max_rel_error(dx.cpu().numpy(), O.adjoint(a, xr, mode)) <= 1e-5
    # forced either way, the same results within FP32 rounding
    x = torch.from_numpy(_rand((n, 64), 9)).to(cuda)
    y_plain, y_cbm = torch.empty(n, 64, device=cuda), torch.empty(n, 64, device=cuda)
    h.attach_uncompressed(h._plain, True)
    m.matmul(x, y_plain)
    h.attach_uncompressed(h._plain, False)
    assert not h.uncompressed_active(torch.float32, 64)
    m.matmul(x, y_cbm)
    h.attach_uncompressed(h._plain, None)
    assert float((y_plain - y_cbm).abs().max()) <= 1e-5 * float(y_cbm.abs().max())


def test_the_model_keeps_compressed_where_delta_saves_gathers(cuda):
    """c2 (19,717 nodes, depth 12 at alpha 0) runs over A; a dense graph
    whose Delta halves the gathers (the c4 shape at 1/8 of its rows) runs
    compressed at F = 256."""
    ei, n, _ = _graph("c2", 0)
    h = cbm4dad(ei, torch.ones(ei.shape[1]), alpha=0, device=cuda, uncompressed=True).cusparse_extension
    h.reserve(128)
    assert h.uncompressed_active(torch.bfloat16, 128)
    nd = 16_000
    ed = structure.copying_model(nd, 8, 0.95, 20.0, 225, 1004)
    hd = cbm4ad(ed, torch.ones(ed.shape[1]), alpha=32, device=cuda, uncompressed=True).cusparse_extension
    hd.reserve(256)
    assert hd.nnz < 0.7 * ed.shape[1]
    assert not hd.uncompressed_active(torch.bfloat16, 256)


def test_routed_step_replays_from_a_graph(cuda):
    ei, n, alpha = _graph("c1")
    m = cbm4dad(ei, torch.ones(ei.shape[1]), alpha=alpha, device=cuda, uncompressed=True)
    h = m.cusparse_extension
    h.reserve(32)
    assert h.uncompressed_active(torch.float32, 32)
    x = torch.from_numpy(_rand((n, 32), 1)).to(cuda)
    y, dx = torch.empty(n, 32, device=cuda), torch.empty(n, 32, device=cuda)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        h.forward(x, y)
        h.matmul_t(x, dx)
    torch.cuda.current_stream().wait_stream(s)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        h.forward(x, y)
        h.matmul_t(x, dx)
    for it in range(3):
        x.copy_(torch.from_numpy(_rand((n, 32), 10 + it)).to(cuda))
        y_ref, dx_ref = torch.empty_like(y), torch.empty_like(dx)
        h.forward(x, y_ref)
        h.matmul_t(x, dx_ref)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(y, y_ref) and torch.equal(dx, dx_ref)
