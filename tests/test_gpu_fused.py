"""The forward product in one launch (k_fused_forward over the plan's
units; opt-in, forced on here): Delta pieces gathered into LDS, propagation level by level in LDS,
outside parents read after their unit's flag.  Checked against the FP64
oracle (normwise relative error <= 1e-5, FP32 and BF16 storage against the
rounded inputs), integer X bitwise exact, equal within FP32 rounding to the
multi-launch forms, deterministic across repeated calls (the flags' epochs
roll over from call to call), replayed from a HIP graph, and with no
dependency wait that gave up (cbm_fused_errors).  Graphs: the smallest
benchmark workload (2,708 nodes: units with outside parents), the hub graph
(rows of up to ~2.4k entries: many pieces per row), the star graph, a chain
of depth 120 (components stacked many levels deep) and the 19,717-node
workload at alpha = 0.  The adjoint of these symmetric graphs runs the same
kernel (M^T = M for A and D A D, D A for A D)."""
import numpy as np
import pytest
import torch

import bench
from cbm import structure
from cbm.cbm4ad import cbm4ad
from cbm.cbm4dad import cbm4dad
from cbm.cbm4mm import cbm4mm
from hubgraph import HUB_N, hub_graph, star_graph
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CLASSES = {"ax": cbm4mm, "adx": cbm4ad, "dadx": cbm4dad}


def _chain(m=120):
    pairs = sorted({(i, m + j) for i in range(m) for j in range(m - i)} |
                   {(m + j, i) for i in range(m) for j in range(m - i)})
    return torch.tensor(pairs, dtype=torch.int32).t().contiguous(), 2 * m


def _graph(name):
    if name == "c1":
        w = bench.WORKLOADS["c1"]
        return structure.copying_model(w["n"], w["blocks"], w["p"], w["lam"], w["seed_deg"], w["seed"]), w["n"]
    if name == "c2":
        w = bench.WORKLOADS["c2"]
        return structure.copying_model(w["n"], w["blocks"], w["p"], w["lam"], w["seed_deg"], w["seed"]), w["n"]
    if name == "hub":
        return hub_graph(), HUB_N
    if name == "star":
        return star_graph()
    return _chain()


_MATS = {}


def _mats(name, cuda):
    if name not in _MATS:
        _MATS.clear()
        ei, n = _graph(name)
        mats = {mode: cls(ei, torch.ones(ei.shape[1]), alpha=0, device=cuda) for mode, cls in CLASSES.items()}
        for m in mats.values():
            if m.cusparse_extension.depth > 32:  # the hub graph (46) and the chain (120)
                m.cusparse_extension.set_kahan_depth(None)  # no compensated levels: one launch possible
        _MATS[name] = (O.csr_of(ei.numpy(), n), n, mats)
    return _MATS[name]


def _rand(shape, seed):
    return np.random.default_rng(seed).standard_normal(shape).astype(np.float32)


def _run(h, x, fused=True):
    """The forward with the one-launch kernel forced on (or off)."""
    n, f = x.shape
    h.set_fused(fused)
    y = torch.full((n, f), float("nan"), device=x.device)
    h.forward(x, y)
    h.set_fused(None)
    return y


GRAPHS = ["c1", "hub", "star", "chain", "c2"]


@pytest.mark.parametrize("graph", GRAPHS)
@pytest.mark.parametrize("mode", ["ax", "adx", "dadx"])
def test_fused_forward_and_adjoint_vs_oracle(cuda, graph, mode):
    a, n, mats = _mats(graph, cuda)
    h = mats[mode].cusparse_extension
    assert h.plan_stats["units"] > 0
    for f, dtype in ((1, torch.float32), (3, torch.float32), (64, torch.float32), (100, torch.float32),
                     (256, torch.float32), (64, torch.bfloat16), (136, torch.bfloat16)):
        x = torch.from_numpy(_rand((n, f), 10 + f)).to(cuda).to(dtype)
        h.reserve(f)
        h.set_fused(True)
        assert h.fused_active(dtype, f), (graph, mode, f)
        y = _run(h, x)
        ref_x = x.float().cpu().numpy()
        err = O.max_rel_error(y.cpu().numpy(), O.forward(a, ref_x, mode))
        assert err <= 1e-5, (graph, mode, f, dtype, err)
        dx = torch.full((n, f), float("nan"), device=cuda)
        h.set_fused(True)
        h.matmul_t(x, dx)  # symmetric: the same one-launch kernel
        assert h.last_launches() == 1
        h.set_fused(None)
        err = O.max_rel_error(dx.cpu().numpy(), O.adjoint(a, ref_x, mode))
        assert err <= 1e-5, (graph, mode, f, dtype, "adjoint", err)
    assert h.fused_errors() == 0


@pytest.mark.parametrize("graph", GRAPHS)
def test_fused_integer_exact_deterministic_and_close_to_multi_launch(cuda, graph):
    a, n, mats = _mats(graph, cuda)
    h = mats["ax"].cusparse_extension
    f = 72
    h.reserve(f)
    x = np.random.default_rng(5).integers(-8, 9, size=(n, f)).astype(np.float32)
    xt = torch.from_numpy(x).to(cuda)
    exact = O.forward(a, x, "ax")
    first = _run(h, xt)
    assert np.array_equal(first.cpu().numpy().astype(np.float64), exact)
    for _ in range(3):  # the epochs roll over call after call
        assert torch.equal(_run(h, xt), first)
    xr = torch.from_numpy(_rand((n, f), 6)).to(cuda)
    for mode, m in mats.items():
        hm = m.cusparse_extension
        hm.reserve(f)
        one, multi = _run(hm, xr), _run(hm, xr, fused=False)
        scale = float(multi.abs().max())
        assert float((one - multi).abs().max()) <= 1e-5 * scale, mode
        assert torch.equal(_run(hm, xr), one)
    assert h.fused_errors() == 0


def test_fused_graph_replay_and_launch_count(cuda):
    """One launch per product; forward + adjoint captured in a HIP graph and
    replayed with eager work between replays, each replay equal to the eager
    results (the flags and counters carry over between replays)."""
    a, n, mats = _mats("c1", cuda)
    h = mats["dadx"].cusparse_extension
    f = 64
    h.reserve(f)
    h.set_fused(True)
    x = torch.from_numpy(_rand((n, f), 1)).to(cuda)
    g = torch.from_numpy(_rand((n, f), 2)).to(cuda)
    y = torch.empty(n, f, device=cuda)
    dx = torch.empty(n, f, device=cuda)
    h.forward(x, y)
    assert h.last_launches() == 1
    h.matmul_t(g, dx)
    assert h.last_launches() == 1
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        h.forward(x, y)
        h.matmul_t(g, dx)
    torch.cuda.current_stream().wait_stream(s)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        h.forward(x, y)
        h.matmul_t(g, dx)
    for it in range(4):
        x.copy_(torch.from_numpy(_rand((n, f), 20 + it)).to(cuda))
        g.copy_(torch.from_numpy(_rand((n, f), 30 + it)).to(cuda))
        y_ref, dx_ref = torch.empty_like(y), torch.empty_like(dx)
        h.forward(x, y_ref)
        h.matmul_t(g, dx_ref)
        torch.cuda.synchronize()
        y.zero_()
        dx.zero_()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(y, y_ref) and torch.equal(dx, dx_ref), it
    h.set_fused(None)
    assert h.fused_errors() == 0


def test_fused_is_not_used_where_it_cannot_be(cuda):
    """Compensated levels, or a width not reserved, fall back to the
    multi-launch forms."""
    ei, n = _chain()
    h = cbm4mm(ei, torch.ones(ei.shape[1]), alpha=0, device=cuda).cusparse_extension
    h.set_kahan_depth(None)
    h.set_fused(True)
    assert not h.fused_active(torch.float32, 16)  # nothing reserved: no flags yet
    h.reserve(16)
    assert h.fused_active(torch.float32, 16)
    assert not h.fused_active(torch.float32, 256)  # flags sized for one 64-column slab
    h.set_kahan_depth(32)
    assert not h.fused_active(torch.float32, 16)
    h.set_kahan_depth(None)
    h.set_fused(False)
    assert not h.fused_active(torch.float32, 16)
