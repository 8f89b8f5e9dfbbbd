"""The GPU candidate-edge scan (cand_api.hip, through the C ABI) against the
host scan and the oracle: the same list entry for entry, by both counter
paths (LDS tiles, global touched lists), on sparse graphs, dense ones (rows
sharing columns with thousands of later rows, several LDS tiles per row),
with 1 to many persistent workgroups; the full build fed by it equals the
host build; writes never pass the caller's array length."""
import ctypes

import numpy as np
import pytest
import torch

from cbm import _lib, structure
from cbm.cbm4mm import cbm4mm
from hubgraph import HUB_N, hub_graph
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ARRAYS = ("delta_row", "delta_col", "delta_val", "branches", "tree_src", "tree_dst", "parent", "csr_ptr", "csr_col",
          "csr_val")


def _random_graph(seed, n, dens):
    rng = np.random.default_rng(seed)
    a = (rng.random((n, n)) < dens).astype(np.int64)
    for i in range(1, n):
        j = int(rng.integers(0, i))
        a[i] = np.where(rng.random(n) < 0.85, a[j], a[i])
    a = np.triu(a, 1)
    a = a + a.T
    r, c = np.nonzero(a)
    return torch.from_numpy(np.stack([r, c]).astype(np.int32)), n


@pytest.fixture(scope="module")
def graphs():
    out = {
        "random40": _random_graph(1, 40, 0.2),
        "random300": _random_graph(2, 300, 0.05),
        "c1": (structure.copying_model(2708, 7, 0.8, 0.58, 0, 1001), 2708),
        "hub": (hub_graph(), HUB_N),
    }
    # isolated trailing rows and a single edge
    out["sparse"] = (torch.tensor([[0, 3], [3, 0]], dtype=torch.int32), 7)
    # not symmetric: a row subset of c1 (the matrices the row clustering
    # builds) and a random directed graph, where row overlaps need A^T
    c1, _ = out["c1"]
    keep = (c1[0] % 3) == 0
    out["c1_rows"] = (c1[:, keep].contiguous(), 2708)
    rng = np.random.default_rng(5)
    d = (rng.random((400, 400)) < 0.04) & ~np.eye(400, dtype=bool)
    r, c = np.nonzero(d)
    out["directed"] = (torch.from_numpy(np.stack([r, c]).astype(np.int32)), 400)
    return out


def _check_same(got, want):
    for g, w in zip(got, want):
        g = g.cpu().numpy() if isinstance(g, torch.Tensor) else g
        w = w.cpu().numpy() if isinstance(w, torch.Tensor) else w
        assert np.array_equal(g, w)


@pytest.mark.parametrize("path", ["tiles", "lists"])
@pytest.mark.parametrize("alpha", [0, 2, 8])
@pytest.mark.parametrize("name", ["random40", "random300", "c1", "sparse", "c1_rows", "directed"])
def test_gpu_scan_equals_host_scan_and_oracle(graphs, cuda, name, alpha, path):
    ei, n = graphs[name]
    print(f"candidates: {name} n={n} alpha={alpha} {path}", flush=True)
    with structure.Adjacency(ei[0], ei[1], torch.ones(ei.shape[1]), n) as adj:
        host = structure.host_candidates(adj, alpha)
        for wgs in (1, 7, 0):
            _check_same(structure.gpu_candidates(adj, alpha, cuda, path=path, n_workgroups=wgs), host)
    _check_same(host, O.candidates(O.csr_of(ei.numpy(), n), alpha))


def _max_passing_columns(src, dst, n):
    """Most distinct columns c with an edge in one row's part of a list."""
    s, d = src.numpy(), dst.numpy()
    root = np.nonzero(s == n)[0]
    ends = np.append(root[1:], len(s))
    best = 0
    for a, b in zip(root, ends):
        r = d[a]
        best = max(best, len(np.unique(np.where(s[a + 1:b] == r, d[a + 1:b], s[a + 1:b]))))
    return best


@pytest.mark.parametrize("alpha", [0, 32])
def test_dense_rows(graphs, cuda, alpha):
    """The hub graph's early rows share columns with nearly every later row:
    thousands of touched columns per row, and rows with more than the 4096
    passing columns the lists path sorts in LDS (it walks their counters)."""
    ei, n = graphs["hub"]
    print(f"candidates: hub n={n} alpha={alpha}", flush=True)
    with structure.Adjacency(ei[0], ei[1], torch.ones(ei.shape[1]), n) as adj:
        a = O.csr_of(ei.numpy(), n)
        s = (a[:50] @ a.T).tocsr()
        touched = [int((s[r].indices > r).sum()) for r in range(50)]
        assert max(touched) > 4096
        host = structure.host_candidates(adj, alpha)
        assert _max_passing_columns(host[0], host[1], n) > 4096
        for path, wgs in (("tiles", 5), ("tiles", 0), ("lists", 5), ("lists", 0)):
            _check_same(structure.gpu_candidates(adj, alpha, cuda, path=path, n_workgroups=wgs), host)


def test_rows_spanning_many_tiles(cuda):
    """A 100,000-row graph: rows cross up to four 32768-column LDS tiles, with
    hub columns in every tile."""
    n = 100_000
    ei = structure.copying_model(n, 4, 0.8, 1.0, 0, 31).long()
    hubs = torch.tensor([5, 40_000, 70_000, 99_990])
    nb = torch.arange(0, n, 37)
    extra = [torch.stack([torch.full_like(nb, h), nb]) for h in hubs]
    extra += [torch.stack([nb, torch.full_like(nb, h)]) for h in hubs]
    allr = torch.cat([ei] + extra, dim=1)
    allr = allr[:, allr[0] != allr[1]]
    key = torch.unique(allr[0] * n + allr[1])
    ei = torch.stack([key // n, key % n]).to(torch.int32)
    print(f"candidates: tiles n={n} nnz={ei.shape[1]}", flush=True)
    with structure.Adjacency(ei[0], ei[1], torch.ones(ei.shape[1]), n) as adj:
        for alpha in (0, 3):
            host = structure.host_candidates(adj, alpha)
            for path in ("tiles", "lists"):
                _check_same(structure.gpu_candidates(adj, alpha, cuda, path=path), host)


def test_large_sparse_graph_takes_the_lists(cuda):
    """Above 262144 rows the automatic choice is the touched-list path."""
    n = 300_000
    ei = structure.copying_model(n, 6, 0.7, 1.0, 0, 17)
    print(f"candidates: lists n={n} nnz={ei.shape[1]}", flush=True)
    hip = _lib.hip()
    assert hip.cbm_candidates_workspace_bytes(n, 0, 3) == hip.cbm_candidates_workspace_bytes(n, 2, 3) > 256
    with structure.Adjacency(ei[0], ei[1], torch.ones(ei.shape[1]), n) as adj:
        host = structure.host_candidates(adj, 1)
        _check_same(structure.gpu_candidates(adj, 1, cuda), host)


@pytest.mark.parametrize("name,alpha", [("c1", 0), ("c1", 4), ("hub", 0), ("random300", 1), ("c1_rows", 0),
                                        ("directed", 2)])
def test_build_with_gpu_candidates_equals_host_build(graphs, cuda, name, alpha):
    ei, n = graphs[name]
    print(f"build: {name} n={n} alpha={alpha}", flush=True)
    gpu = structure.build(ei[0], ei[1], torch.ones(ei.shape[1]), n, alpha, candidates_on=cuda)
    host = structure.build(ei[0], ei[1], torch.ones(ei.shape[1]), n, alpha)
    for k in ARRAYS:
        assert torch.equal(getattr(gpu, k), getattr(host, k)), k
    assert gpu.stats == host.stats


def test_classes_scan_on_the_device_by_default(graphs, cuda):
    ei, n = graphs["c1"]
    m_dev = cbm4mm(ei, torch.ones(ei.shape[1]), alpha=2, device=cuda)
    m_host = cbm4mm(ei, torch.ones(ei.shape[1]), alpha=2, device=cuda, candidates_on=None)
    assert m_dev.build_stats == m_host.build_stats
    assert torch.equal(m_dev.mca_src_idx, m_host.mca_src_idx)
    assert torch.equal(m_dev.mca_dst_idx, m_host.mca_dst_idx)
    assert torch.equal(m_dev.deltas.col_indices(), m_host.deltas.col_indices())


def test_emit_stays_inside_the_callers_arrays(graphs, cuda):
    """An n_cand shorter than the list: the first n_cand entries are written,
    nothing after them."""
    ei, n = graphs["random300"]
    hip = _lib.hip()
    with structure.Adjacency(ei[0], ei[1], torch.ones(ei.shape[1]), n) as adj:
        assert adj.symmetric  # A^T is A itself
        host = structure.host_candidates(adj, 0)
        total = host[0].numel()
        cap = total // 2
        ptr, col = adj.ptr.to(cuda), adj.col.to(cuda)
        stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        for path in (1, 2):
            ws_bytes = int(hip.cbm_candidates_workspace_bytes(n, path, 3))
            ws = torch.empty(ws_bytes, dtype=torch.uint8, device=cuda)
            assert ws.data_ptr() % 256 == 0
            counts = torch.empty(n, dtype=torch.int64, device=cuda)
            _lib.check_candidates(hip.cbm_candidates_count(n, ptr.data_ptr(), col.data_ptr(), ptr.data_ptr(),
                                                           col.data_ptr(), 0, path, 3,
                                                           counts.data_ptr(), ws.data_ptr(), ws_bytes, stream))
            offsets = torch.cumsum(counts, 0) - counts
            assert int(counts.sum()) == total
            out = [torch.full((total + 64,), -7, dtype=torch.int32, device=cuda) for _ in range(3)]
            _lib.check_candidates(hip.cbm_candidates_emit(n, ptr.data_ptr(), col.data_ptr(), ptr.data_ptr(),
                                                          col.data_ptr(), 0, path, 3,
                                                          offsets.data_ptr(), out[0].data_ptr(), out[1].data_ptr(),
                                                          out[2].data_ptr(), cap, ws.data_ptr(), ws_bytes, stream))
            torch.cuda.synchronize()
            for o, h in zip(out, host):
                o = o.cpu()
                assert torch.equal(o[:cap], h[:cap])
                assert bool((o[cap:] == -7).all())


def test_bad_arguments_are_rejected(cuda):
    hip = _lib.hip()
    assert hip.cbm_candidates_workspace_bytes(0, 0, 0) == -1
    assert hip.cbm_candidates_workspace_bytes(100, 3, 0) == -1
    assert hip.cbm_candidates_workspace_bytes(100, 2, 4) == 256 + 4 * 100 * 8
    assert hip.cbm_candidates_workspace_bytes(100, 1, 4) == 256
    ptr = torch.zeros(3, dtype=torch.int64, device=cuda)
    col = torch.zeros(1, dtype=torch.int32, device=cuda)
    counts = torch.empty(2, dtype=torch.int64, device=cuda)
    ws = torch.empty(4096, dtype=torch.uint8, device=cuda)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    with pytest.raises(ValueError, match="aligned"):
        _lib.check_candidates(hip.cbm_candidates_count(2, ptr.data_ptr(), col.data_ptr(), ptr.data_ptr(), col.data_ptr(), 0, 2, 1,
                                                       counts.data_ptr(), ws.data_ptr() + 4, 1024, stream))
    with pytest.raises(ValueError, match="needed"):
        _lib.check_candidates(hip.cbm_candidates_count(2, ptr.data_ptr(), col.data_ptr(), ptr.data_ptr(), col.data_ptr(), 0, 2, 1,
                                                       counts.data_ptr(), ws.data_ptr(), 260, stream))
    with pytest.raises(ValueError, match="invalid"):
        _lib.check_candidates(hip.cbm_candidates_count(2, ptr.data_ptr(), col.data_ptr(), ptr.data_ptr(), col.data_ptr(), 0, 7, 1,
                                                       counts.data_ptr(), ws.data_ptr(), 4096, stream))
