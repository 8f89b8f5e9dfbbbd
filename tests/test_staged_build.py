"""The host build in stages (adjacency CSR, candidate edges, structure from a
candidate list), the path the GPU candidate scan feeds: each stage against the
oracle and the one-call build, and the error cases of a caller-supplied list
(CPU only)."""
import numpy as np
import pytest
import torch

from cbm import structure
from hubgraph import hub_graph
from oracle import oracle as O

ARRAYS = ("delta_row", "delta_col", "delta_val", "branches", "tree_src", "tree_dst", "parent", "csr_ptr", "csr_col",
          "csr_val")


def _random_graph(rng, n, dens, copied):
    a = (rng.random((n, n)) < dens).astype(np.int64)
    if copied:
        for i in range(1, n):
            j = int(rng.integers(0, i))
            a[i] = np.where(rng.random(n) < 0.9, a[j], a[i])
    a = np.triu(a, 1)
    a = a + a.T
    r, c = np.nonzero(a)
    return torch.from_numpy(np.stack([r, c]).astype(np.int32))


def _graphs():
    rng = np.random.default_rng(123)
    out = [(_random_graph(rng, n, d, cp), n) for n, d, cp in
           ((40, 0.2, False), (90, 0.1, True), (150, 0.05, True), (64, 0.5, False))]
    c1 = structure.copying_model(2708, 7, 0.8, 0.58, 0, 1001)
    out.append((c1, 2708))
    # not symmetric: a row subset (as the row clustering builds) and a directed graph
    out.append((c1[:, (c1[0] % 3) == 0].contiguous(), 2708))
    d = (rng.random((150, 150)) < 0.06) & ~np.eye(150, dtype=bool)
    r, c = np.nonzero(d)
    out.append((torch.from_numpy(np.stack([r, c]).astype(np.int32)), 150))
    return out


def _stages(ei, n, alpha):
    adj = structure.Adjacency(ei[0], ei[1], torch.ones(ei.shape[1]), n)
    cand = structure.host_candidates(adj, alpha)
    return adj, cand


@pytest.mark.parametrize("alpha", [0, 1, 3, 8])
def test_host_candidates_equal_the_oracle(alpha):
    """The candidate list entry for entry (src, dst, weight, creation order)."""
    for ei, n in _graphs():
        adj, cand = _stages(ei, n, alpha)
        want = O.candidates(O.csr_of(ei.numpy(), n), alpha)
        for got, ref in zip(cand, want):
            assert np.array_equal(got.numpy(), ref)
        adj.close()


def test_adjacency_is_the_sorted_binary_csr():
    ei = structure.copying_model(500, 2, 0.8, 1.0, 0, 5)
    dup = torch.cat([ei, ei[:, :50]], dim=1)  # repeated entries are merged
    perm = torch.randperm(dup.shape[1], generator=torch.Generator().manual_seed(1))
    dup = dup[:, perm]
    with structure.Adjacency(dup[0], dup[1], torch.ones(dup.shape[1]), 500) as adj:
        a = O.csr_of(ei.numpy(), 500)
        assert adj.n == 500 and adj.nnz == a.nnz
        assert np.array_equal(adj.ptr.numpy(), a.indptr)
        assert np.array_equal(adj.col.numpy(), a.indices)
    with pytest.raises(ValueError, match="binary"):
        structure.Adjacency(ei[0], ei[1], torch.full((ei.shape[1],), 2.0), 500)


def test_adjacency_transpose():
    """A^T for a directed matrix; A itself for a symmetric one."""
    ei = structure.copying_model(500, 2, 0.8, 1.0, 0, 5)
    with structure.Adjacency(ei[0], ei[1], torch.ones(ei.shape[1]), 500) as adj:
        assert adj.symmetric and adj.t_col is adj.col
    half = ei[:, ei[0] < ei[1]]
    with structure.Adjacency(half[0], half[1], torch.ones(half.shape[1]), 500) as adj:
        assert not adj.symmetric
        at = O.csr_of(half.numpy(), 500).T.tocsr()
        at.sort_indices()
        assert np.array_equal(adj.t_ptr.numpy(), at.indptr)
        assert np.array_equal(adj.t_col.numpy(), at.indices)


@pytest.mark.parametrize("alpha", [0, 4])
def test_staged_build_equals_one_call_build(alpha):
    """Every init array, the parent array, Delta as CSR and the statistics."""
    graphs = _graphs() + [(hub_graph(1200), 1200)]
    for ei, n in graphs:
        adj, cand = _stages(ei, n, alpha)
        staged = structure.build_from_candidates(adj, *cand)
        whole = structure.build(ei[0], ei[1], torch.ones(ei.shape[1]), n, alpha)
        for k in ARRAYS:
            assert torch.equal(getattr(staged, k), getattr(whole, k)), k
        assert staged.stats == whole.stats
        adj.close()


def test_caller_candidate_lists_are_checked():
    ei = structure.copying_model(300, 2, 0.8, 0.6, 0, 3)
    adj, (src, dst, w) = _stages(ei, 300, 0)
    bad = src.clone()
    bad[5] = 301  # beyond the virtual root N = 300
    with pytest.raises(ValueError, match="not an edge"):
        structure.build_from_candidates(adj, bad, dst, w)
    loop = src.clone()
    loop[3] = dst[3]
    with pytest.raises(ValueError, match="not an edge"):
        structure.build_from_candidates(adj, loop, dst, w)
    # without the root edges nothing reaches the rows that have no candidate parent
    keep = src != 300
    with pytest.raises(ValueError, match="spanning"):
        structure.build_from_candidates(adj, src[keep], dst[keep], w[keep])
    # weights one too large everywhere: the chosen tree weighs nnz(Delta) + N
    with pytest.raises(ValueError, match="row distances"):
        structure.build_from_candidates(adj, src, dst, w + 1)
    with pytest.raises(ValueError, match="differ in length"):
        structure.build_from_candidates(adj, src, dst[:-1], w)
    adj.close()
