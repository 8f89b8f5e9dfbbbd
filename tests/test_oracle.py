"""Pins the oracle (CPU only).

The reference ships no tests, fixtures or golden vectors, so each oracle
function is checked against something the mathematics fixes:
  * golden fixtures written by brute force (tests/golden/make_golden.py);
  * networkx's Edmonds implementation and exhaustive enumeration;
  * the identity A = (I - P)^-1 Delta and the FP64 products of dense numpy;
  * planted mistakes in a copy of the oracle's logic that these checks catch.
"""
import itertools
import json
import os

import networkx as nx
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "cbm_small.json")


@pytest.fixture(scope="module")
def golden():
    with open(GOLDEN) as fh:
        return json.load(fh)["cases"]


def _dense(case):
    n = case["n"]
    A = np.zeros((n, n), np.int64)
    r, c = case["edge_index"]
    A[r, c] = 1
    return A


def test_golden_candidates_in_creation_order(golden):
    for case in golden:
        a = O.csr_of(np.array(case["edge_index"]), case["n"])
        s, d, w = O.candidates(a, case["alpha"])
        got = [list(t) for t in zip(s.tolist(), d.tolist(), w.tolist())]
        assert got == case["candidates"]


def test_golden_unique_arborescence(golden):
    for case in golden:
        out = O.cbm_init(np.array(case["edge_index"]), case["n"], case["alpha"])
        assert out["weight"] == case["weight"]
        assert out["parent"].tolist() == case["parent"]


def test_golden_delta_and_products(golden):
    for case in golden:
        n = case["n"]
        out = O.cbm_init(np.array(case["edge_index"]), n, case["alpha"])
        delta = O.delta_matrix(out["delta_row"], out["delta_col"], out["delta_val"], n).toarray()
        assert np.array_equal(delta, np.array(case["delta_dense"], np.float64))
        a = O.csr_of(np.array(case["edge_index"]), n)
        x = np.array(case["x"])
        for mode in O.MODES:
            assert np.allclose(O.forward(a, x, mode), case["products"][mode], rtol=1e-14, atol=1e-14)
            assert np.allclose(O.adjoint(a, x, mode), case["products"][mode + "_t"], rtol=1e-14, atol=1e-14)
            assert np.allclose(O.cbm_forward(a, out, x, mode), case["products"][mode], rtol=1e-13, atol=1e-13)


def _random_symmetric(rng, n, p):
    A = (rng.random((n, n)) < p).astype(np.int64)
    A = np.triu(A, 1)
    return A + A.T


def _candidate_graph(A, alpha):
    n = A.shape[0]
    deg = A.sum(1)
    S = A @ A.T
    G = nx.MultiDiGraph()
    G.add_nodes_from(range(n + 1))
    for v in range(n):
        G.add_edge(n, v, weight=int(deg[v]))
    for r in range(n):
        for c in range(r + 1, n):
            if S[r, c] == 0:
                continue
            h = int(deg[r] + deg[c] - 2 * S[r, c])
            if h < deg[r] - alpha and h != deg[r]:
                G.add_edge(c, r, weight=h)
            if h < deg[c] - alpha and h != deg[c]:
                G.add_edge(r, c, weight=h)
    return G


@pytest.mark.parametrize("seed", range(6))
def test_edmonds_weight_matches_networkx(seed):
    rng = np.random.default_rng(seed)
    for _ in range(25):
        n = int(rng.integers(2, 40))
        A = _random_symmetric(rng, n, float(rng.uniform(0.05, 0.5)))
        alpha = int(rng.integers(0, 3))
        r, c = np.nonzero(A)
        out = O.cbm_init(np.stack([r, c]), n, alpha)
        G = nx.DiGraph()
        for u, v, data in _candidate_graph(A, alpha).edges(data=True):
            if not G.has_edge(u, v) or G[u][v]["weight"] > data["weight"]:
                G.add_edge(u, v, weight=data["weight"])
        T = nx.minimum_spanning_arborescence(G)
        assert out["weight"] == sum(d["weight"] for _, _, d in T.edges(data=True))
        # nnz(Delta) equals the arborescence weight
        assert len(out["delta_row"]) == out["weight"]


def test_edmonds_matches_exhaustive_search():
    rng = np.random.default_rng(123)
    for _ in range(40):
        n = int(rng.integers(2, 6))
        m = int(rng.integers(n, 3 * n + 3))
        src = [n] * n + [int(rng.integers(0, n)) for _ in range(m)]
        dst = list(range(n)) + [int(rng.integers(0, n)) for _ in range(m)]
        w = [int(rng.integers(0, 9)) for _ in range(n + m)]
        keep = [k for k in range(len(src)) if src[k] != dst[k]]
        src, dst, w = ([x[k] for k in keep] for x in (src, dst, w))
        weight, in_edge = O.msa(n + 1, n, src, dst, w)
        best = None
        ins = [[k for k in range(len(src)) if dst[k] == v] for v in range(n)]
        for choice in itertools.product(*ins):
            par = [src[k] for k in choice]
            ok = True
            for v in range(n):
                seen, u = set(), v
                while u != n and ok:
                    ok = u not in seen
                    seen.add(u)
                    u = par[u]
            if ok:
                tot = sum(w[k] for k in choice)
                best = tot if best is None else min(best, tot)
        assert weight == best
        assert sum(w[in_edge[v]] for v in range(n)) == weight


def test_delta_reconstructs_a_exactly():
    rng = np.random.default_rng(5)
    for _ in range(20):
        n = int(rng.integers(5, 60))
        A = _random_symmetric(rng, n, float(rng.uniform(0.05, 0.4)))
        r, c = np.nonzero(A)
        for alpha in (0, 1):
            out = O.cbm_init(np.stack([r, c]), n, alpha)
            delta = O.delta_matrix(out["delta_row"], out["delta_col"], out["delta_val"], n)
            P = O.parent_matrix(out["parent"], n)
            # (I - P) A = Delta   <=>   A = (I - P)^-1 Delta
            lhs = (sp.identity(n) - P) @ sp.csr_matrix(A.astype(np.float64))
            assert np.array_equal(lhs.toarray(), delta.toarray())
            assert set(np.unique(out["delta_val"]).tolist()) <= {-1.0, 1.0}


def test_branch_arrays_are_bfs_per_branch():
    rng = np.random.default_rng(9)
    A = np.zeros((60, 60), np.int64)
    base = (rng.random(60) < 0.3).astype(np.int64)
    for i in range(60):
        A[i] = np.where(rng.random(60) < 0.9, base, 1 - base)
    A = np.triu(A, 1)
    A = A + A.T
    r, c = np.nonzero(A)
    out = O.cbm_init(np.stack([r, c]), 60, 0)
    br, src, dst = out["branches"], out["tree_src"], out["tree_dst"]
    assert br[0] == 0 and br[-1] == len(src) and np.all(np.diff(br) > 0)
    parent = out["parent"]
    for b in range(len(br) - 1):
        seen = {int(src[br[b]])}
        assert parent[src[br[b]]] == -1  # a branch starts at a root child
        for k in range(br[b], br[b + 1]):
            assert int(src[k]) in seen  # parents are visited before their children
            assert parent[dst[k]] == src[k]
            seen.add(int(dst[k]))


def test_cbm_forward_equals_direct_products():
    rng = np.random.default_rng(2)
    n = 80
    A = _random_symmetric(rng, n, 0.15)
    A[7] = 0
    A[:, 7] = 0  # an isolated node: D = 0 there
    r, c = np.nonzero(A)
    out = O.cbm_init(np.stack([r, c]), n, 0)
    a = O.csr_of(np.stack([r, c]), n)
    x = rng.standard_normal((n, 5))
    deg = A.sum(1).astype(float)
    d = np.where(deg > 0, 1 / np.sqrt(np.maximum(deg, 1)), 0.0)
    dense = {"ax": A, "adx": A @ np.diag(d), "dadx": np.diag(d) @ A @ np.diag(d)}
    for mode in O.MODES:
        np.testing.assert_allclose(O.forward(a, x, mode), dense[mode] @ x, rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(O.adjoint(a, x, mode), dense[mode].T @ x, rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(O.cbm_forward(a, out, x, mode), dense[mode] @ x, rtol=1e-12, atol=1e-12)


def test_planted_mistakes_are_caught(golden):
    """Plausible slips in the oracle's rules fail the checks above."""
    case = next(c for c in golden if sum(p >= 0 for p in c["parent"]) >= 2)
    A = _dense(case)
    n = case["n"]
    # 1. a relaxed candidate rule (<= instead of <) changes the candidate list
    #    of some fixture, so the exact-list check would catch it
    differing = 0
    for cs in golden:
        Ac, nc, al = _dense(cs), cs["n"], cs["alpha"]
        deg, S = Ac.sum(1), Ac @ Ac.T
        relaxed = []
        for r in range(nc):
            relaxed.append([nc, r, int(deg[r])])
            for c in range(r + 1, nc):
                if S[r, c]:
                    h = int(deg[r] + deg[c] - 2 * S[r, c])
                    if h <= deg[r] - al and h != deg[r]:
                        relaxed.append([c, r, h])
                    if h <= deg[c] - al and h != deg[c]:
                        relaxed.append([r, c, h])
        differing += relaxed != cs["candidates"]
    assert differing > 0
    # 2. Delta with the sign convention flipped no longer reconstructs A
    out = O.cbm_init(np.array(case["edge_index"]), n, case["alpha"])
    flipped = O.delta_matrix(out["delta_row"], out["delta_col"], -out["delta_val"], n)
    P = O.parent_matrix(out["parent"], n)
    assert not np.array_equal(((sp.identity(n) - P) @ sp.csr_matrix(A.astype(float))).toarray(), flipped.toarray())
    # 3. propagation in the wrong direction (child into parent) breaks A X
    a = O.csr_of(np.array(case["edge_index"]), n)
    x = np.array(case["x"])
    swapped = dict(out, tree_src=out["tree_dst"], tree_dst=out["tree_src"])
    assert not np.allclose(O.cbm_forward(a, swapped, x, "ax"), case["products"]["ax"])
    # 4. the transposed operand in the adjoint breaks A D (not symmetric)
    assert not np.allclose(O.forward(a, x, "adx"), case["products"]["adx_t"])


def test_sorted_edge_list_csr_equals_coo_build():
    """csr_of_sorted (counting build for sorted unique edges) gives csr_of's
    matrix, and refuses unsorted or duplicated input."""
    rng = np.random.default_rng(5)
    n = 300
    pairs = np.unique(rng.integers(0, n, size=(2000, 2)), axis=0)
    ei = pairs.T.copy()
    a, b = O.csr_of(ei, n), O.csr_of_sorted(ei, n)
    assert np.array_equal(a.indptr, b.indptr) and np.array_equal(a.indices, b.indices)
    assert np.array_equal(a.data, b.data)
    with pytest.raises(ValueError):
        O.csr_of_sorted(ei[:, ::-1], n)
    with pytest.raises(ValueError):
        O.csr_of_sorted(np.concatenate([ei, ei[:, :1]], axis=1), n)
