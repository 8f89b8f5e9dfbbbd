"""Fingerprint row clustering, host side (CPU only): the oracle's
restatement against hand-worked cases, and the product's merge against the
oracle's."""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

from cbm.clustering import merge_clusters
from oracle import cluster as OC


def _csr(rows, n):
    r = [i for i, cols in enumerate(rows) for _ in cols]
    c = [x for cols in rows for x in cols]
    a = sp.csr_matrix((np.ones(len(r)), (r, c)), shape=(n, n))
    a.sort_indices()
    return a.indptr.astype(np.int64), a.indices.astype(np.int32)


def test_two_blocks_split_on_their_columns():
    """Rows 0-2 hold {0, 1}, rows 3-5 hold {2, 3}: columns 0-3 have degree 3 =
    m / 2, 4 and 5 degree 0; with two hashes {0, 1} the fingerprints are 3 and
    0, so the fingerprint-0 block closes first."""
    indptr, indices = _csr([[0, 1]] * 3 + [[2, 3]] * 3, 6)
    hashes, counts = OC.hash_columns(indptr, indices, list(range(6)), 6, 2)
    assert counts == [3, 3, 3, 3, 0, 0] and hashes == [0, 1]
    out = OC.clusters(indptr, indices, 6, 3, 2)
    assert [c.tolist() for c in out] == [[3, 4, 5], [0, 1, 2]]


def test_recursion_pops_the_last_group_first():
    """max 2: the first round gives groups {3,4,5} (fp 0) and {0,1,2} (fp 3),
    both too large; {0,1,2} is split first.  Inside it every column has
    degree 3 or 0 against m = 3 (distance 3 everywhere), so hashes {0, 1}
    again give one group: it closes as it is, then {3,4,5} likewise."""
    indptr, indices = _csr([[0, 1]] * 3 + [[2, 3]] * 3, 6)
    out = OC.clusters(indptr, indices, 6, 2, 2)
    assert [c.tolist() for c in out] == [[0, 1, 2], [3, 4, 5]]


def test_rows_without_entries_belong_to_no_cluster():
    indptr, indices = _csr([[1], [], [0], []], 4)
    out = OC.clusters(indptr, indices, 4, 10, 1)
    assert sorted(np.concatenate(out).tolist()) == [0, 2]


def test_merge_by_hand():
    """Sizes 1, 2, 3, 10 with cap 6: the 10 stays alone; the 3 absorbs the 1
    (in front), then the 2 (in front again)."""
    cs = [np.arange(k) + 100 * k for k in (3, 1, 10, 2)]
    merged = OC.merge(cs, 6)
    assert [len(c) for c in merged] == [10, 6]
    assert merged[1].tolist() == [200, 201, 100, 300, 301, 302]


def test_product_merge_equals_the_oracle():
    rng = np.random.default_rng(4)
    for trial in range(40):
        sizes = rng.integers(1, 40, size=int(rng.integers(1, 12)))
        start = np.concatenate([[0], np.cumsum(sizes)])
        cs = [np.arange(start[i], start[i + 1]) for i in range(len(sizes))]
        cap = int(rng.integers(1, 80))
        want = OC.merge(cs, cap)
        got = merge_clusters([torch.from_numpy(c) for c in cs], cap)
        assert len(got) == len(want)
        for g, w in zip(got, want):
            assert g.tolist() == w.tolist()
    assert merge_clusters([], 5) == []


def test_oracle_partitions_random_graphs():
    rng = np.random.default_rng(8)
    for trial in range(10):
        n = int(rng.integers(20, 120))
        a = sp.random(n, n, density=0.08, random_state=int(rng.integers(1 << 30)), format="csr")
        a = ((a + a.T) > 0).astype(np.float64).tocsr()
        a.sort_indices()
        indptr, indices = a.indptr.astype(np.int64), a.indices.astype(np.int32)
        cap = int(rng.integers(2, 30))
        out = OC.clusters(indptr, indices, n, cap, int(rng.integers(1, 6)))
        allr = np.concatenate(out) if out else np.zeros(0, np.int64)
        nonempty = np.nonzero(np.diff(indptr))[0]
        assert sorted(allr.tolist()) == nonempty.tolist()
        for c in out:
            assert len(c) >= 1 and c.tolist() == sorted(c.tolist())
