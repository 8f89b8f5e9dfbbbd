"""CPU oracle of the fingerprint row clustering -- TEST INFRASTRUCTURE ONLY.

Only tests/ may use this module, and only as the checker; the product never
imports it.  A plain-Python / numpy restatement of
benchmark/clustering_gpu.py:86-197 (the reference runs it with torch ops on
the GPU):

* in-cluster column degrees (:99-100) by counting each row's columns;
* hash columns (:101-104): the ``num_hashes`` columns whose degree is closest
  to half the cluster size, ties to the smaller column (torch.topk leaves the
  tie order open; this is the rule both sides state), in ascending order;
* fingerprints (:106-110): bit j set when the row holds hash column j;
* groups in ascending fingerprint, rows ascending inside a group; groups of
  at most ``max_cluster_size`` rows close, larger ones go on a stack, last
  pushed first (:112-138); a group that cannot be split closes as it is (the
  reference would push it again forever);
* the two-pointer merge of the size-sorted clusters (:153-176).
"""
import numpy as np


def _rows_of(indptr, indices, r):
    return indices[indptr[r]:indptr[r + 1]]


def hash_columns(indptr, indices, rows, n_cols, num_hashes):
    counts = [0] * n_cols
    for r in rows:
        for c in _rows_of(indptr, indices, int(r)):
            counts[int(c)] += 1
    m = len(rows)
    ranked = sorted(range(n_cols), key=lambda c: (abs(2 * counts[c] - m), c))
    return sorted(ranked[:num_hashes]), counts


def fingerprint(indptr, indices, r, hashes):
    held = set(int(c) for c in _rows_of(indptr, indices, int(r)))
    return sum(1 << j for j, h in enumerate(hashes) if h in held)


def clusters(indptr, indices, n_cols, max_cluster_size, num_hashes):
    n = len(indptr) - 1
    start = [r for r in range(n) if indptr[r + 1] > indptr[r]]
    stack, closed = [start], []
    while stack:
        rows = stack.pop()
        hashes, _ = hash_columns(indptr, indices, rows, n_cols, num_hashes)
        groups = {}
        for r in rows:
            groups.setdefault(fingerprint(indptr, indices, r, hashes), []).append(r)
        parts = [sorted(groups[k]) for k in sorted(groups)]
        if len(parts) == 1 and len(rows) > max_cluster_size:
            closed.append(parts[0])
            continue
        for g in parts:
            (closed if len(g) <= max_cluster_size else stack).append(g)
    return [np.array(g, dtype=np.int64) for g in closed]


def merge(clusters, max_cluster_size):
    if not clusters:
        return []
    cs = sorted(clusters, key=len)
    merged = []
    b, e = 0, len(cs) - 1
    ec = cs[e]
    while b < e:
        bc = cs[b]
        if len(bc) + len(ec) <= max_cluster_size:
            ec = np.concatenate([bc, ec])
            b += 1
        else:
            merged.append(ec)
            e -= 1
            ec = cs[e]
    merged.append(ec)
    return merged
