"""Fingerprint row clustering and per-cluster CBM compression: the
experiment of benchmark/clustering_gpu.py:40-215.

Rows of A are split recursively: for a cluster of m rows, the ``num_hashes``
columns whose degree inside the cluster is closest to m / 2 are hash columns,
and each row's fingerprint has bit j set when it holds hash column j; rows
with equal fingerprints form the next clusters (ascending fingerprint), which
close at ``max_cluster_size`` rows or less and are split again otherwise
(last one first).  Small clusters are then merged (two pointers over the
size-sorted list), and each merged cluster's rows get their own CBM.  The
data passes (in-cluster column degrees, fingerprints) are HIP kernels
(cluster_api.hip); the recursion and the merge are host logic.

Deliberate differences: hash columns of equal distance go to the smaller
column (torch.topk leaves that order open), rows are ascending inside a
cluster, and a cluster that one round cannot split closes as it is (the
reference would push it again forever).
"""
import ctypes

import torch

from . import _lib, structure
from .cbm_cusparse import current_stream_handle

MAX_HASHES = 30


def _ok(rc, what):
    if rc != _lib.CBM_OK:
        _lib._raise(rc, _lib.hip().cbm_cluster_last_error().decode(errors="replace"), what)


def _edges(edge_index, num_nodes):
    ei = torch.as_tensor(edge_index).to("cpu", torch.int32)
    n = int(num_nodes) if num_nodes is not None else int(ei.max()) + 1 if ei.numel() else 0
    return ei, n


def lsh_clusters(edge_index, num_nodes=None, *, max_cluster_size=100000, num_hashes=4, device="cuda"):
    """Row clusters of the binary matrix given by ``edge_index`` ``[2, nnz]``
    (clustering_gpu.py:86-138); returns a list of int64 CPU tensors of row
    ids, in the order the clusters closed.  Rows without entries belong to
    no cluster."""
    if not 1 <= int(num_hashes) <= MAX_HASHES:
        raise ValueError(f"num_hashes must be in 1..{MAX_HASHES}")
    if int(max_cluster_size) < 1:
        raise ValueError("max_cluster_size must be positive")
    ei, n = _edges(edge_index, num_nodes)
    if n == 0:
        return []
    dev = torch.device(device)
    hip = _lib.hip()
    with structure.Adjacency(ei[0], ei[1], torch.ones(ei.shape[1]), n) as adj, torch.cuda.device(dev):
        ptr = adj.ptr.to(dev)
        col = adj.col.to(dev) if adj.nnz else torch.zeros(1, dtype=torch.int32, device=dev)
        stream = current_stream_handle(dev)
        counts = torch.empty(n, dtype=torch.int32, device=dev)
        col_ids = torch.arange(n, dtype=torch.int64, device=dev)
        stack = [torch.nonzero(ptr.diff() > 0).view(-1).to(torch.int32)]
        closed = []
        while stack:
            rows = stack.pop()
            m = int(rows.numel())
            _ok(hip.cbm_cluster_column_counts(n, n, ptr.data_ptr(), col.data_ptr(), rows.data_ptr(), m,
                                              counts.data_ptr(), stream), "column counts")
            # the hash columns: |2 degree - m| smallest, ties to the smaller column
            key = (2 * counts.to(torch.int64) - m).abs() * n + col_ids
            k = min(int(num_hashes), n)
            hashes = torch.sort(torch.topk(key, k, largest=False).values % n).values.to(torch.int32)
            fp = torch.empty(m, dtype=torch.int32, device=dev)
            _ok(hip.cbm_cluster_fingerprints(n, ptr.data_ptr(), col.data_ptr(), rows.data_ptr(), m,
                                             hashes.data_ptr(), k, fp.data_ptr(), stream), "fingerprints")
            order = torch.argsort(fp.to(torch.int64) * n + rows.to(torch.int64))
            sizes = torch.unique_consecutive(fp[order], return_counts=True)[1].tolist()
            groups = torch.split(rows[order], sizes)
            if len(groups) == 1 and m > max_cluster_size:
                closed.append(groups[0])  # no column splits these rows
                continue
            for g in groups:
                (closed if g.numel() <= max_cluster_size else stack).append(g)
        return [g.to("cpu", torch.int64) for g in closed]


def merge_clusters(clusters, max_cluster_size=100000):
    """Two-pointer merge of the size-sorted clusters (clustering_gpu.py:153-176):
    the largest open cluster absorbs the smallest ones while it stays within
    ``max_cluster_size``."""
    if not clusters:
        return []
    cs = sorted(clusters, key=lambda c: int(c.numel()))
    merged = []
    b, e = 0, len(cs) - 1
    ec = cs[e]
    while b < e:
        bc = cs[b]
        if int(bc.numel()) + int(ec.numel()) <= max_cluster_size:
            ec = torch.cat((bc, ec))
            b += 1
        else:
            merged.append(ec)
            e -= 1
            ec = cs[e]
    merged.append(ec)
    return merged


def cluster_compression(edge_index, clusters, alpha=0, *, num_nodes=None, device="cuda"):
    """nnz of A, of its CBM, and of one CBM per cluster of rows
    (clustering_gpu.py:178-215): each cluster's rows with all their entries
    form a matrix of their own (N = its largest index + 1, as the reference's
    cbm4mm infers it), built with the candidate scan on ``device``."""
    ei, n = _edges(edge_index, num_nodes)
    ones = torch.ones(ei.shape[1])
    whole = structure.build(ei[0], ei[1], ones, n, alpha, candidates_on=device).stats["nnz_delta"]
    per = []
    for c in clusters:
        mask = torch.isin(ei[0], torch.as_tensor(c, dtype=torch.int32))
        sub = ei[:, mask]
        if sub.shape[1] == 0:
            per.append(0)
            continue
        n_c = int(sub.max()) + 1
        per.append(structure.build(sub[0], sub[1], torch.ones(sub.shape[1]), n_c, alpha,
                                   candidates_on=device).stats["nnz_delta"])
    total = int(ei.shape[1])
    clustered = sum(per)
    return {"nnz": total, "nnz_delta": int(whole), "nnz_delta_per_cluster": per, "nnz_delta_clusters": clustered,
            "reduction_vs_csr": total / max(1, clustered), "reduction_vs_cbm": int(whole) / max(1, clustered),
            "cbm_reduction_vs_csr": total / max(1, int(whole))}
