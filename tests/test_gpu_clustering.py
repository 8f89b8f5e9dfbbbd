"""Fingerprint row clustering on the GPU (cluster_api.hip through the C ABI
and cbm/clustering.py): the kernels' column counts and fingerprints equal
numpy's, the clusters equal the oracle restatement's (same clusters, same
order), and the per-cluster compression equals host builds of each cluster."""
import ctypes

import numpy as np
import pytest
import torch

from cbm import _lib, structure
from cbm.clustering import cluster_compression, lsh_clusters, merge_clusters
from oracle import cluster as OC
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c1():
    n = 2708
    ei = structure.copying_model(n, 7, 0.8, 0.58, 0, 1001)
    return ei, n, O.csr_of(ei.numpy(), n)


def test_kernels_equal_numpy(c1, cuda):
    ei, n, a = c1
    hip = _lib.hip()
    ptr = torch.from_numpy(a.indptr.astype(np.int64)).to(cuda)
    col = torch.from_numpy(a.indices.astype(np.int32)).to(cuda)
    rng = np.random.default_rng(3)
    rows = np.sort(rng.choice(n, 900, replace=False)).astype(np.int32)
    rows_t = torch.from_numpy(rows).to(cuda)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    counts = torch.full((n,), -5, dtype=torch.int32, device=cuda)
    assert hip.cbm_cluster_column_counts(n, n, ptr.data_ptr(), col.data_ptr(), rows_t.data_ptr(), len(rows),
                                         counts.data_ptr(), stream) == 0
    want = np.bincount(a[rows].indices, minlength=n)
    assert np.array_equal(counts.cpu().numpy(), want)
    hashes = np.array(sorted(rng.choice(n, 11, replace=False)), dtype=np.int32)
    h_t = torch.from_numpy(hashes).to(cuda)
    fp = torch.empty(len(rows), dtype=torch.int32, device=cuda)
    assert hip.cbm_cluster_fingerprints(n, ptr.data_ptr(), col.data_ptr(), rows_t.data_ptr(), len(rows),
                                        h_t.data_ptr(), len(hashes), fp.data_ptr(), stream) == 0
    dense = a[rows][:, hashes].toarray() > 0
    want_fp = (dense * (1 << np.arange(len(hashes)))).sum(axis=1)
    assert np.array_equal(fp.cpu().numpy(), want_fp)
    assert hip.cbm_cluster_fingerprints(n, ptr.data_ptr(), col.data_ptr(), rows_t.data_ptr(), len(rows),
                                        h_t.data_ptr(), 31, fp.data_ptr(), stream) != 0


@pytest.mark.parametrize("cap,hashes", [(500, 4), (100, 6), (40, 3)])
def test_clusters_equal_the_oracle(c1, cuda, cap, hashes):
    ei, n, a = c1
    got = lsh_clusters(ei, n, max_cluster_size=cap, num_hashes=hashes, device=cuda)
    want = OC.clusters(a.indptr.astype(np.int64), a.indices.astype(np.int32), n, cap, hashes)
    assert len(got) == len(want)
    for g, w in zip(got, want):
        assert g.tolist() == w.tolist()
    mg = merge_clusters(got, cap)
    mw = OC.merge(want, cap)
    assert [m.tolist() for m in mg] == [m.tolist() for m in mw]


def test_cluster_compression_equals_host_builds(c1, cuda):
    ei, n, _ = c1
    clusters = merge_clusters(lsh_clusters(ei, n, max_cluster_size=400, num_hashes=4, device=cuda), 400)
    rep = cluster_compression(ei, clusters, alpha=0, num_nodes=n, device=cuda)
    assert rep["nnz"] == ei.shape[1]
    whole = structure.build(ei[0], ei[1], torch.ones(ei.shape[1]), n, 0)
    assert rep["nnz_delta"] == whole.stats["nnz_delta"]
    for c, got in zip(clusters, rep["nnz_delta_per_cluster"]):
        sub = ei[:, torch.isin(ei[0], c.to(torch.int32))]
        b = structure.build(sub[0], sub[1], torch.ones(sub.shape[1]), int(sub.max()) + 1, 0)
        assert got == b.stats["nnz_delta"]
    assert rep["nnz_delta_clusters"] == sum(rep["nnz_delta_per_cluster"])
