"""Graphs from local files (cbm/datasets.py): MatrixMarket and edge-list
parsing against hand-written files, the edge_index contract of the cbm4*
classes, and the structure build on a file's graph (CPU only)."""
import numpy as np
import pytest
import torch

from cbm import cbm_mkl_cpp, datasets, structure


def _write(path, text):
    path.write_text(text)
    return path


def test_symmetric_pattern_file_is_mirrored(tmp_path):
    f = _write(tmp_path / "g.mtx", "%%MatrixMarket matrix coordinate pattern symmetric\n"
               "% a comment\n4 4 3\n2 1\n3 1\n4 3\n")
    ei, n = datasets.read_matrix_market(f)
    assert n == 4
    assert ei.dtype == torch.int32
    assert ei.tolist() == [[0, 0, 1, 2, 2, 3], [1, 2, 0, 0, 3, 2]]


def test_general_valued_file_drops_zeros_and_keeps_direction(tmp_path):
    f = _write(tmp_path / "g.mtx", "%%MatrixMarket matrix coordinate real general\n"
               "3 3 4\n1 2 0.5\n2 3 0\n3 1 -2\n1 2 1\n")
    ei, n = datasets.read_matrix_market(f)
    assert (n, ei.tolist()) == (3, [[0, 2], [1, 0]])  # (2,3) is an explicit zero; (1,2) appears twice


def test_self_loops_kept_or_dropped(tmp_path):
    f = _write(tmp_path / "g.mtx", "%%MatrixMarket matrix coordinate integer symmetric\n3 3 2\n1 1 1\n3 2 7\n")
    ei, _ = datasets.read_matrix_market(f)
    assert ei.tolist() == [[0, 1, 2], [0, 2, 1]]
    ei, _ = datasets.read_matrix_market(f, self_loops=False)
    assert ei.tolist() == [[1, 2], [2, 1]]


@pytest.mark.parametrize("text,match", [
    ("%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n", "coordinate"),
    ("%%MatrixMarket matrix coordinate complex general\n2 2 1\n1 2 1 0\n", "complex"),
    ("%%MatrixMarket matrix coordinate pattern general\n2 3 1\n1 2\n", "square"),
    ("%%MatrixMarket matrix coordinate pattern general\n3 3 2\n1 2\n", "entries"),
    ("%%MatrixMarket matrix coordinate pattern general\n3 3 1\n0 2\n", "from 1"),
    ("%%MatrixMarket matrix coordinate pattern general\n3 3 1\n4 2\n", "outside"),
    ("not a header\n", "header"),
])
def test_bad_matrix_market_files_are_refused(tmp_path, text, match):
    f = _write(tmp_path / "bad.mtx", text)
    with pytest.raises(datasets.GraphFileError, match=match):
        datasets.read_matrix_market(f)


def test_snap_edge_list_with_gaps(tmp_path):
    f = _write(tmp_path / "ca.txt", "# Directed graph: toy\n# FromNodeId\tToNodeId\n10\t20\n20\t10\n20\t35\n"
               "35\t20\n10\t10\n")
    ei, n = datasets.read_edge_list(f)
    assert n == 36 and int(ei.shape[1]) == 5  # ids kept: N = max id + 1, the self-loop kept
    ei, n = datasets.read_edge_list(f, compact=True, self_loops=False)
    assert n == 3 and ei.tolist() == [[0, 1, 1, 2], [1, 0, 2, 1]]
    one_way = _write(tmp_path / "dir.txt", "0 1\n1 2\n")
    assert datasets.read_edge_list(one_way)[0].tolist() == [[0, 1, 1, 2], [1, 0, 2, 1]]
    assert datasets.read_edge_list(one_way, symmetric=False)[0].tolist() == [[0, 1], [1, 2]]
    with pytest.raises(datasets.GraphFileError, match="num_nodes"):
        datasets.read_edge_list(one_way, num_nodes=2)


def test_edge_index_arrays_and_dataset_lookup(tmp_path):
    ei = structure.copying_model(500, 3, 0.8, 0.5, 0, 9)
    np.save(tmp_path / "toy.npy", ei.numpy())
    np.savez(tmp_path / "toy2.npz", edge_index=ei.numpy().astype(np.int64))
    for name in ("toy", "toy2"):
        got, n, alpha = datasets.load_dataset(name, tmp_path)
        assert torch.equal(got, ei) and n == int(ei.max()) + 1 and alpha == 32
    with pytest.raises(FileNotFoundError, match="no network"):
        datasets.load_dataset("ca-HepPh", tmp_path)
    # the SuiteSparse layout <root>/<group>/<name>/<name>.mtx
    d = tmp_path / "SNAP" / "ca-HepPh"
    d.mkdir(parents=True)
    rows, cols = ei.numpy()
    low = rows > cols
    body = "".join(f"{r + 1} {c + 1}\n" for r, c in zip(rows[low], cols[low]))
    _write(d / "ca-HepPh.mtx", f"%%MatrixMarket matrix coordinate pattern symmetric\n{500} {500} {int(low.sum())}\n"
           + body)
    got, n, alpha = datasets.load_dataset("ca-HepPh", tmp_path)
    assert torch.equal(got, ei) and n == 500 and alpha == 32


def test_a_file_graph_builds_like_its_edge_index(tmp_path):
    """The structure built from a file's graph equals the one built from the
    same edges given directly."""
    ei = structure.copying_model(800, 4, 0.85, 0.5, 0, 21)
    text = "".join(f"{r} {c}\n" for r, c in ei.t().tolist())
    got, n = datasets.read_edge_list(_write(tmp_path / "g.txt", text), num_nodes=800)
    a = cbm_mkl_cpp.init(got[0], got[1], torch.ones(got.shape[1]), n, 4)
    b = cbm_mkl_cpp.init(ei[0], ei[1], torch.ones(ei.shape[1]), 800, 4)
    for x, y in zip(a, b):
        assert torch.equal(x, y)


def test_largest_workload_links_ignore_the_communities():
    """The largest workload's graph (N = 2,449,029, 47 blocks): the share of
    its entries whose two ends lie in the same block is that of uniformly
    random links (1/47, about 2%), so no ordering of rows or columns by block
    gives its gathers locality."""
    import bench
    w = bench.WORKLOADS["c5"]
    ei = structure.copying_model(w["n"], w["blocks"], w["p"], w["lam"], w["seed_deg"], w["seed"]).numpy()
    bsize = w["n"] // w["blocks"]
    blk = lambda v: np.minimum(v // bsize, w["blocks"] - 1)
    inside = float(np.mean(blk(ei[0]) == blk(ei[1])))
    assert ei.shape[1] > 100_000_000
    assert 0.01 < inside < 0.05, inside
