"""Graphs from local files, in place of the reference's network loaders.

The reference's benchmarks fetch their graphs through PyG and OGB
(``benchmark/utilities.py:62-198``: SuiteSparse ``SNAP`` and ``DIMACS10``
matrices, Planetoid, OGB and TU datasets) and pair each with alpha = 32
(``:183-198``).  Nothing here touches the network: a graph is read from a file
the user already has, and comes back as the ``edge_index`` the ``cbm4*`` and
``csr4*`` classes take (int32 ``[2, nnz]`` on the CPU, rows ascending, columns
ascending within a row, no duplicates).

* :func:`read_matrix_market` -- a MatrixMarket coordinate file, the format of
  the SuiteSparse collection that ``SuiteSparseMatrixCollection`` downloads
  (``pattern`` / ``integer`` / ``real``; ``general`` / ``symmetric`` /
  ``skew-symmetric``; a symmetric file stores one triangle and is mirrored).
* :func:`read_edge_list` -- a whitespace-separated edge list, one ``u v`` pair
  per line (SNAP's ``.txt`` files; ``#`` and ``%`` lines are comments), with
  optional compaction of sparse node ids.
* :func:`read_edge_index` -- an ``.npy`` / ``.npz`` array of shape ``[2, nnz]``.
* :func:`load_dataset` -- the reference's dataset names mapped onto files
  under a local root, returning ``(edge_index, num_nodes, alpha)``.
"""
import os

import numpy as np
import torch

# The reference's dataset names (benchmark/utilities.py:183-198) and where a
# SuiteSparse download keeps them: <root>/<group>/<name>/<name>.mtx.
SUITESPARSE = {"ca-HepPh": "SNAP", "ca-AstroPh": "SNAP", "coPapersCiteseer": "DIMACS10",
               "coPapersDBLP": "DIMACS10"}
ALPHA = 32  # the alpha every dataset of benchmark/utilities.py:183-198 is run with


class GraphFileError(ValueError):
    """A graph file that this module cannot read."""


def _finish(rows, cols, num_nodes, symmetric, self_loops):
    """Coalesced edge_index (sorted by (row, col), unique) as int32 [2, nnz]."""
    rows = np.asarray(rows, dtype=np.int64)
    cols = np.asarray(cols, dtype=np.int64)
    if symmetric:
        rows, cols = np.concatenate([rows, cols]), np.concatenate([cols, rows])
    if not self_loops:
        keep = rows != cols
        rows, cols = rows[keep], cols[keep]
    if rows.size and (rows.min() < 0 or cols.min() < 0 or max(rows.max(), cols.max()) >= num_nodes):
        raise GraphFileError(f"node ids outside 0 .. {num_nodes - 1}")
    if num_nodes > np.iinfo(np.int32).max:
        raise GraphFileError(f"{num_nodes} nodes do not fit int32 indices")
    key = np.unique(rows * np.int64(num_nodes) + cols)
    out = np.stack([key // num_nodes, key % num_nodes]).astype(np.int32)
    return torch.from_numpy(out)


def _data_lines(fh):
    """The lines of an open text file that are neither blank nor comments."""
    for line in fh:
        s = line.strip()
        if s and s[0] not in "%#":
            yield s


def read_matrix_market(path, *, self_loops=True):
    """Reads a MatrixMarket coordinate file as a binary adjacency.

    Returns ``(edge_index, num_nodes)``.  Entries are 1-based in the file and
    0-based here; values are read only to drop explicit zeros; a symmetric or
    skew-symmetric file is mirrored.  The matrix must be square.
    ``self_loops=False`` drops diagonal entries."""
    with open(path, "r") as fh:
        head = fh.readline().split()
        if len(head) < 5 or head[0].lower() != "%%matrixmarket" or head[1].lower() != "matrix":
            raise GraphFileError(f"{path}: no MatrixMarket matrix header")
        fmt, field, sym = head[2].lower(), head[3].lower(), head[4].lower()
        if fmt != "coordinate":
            raise GraphFileError(f"{path}: '{fmt}' storage; only coordinate files hold a sparse graph")
        if field not in ("pattern", "integer", "real"):
            raise GraphFileError(f"{path}: '{field}' entries are not supported")
        if sym not in ("general", "symmetric", "skew-symmetric"):
            raise GraphFileError(f"{path}: '{sym}' symmetry is not supported")
        lines = _data_lines(fh)
        try:
            m, n, nnz = (int(t) for t in next(lines).split()[:3])
        except (StopIteration, ValueError) as exc:
            raise GraphFileError(f"{path}: no size line") from exc
        if m != n:
            raise GraphFileError(f"{path}: {m} x {n} is not square")
        want = 2 if field == "pattern" else 3
        data = np.loadtxt(lines, dtype=np.float64, ndmin=2, usecols=range(want)) if nnz else np.zeros((0, want))
    if data.shape[0] != nnz:
        raise GraphFileError(f"{path}: {data.shape[0]} entries, the size line says {nnz}")
    if want == 3:
        data = data[data[:, 2] != 0]
    ij = data[:, :2]
    if ij.size and (np.any(ij != np.floor(ij)) or ij.min() < 1):
        raise GraphFileError(f"{path}: entry indices must be integers from 1")
    rows, cols = ij[:, 0].astype(np.int64) - 1, ij[:, 1].astype(np.int64) - 1
    return _finish(rows, cols, n, sym != "general", self_loops), n


def read_edge_list(path, *, symmetric=True, compact=False, self_loops=True, num_nodes=None):
    """Reads ``u v`` pairs (one per line, whitespace-separated, ``#``/``%``
    comments) as a binary adjacency; further columns are ignored.

    Returns ``(edge_index, num_nodes)``.  ``symmetric`` adds every reverse
    edge (SNAP lists undirected graphs that way already; it then changes
    nothing).  ``compact`` renumbers the node ids that occur to 0 .. k-1 in
    ascending order (SNAP ids have gaps); otherwise ids are kept and
    ``num_nodes`` defaults to the largest id + 1, as the reference's classes
    infer it (cbm/cbm4mm.py:12)."""
    with open(path, "r") as fh:
        data = np.loadtxt(_data_lines(fh), dtype=np.int64, ndmin=2, usecols=(0, 1))
    if data.size == 0:
        data = np.zeros((0, 2), np.int64)
    if data.size and data.min() < 0:
        raise GraphFileError(f"{path}: negative node id")
    rows, cols = data[:, 0], data[:, 1]
    if compact:
        ids, inv = np.unique(np.concatenate([rows, cols]), return_inverse=True)
        rows, cols = inv[: rows.size], inv[rows.size:]
        n = int(ids.size)
    else:
        n = int(max(rows.max(), cols.max())) + 1 if rows.size else 0
    if num_nodes is not None:
        if int(num_nodes) < n:
            raise GraphFileError(f"{path}: num_nodes={num_nodes} but ids reach {n - 1}")
        n = int(num_nodes)
    return _finish(rows, cols, n, symmetric, self_loops), n


def read_edge_index(path, *, num_nodes=None, symmetric=False, self_loops=True):
    """Reads an ``edge_index`` saved with numpy (``.npy``, or ``.npz`` with an
    ``edge_index`` member), shape ``[2, nnz]``; returns ``(edge_index,
    num_nodes)``.  Loaded with ``allow_pickle=False``."""
    z = np.load(path, allow_pickle=False)
    if isinstance(z, np.lib.npyio.NpzFile):
        with z:
            if "edge_index" not in z.files:
                raise GraphFileError(f"{path}: no 'edge_index' member")
            ei = np.asarray(z["edge_index"])
    else:
        ei = np.asarray(z)
    if ei.ndim != 2 or ei.shape[0] != 2 or not np.issubdtype(ei.dtype, np.integer):
        raise GraphFileError(f"{path}: expected an integer array of shape [2, nnz], got {ei.dtype}{list(ei.shape)}")
    n = int(ei.max()) + 1 if ei.size else 0
    if num_nodes is not None:
        n = int(num_nodes)
    return _finish(ei[0], ei[1], n, symmetric, self_loops), n


def load_dataset(name, root):
    """The graph the reference's ``load_dataset(name)`` would download
    (benchmark/utilities.py:183-198), read from ``root`` instead:
    ``<root>/<group>/<name>/<name>.mtx`` for the SuiteSparse ones (the layout
    of the collection's archives; also accepted directly under ``root``), or
    ``<root>/<name>.npz`` / ``.npy`` holding ``edge_index`` for any name.
    Returns ``(edge_index, num_nodes, alpha)``."""
    candidates = []
    if name in SUITESPARSE:
        candidates += [os.path.join(root, SUITESPARSE[name], name, name + ".mtx"),
                       os.path.join(root, name, name + ".mtx"), os.path.join(root, name + ".mtx")]
    candidates += [os.path.join(root, name + ".npz"), os.path.join(root, name + ".npy")]
    for path in candidates:
        if os.path.exists(path):
            if path.endswith(".mtx"):
                ei, n = read_matrix_market(path)
            else:
                ei, n = read_edge_index(path)
            return ei, n, ALPHA
    raise FileNotFoundError(f"dataset {name!r}: none of {candidates} exists (no network access is attempted)")
