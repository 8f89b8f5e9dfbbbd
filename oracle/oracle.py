"""CPU oracle of the CBM product -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this module, and only as the checker; the product never imports it.

* Structure (cbm/cbm_extensions.cpp:60-345) through the C restatement in
  cbm_oracle.c: candidate edges in the reference creation order, the
  minimum arborescence by Chu-Liu/Edmonds, Delta and the branch arrays.
* Products in FP64 with numpy/scipy: ``A X``, ``A D X``, ``D A D X`` and their
  transposes straight from the edge list (cbm/csr4*.py semantics), and the
  CBM form ``Delta' X`` followed by the reference propagation of
  cbm/cbm.cu:25-91 (branch by branch, in order, ``(y * D_dst) / D_src``).
"""
import ctypes
import os

import numpy as np
import scipy.sparse as sp

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


def build_lib():
    """Compiles cbm_oracle.c (gcc) into oracle/_build/liboracle.so."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build_lib()
        L = ctypes.CDLL(LIB_PATH)
        L.oracle_candidates.restype = ctypes.c_int64
        L.oracle_candidates.argtypes = [ctypes.c_int32, _i64p, _i32p, ctypes.c_int32, ctypes.c_int64, _i32p, _i32p,
                                        _i32p]
        L.oracle_msa.restype = ctypes.c_int64
        L.oracle_msa.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int64, _i32p, _i32p, _i32p, _i32p]
        L.oracle_delta.restype = ctypes.c_int64
        L.oracle_delta.argtypes = [ctypes.c_int32, _i64p, _i32p, _i32p, _i32p, _i32p, _f32p, _i32p, _i32p,
                                   ctypes.POINTER(ctypes.c_int32), _i32p, ctypes.POINTER(ctypes.c_int32)]
        L.oracle_update.restype = None
        L.oracle_update.argtypes = [ctypes.c_int32, _f64p, ctypes.c_int32, _i32p, _i32p, _i32p, ctypes.c_void_p]
        _lib = L
    return _lib


# ----------------------------------------------------------------- matrix

def csr_of(edge_index, n):
    """Binary CSR (scipy, FP64 ones, sorted unique columns) of an edge list."""
    ei = np.asarray(edge_index, dtype=np.int64)
    a = sp.coo_matrix((np.ones(ei.shape[1]), (ei[0], ei[1])), shape=(n, n)).tocsr()
    a.sum_duplicates()
    a.data[:] = 1.0
    a.sort_indices()
    return a


def csr_of_sorted(edge_index, n):
    """The same binary CSR for an edge list already sorted by (row, col)
    without duplicates (what the copying-model generator returns), built
    without the COO sort: checked, then indptr by counting.  Raises
    ValueError for input that is not sorted and unique."""
    ei = np.asarray(edge_index, dtype=np.int64)
    key = ei[0] * n + ei[1]
    if key.size > 1 and not bool(np.all(key[1:] > key[:-1])):
        raise ValueError("edge list is not sorted by (row, col) without duplicates")
    indptr = np.zeros(n + 1, np.int64)
    np.cumsum(np.bincount(ei[0], minlength=n), out=indptr[1:])
    return sp.csr_matrix((np.ones(ei.shape[1]), ei[1].astype(np.int32), indptr), shape=(n, n))


def degree_scale(a):
    """D^-1/2 from integer row degrees in FP64; 0 for degree-0 rows."""
    deg = np.diff(a.indptr).astype(np.float64)
    out = np.zeros_like(deg)
    out[deg > 0] = deg[deg > 0] ** -0.5
    return out


# -------------------------------------------------------------- structure

def candidates(a, alpha):
    """Candidate edges (src, dst, w) in creation order (cbm_extensions.cpp:122-182)."""
    n = a.shape[0]
    indptr = np.ascontiguousarray(a.indptr, dtype=np.int64)
    indices = np.ascontiguousarray(a.indices, dtype=np.int32)
    cap = 4 * n + 16
    while True:
        s, d, w = (np.empty(cap, np.int32) for _ in range(3))
        m = lib().oracle_candidates(n, indptr, indices, int(alpha), cap, s, d, w)
        if m < 0:
            raise MemoryError("oracle_candidates")
        if m <= cap:
            return s[:m].copy(), d[:m].copy(), w[:m].copy()
        cap = int(m)


def msa(n_vertices, root, src, dst, w):
    """Chu-Liu/Edmonds minimum arborescence; returns (weight, in_edge)."""
    in_edge = np.empty(n_vertices, np.int32)
    weight = lib().oracle_msa(n_vertices, root, len(src), np.ascontiguousarray(src, np.int32),
                              np.ascontiguousarray(dst, np.int32), np.ascontiguousarray(w, np.int32), in_edge)
    if weight < 0:
        raise ValueError("no spanning arborescence")
    return int(weight), in_edge


def delta_for_parent(a, parent):
    """Delta COO in construction order plus the tree/branch arrays for a
    given parent array (-1 = child of the virtual root)."""
    n = a.shape[0]
    indptr = np.ascontiguousarray(a.indptr, dtype=np.int64)
    indices = np.ascontiguousarray(a.indices, dtype=np.int32)
    cap = 2 * int(a.nnz) + 1
    drow, dcol = np.empty(cap, np.int32), np.empty(cap, np.int32)
    dval = np.empty(cap, np.float32)
    bsrc, bdst = np.empty(max(n, 1), np.int32), np.empty(max(n, 1), np.int32)
    branches = np.empty(n + 1, np.int32)
    ne, nb = ctypes.c_int32(0), ctypes.c_int32(0)
    nd = lib().oracle_delta(n, indptr, indices, np.ascontiguousarray(parent, np.int32), drow, dcol, dval, bsrc, bdst,
                            ctypes.byref(ne), branches, ctypes.byref(nb))
    return {"delta_row": drow[:nd].copy(), "delta_col": dcol[:nd].copy(), "delta_val": dval[:nd].copy(),
            "branches": branches[:nb.value + 1].copy(), "tree_src": bsrc[:ne.value].copy(),
            "tree_dst": bdst[:ne.value].copy()}


def cbm_init(edge_index, n, alpha):
    """Full restatement of ``init``: candidates, Edmonds, Delta.  Returns a dict
    with the six init arrays, the parent array and the arborescence weight."""
    a = csr_of(edge_index, n)
    s, d, w = candidates(a, alpha)
    weight, in_edge = msa(n + 1, n, s, d, w)
    parent = s[in_edge[:n]].astype(np.int32)
    parent[parent == n] = -1
    out = delta_for_parent(a, parent)
    out.update(parent=parent, weight=weight, candidates=(s, d, w))
    return out


def delta_matrix(delta_row, delta_col, delta_val, n):
    return sp.csr_matrix((np.asarray(delta_val, np.float64), (delta_row, delta_col)), shape=(n, n))


def parent_matrix(parent, n):
    """P with P[v, parent(v)] = 1, so that A = (I - P)^-1 Delta."""
    v = np.nonzero(np.asarray(parent) >= 0)[0]
    return sp.csr_matrix((np.ones(len(v)), (v, np.asarray(parent)[v])), shape=(n, n))


# ---------------------------------------------------------------- products

MODES = ("ax", "adx", "dadx")


def product_matrix(a, mode):
    """A, A D or D A D (D = D^-1/2) as an FP64 scipy matrix (cbm/csr4*.py)."""
    if mode == "ax":
        return a.astype(np.float64)
    dg = sp.diags(degree_scale(a))
    if mode == "adx":
        return (a @ dg).tocsr()
    if mode == "dadx":
        return (dg @ a @ dg).tocsr()
    raise ValueError(mode)


def forward(a, x, mode="ax"):
    """M X in FP64."""
    return product_matrix(a, mode) @ np.asarray(x, np.float64)


def adjoint(a, g, mode="ax"):
    """M^T G in FP64."""
    return product_matrix(a, mode).T @ np.asarray(g, np.float64)


def cbm_forward(a, structure, x, mode="ax"):
    """The reference GPU path in FP64: Y = Delta' X (cbm/cbm.cu:184-189 /
    :244-249), then ``update`` / ``fused_update`` branch by branch
    (cbm/cbm.cu:25-91) with Delta' = Delta, Delta D or D Delta D
    (cbm/cbm4ad.py:25, cbm/cbm4dad.py:25-34)."""
    n = a.shape[0]
    delta = delta_matrix(structure["delta_row"], structure["delta_col"], structure["delta_val"], n)
    mult = None
    if mode != "ax":
        dvec = degree_scale(a)
        delta = delta @ sp.diags(dvec)
        if mode == "dadx":
            delta = sp.diags(dvec) @ delta
            mult = np.ascontiguousarray(dvec)
    y = np.ascontiguousarray(delta @ np.asarray(x, np.float64))
    f = y.shape[1]
    br = np.ascontiguousarray(structure["branches"], np.int32)
    lib().oracle_update(f, y, len(br) - 1, br, np.ascontiguousarray(structure["tree_src"], np.int32),
                        np.ascontiguousarray(structure["tree_dst"], np.int32),
                        mult.ctypes.data_as(ctypes.c_void_p) if mult is not None else None)
    return y


def max_rel_error(y, ref):
    """max |y - ref| / max |ref| (normwise relative error)."""
    ref = np.asarray(ref, np.float64)
    scale = float(np.max(np.abs(ref))) if ref.size else 0.0
    err = float(np.max(np.abs(np.asarray(y, np.float64) - ref))) if ref.size else 0.0
    return err / scale if scale > 0 else err
