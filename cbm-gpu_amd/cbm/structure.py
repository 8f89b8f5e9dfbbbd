"""Host-side CBM structure: synthetic graphs, the ``init`` build and the
device plan, as CPU torch tensors.  Thin wrappers over libcbm_host.so."""
import ctypes
from dataclasses import dataclass, field

import torch

from . import _lib

# Array and statistic slots of include/cbm_host.h.
(_B_DELTA_ROW, _B_DELTA_COL, _B_DELTA_VAL, _B_BRANCHES, _B_TREE_SRC, _B_TREE_DST, _B_PARENT, _B_CSR_PTR, _B_CSR_COL,
 _B_CSR_VAL) = range(10)
_STAT_NAMES = ("n", "nnz_a", "nnz_delta", "candidates", "msa_weight", "tree_edges", "branches", "depth", "symmetric")
_PLAN_NAMES = ("lvl_ptr", "lvl_node", "lvl_parent", "fwd_ptr", "fwd_col", "bwd_ptr", "bwd_row", "ch_ptr",
               "ch_node", "lvl_nlong", "fck_ptr", "fck_pos", "fck_meta", "fck_slot", "fred_ptr", "fred_pos",
               "fred_slot", "fred_cnt", "bwd_node", "bck_pos", "bck_meta", "bck_slot", "bred_pos", "bred_slot",
               "bred_cnt", "rev_ptr", "rev_node", "bwd_mid", "ch_mid", "lvl_ppos", "fck_gslot", "fred_gslot", "rev_nlong", "rck_ptr", "rck_pos", "rck_meta", "rck_slot",
               "rred_ptr", "rred_pos", "rred_slot", "rred_cnt", "fpc_b", "fpc_e", "fpc_slot", "fpc_range", "frs_slot",
               "frs_cnt", "bpc_b", "bpc_e", "bpc_mid", "bpc_slot", "bpc_range", "brs_slot", "brs_cnt",
               "bpc_row", "fpc4_b", "fpc4_e", "fpc4_slot", "fpc4_range", "frs4_slot", "frs4_cnt", "bpc4_b", "bpc4_e",
               "bpc4_mid", "bpc4_slot", "bpc4_range", "brs4_slot", "brs4_cnt", "bpc4_row", "tf_frag", "tf_seg", "tf_pos",
               "tr_frag", "tr_seg", "tr_pos", "un_ptr", "un_row", "un_lpar", "un_pc", "un_pcb", "un_pce", "un_seg_ptr",
               "un_seg", "un_of", "un_waited")
_PLAN_STATS = ("depth", "fwd_slots", "bwd_slots", "bwd_nlong", "fwd_gslots", "rev_slots", "fwd_pieces", "bwd_pieces",
               "fwd_pieces4", "bwd_pieces4", "tf_level", "tr_level", "units", "unit_levels")
# CBM_XCD_RANGES, CBM_XCD_RANGES_WIDE, CBM_PIECE, CBM_RANGE_MAX_ROWS
XCD_RANGES, XCD_RANGES_WIDE, PIECE, RANGE_MAX_ROWS = 8, 4, 256, 524288
# CBM_UNIT_ROWS, CBM_UNIT_PIECES, CBM_UNIT_PIECE, CBM_UNIT_MAX_ROWS
UNIT_ROWS, UNIT_PIECES, UNIT_PIECE, UNIT_MAX_ROWS = 64, 256, 16, 524288
PLAN_NAMES = _PLAN_NAMES


def _copy_out(ptr, n, dtype):
    t = torch.empty(int(n), dtype=dtype)
    if n > 0:
        ctypes.memmove(t.data_ptr(), ptr, int(n) * t.element_size())
    return t


def _cpu_contig(t, dtype, name):
    if not isinstance(t, torch.Tensor):
        t = torch.as_tensor(t)
    if t.dtype != dtype:
        raise TypeError(f'"{name}" is not a tensor of type "{dtype}"')
    return t.detach().to("cpu").contiguous()


def copying_model(n_nodes, n_blocks, p_copy, lam, seed_deg=0, seed=0):
    """Seeded community copying model; returns ``edge_index`` int32 ``[2, nnz]``
    (CPU), symmetric, coalesced, sorted by (row, col), without self-loops."""
    lib = _lib.host()
    g = ctypes.c_void_p()
    rc = lib.cbm_graph_copying_model(int(n_nodes), int(n_blocks), float(p_copy), float(lam), int(seed_deg),
                                     int(seed) & (2**64 - 1), ctypes.byref(g))
    _lib.check_host(rc, "copying_model")
    try:
        nnz = lib.cbm_graph_nnz(g)
        ei = torch.empty((2, nnz), dtype=torch.int32)
        if nnz:
            ctypes.memmove(ei[0].data_ptr(), lib.cbm_graph_row(g), nnz * 4)
            ctypes.memmove(ei[1].data_ptr(), lib.cbm_graph_col(g), nnz * 4)
        return ei
    finally:
        lib.cbm_graph_free(g)


@dataclass
class Build:
    """Result of the host build (the six ``init`` arrays plus extras)."""
    delta_row: torch.Tensor
    delta_col: torch.Tensor
    delta_val: torch.Tensor
    branches: torch.Tensor
    tree_src: torch.Tensor
    tree_dst: torch.Tensor
    parent: torch.Tensor
    csr_ptr: torch.Tensor
    csr_col: torch.Tensor
    csr_val: torch.Tensor
    stats: dict = field(default_factory=dict)

    def init_tuple(self):
        """The list ``cbm_mkl_cpp.init`` returns (cbm/cbm_extensions.cpp:300-344)."""
        return [self.delta_row, self.delta_col, self.delta_val, self.branches, self.tree_src, self.tree_dst]


def _read_build(lib, b):
    """Copies a cbm_build result into a :class:`Build` (CPU tensors)."""
    def arr(which, dtype):
        return _copy_out(lib.cbm_build_array(b, which), lib.cbm_build_size(b, which), dtype)
    out = Build(arr(_B_DELTA_ROW, torch.int32), arr(_B_DELTA_COL, torch.int32), arr(_B_DELTA_VAL, torch.float32),
                arr(_B_BRANCHES, torch.int32), arr(_B_TREE_SRC, torch.int32), arr(_B_TREE_DST, torch.int32),
                arr(_B_PARENT, torch.int32), arr(_B_CSR_PTR, torch.int32), arr(_B_CSR_COL, torch.int32),
                arr(_B_CSR_VAL, torch.float32))
    out.stats = {name: int(lib.cbm_build_stat(b, i)) for i, name in enumerate(_STAT_NAMES)}
    return out


def _coo(row, col, val):
    row = _cpu_contig(row, torch.int32, "src_row_idx")
    col = _cpu_contig(col, torch.int32, "src_col_idx")
    val = _cpu_contig(val, torch.float32, "src_values")
    nnz = row.numel()
    if col.numel() != nnz or val.numel() != nnz:
        raise ValueError(f"row/col/val lengths differ: {row.numel()}, {col.numel()}, {val.numel()}")
    return row, col, val


class Adjacency:
    """Binary CSR of A on the host (``cbm_adjacency_create``): rows sorted,
    duplicates merged, values all ones.  ``ptr`` int64 ``[n+1]``, ``col``
    int32 ``[nnz]`` as CPU tensors, and ``t_ptr``, ``t_col`` for A^T (the
    same tensors when ``symmetric``); the native object is released by
    :meth:`close` (or ``with``)."""

    def __init__(self, row, col, val, n_rows):
        row, col, val = _coo(row, col, val)
        self._lib = _lib.host()
        self._p = ctypes.c_void_p()
        rc = self._lib.cbm_adjacency_create(row.data_ptr(), col.data_ptr(), val.data_ptr(), row.numel(),
                                            int(n_rows), ctypes.byref(self._p))
        _lib.check_host(rc, "adjacency")
        self.n = int(self._lib.cbm_adjacency_rows(self._p))
        self.nnz = int(self._lib.cbm_adjacency_nnz(self._p))
        self.ptr = _copy_out(self._lib.cbm_adjacency_ptr(self._p), self.n + 1, torch.int64)
        self.col = _copy_out(self._lib.cbm_adjacency_col(self._p), self.nnz, torch.int32)
        self.symmetric = bool(self._lib.cbm_adjacency_symmetric(self._p))
        if self.symmetric:
            self.t_ptr, self.t_col = self.ptr, self.col
        else:
            self.t_ptr = _copy_out(self._lib.cbm_adjacency_t_ptr(self._p), self.n + 1, torch.int64)
            self.t_col = _copy_out(self._lib.cbm_adjacency_t_col(self._p), self.nnz, torch.int32)

    def close(self):
        # Also runs from __del__ at interpreter exit, when module globals
        # (ctypes among them) may already be gone: touch none of them.
        p = getattr(self, "_p", None)
        if p:
            self._p = None
            self._lib.cbm_adjacency_free(p)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        self.close()


def host_candidates(adj, alpha, n_threads=0):
    """Candidate edges of the host scan (``cbm_candidates_create``) in the
    reference creation order: ``(src, dst, w)`` int32 CPU tensors."""
    lib = _lib.host()
    c = ctypes.c_void_p()
    _lib.check_host(lib.cbm_candidates_create(adj._p, int(alpha), int(n_threads), ctypes.byref(c)), "candidates")
    try:
        m = lib.cbm_candidates_size(c)
        return tuple(_copy_out(lib.cbm_candidates_array(c, k), m, torch.int32) for k in range(3))
    finally:
        lib.cbm_candidates_free(c)


CANDIDATE_PATHS = {"auto": 0, "tiles": 1, "lists": 2}


def gpu_candidates(adj, alpha, device, path="auto", n_workgroups=0):
    """The same candidate edges from the GPU scan (``cbm_candidates_count`` /
    ``_emit``, HIP kernels) on ``device``; returned as ``(src, dst, w)`` int32
    tensors on that device.  ``path``: "auto", "tiles" (LDS counters) or
    "lists" (global counters and touched lists); same list either way."""
    dev = torch.device(device)
    if dev.type != "cuda":
        raise RuntimeError("tensor is not a CUDA tensor")
    hip = _lib.hip()
    n = adj.n
    pcode = CANDIDATE_PATHS[path]
    with torch.cuda.device(dev):
        ws_bytes = int(hip.cbm_candidates_workspace_bytes(n, pcode, int(n_workgroups)))
        if ws_bytes < 0:
            raise ValueError(f"no candidate workspace for {n} rows ({path}, {n_workgroups} workgroups)")
        stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
        ptr = adj.ptr.to(dev)
        col = adj.col.to(dev)
        if col.numel() == 0:
            col = torch.zeros(1, dtype=torch.int32, device=dev)  # a valid pointer for an empty matrix
        if adj.symmetric:
            t_ptr, t_col = ptr, col
        else:
            t_ptr = adj.t_ptr.to(dev)
            t_col = adj.t_col.to(dev) if adj.nnz else col
        ws = torch.empty(ws_bytes + 256, dtype=torch.uint8, device=dev)
        ws_ptr = (ws.data_ptr() + 255) // 256 * 256
        counts = torch.empty(n, dtype=torch.int64, device=dev)
        _lib.check_candidates(hip.cbm_candidates_count(n, ptr.data_ptr(), col.data_ptr(), t_ptr.data_ptr(),
                                                       t_col.data_ptr(), int(alpha), pcode,
                                                       int(n_workgroups), counts.data_ptr(), ws_ptr, ws_bytes,
                                                       stream), "candidates (count)")
        offsets = torch.cumsum(counts, 0) - counts
        total = int(counts.sum().item())
        src, dst, w = (torch.empty(max(total, 1), dtype=torch.int32, device=dev) for _ in range(3))
        _lib.check_candidates(hip.cbm_candidates_emit(n, ptr.data_ptr(), col.data_ptr(), t_ptr.data_ptr(),
                                                      t_col.data_ptr(), int(alpha), pcode,
                                                      int(n_workgroups), offsets.data_ptr(), src.data_ptr(),
                                                      dst.data_ptr(), w.data_ptr(), total, ws_ptr, ws_bytes,
                                                      stream), "candidates (emit)")
        del ws
        return src[:total], dst[:total], w[:total]


def build_from_candidates(adj, src, dst, w, n_threads=0):
    """The structure of ``adj`` from a candidate list (``cbm_build_from_candidates``)."""
    src, dst, w = (_cpu_contig(t, torch.int32, name) for t, name in ((src, "src"), (dst, "dst"), (w, "w")))
    if not (src.numel() == dst.numel() == w.numel()):
        raise ValueError("candidate arrays differ in length")
    lib = _lib.host()
    b = ctypes.c_void_p()
    rc = lib.cbm_build_from_candidates(adj._p, src.data_ptr(), dst.data_ptr(), w.data_ptr(), src.numel(),
                                       int(n_threads), ctypes.byref(b))
    _lib.check_host(rc, "init from candidates")
    try:
        return _read_build(lib, b)
    finally:
        lib.cbm_build_free(b)


def build(row, col, val, n_rows, alpha, n_threads=0, candidates_on=None):
    """Builds the CBM structure of the binary matrix given as COO.

    ``candidates_on``: None scans the candidate edges on the host; a CUDA
    device scans them there with the HIP kernels.  The arborescence and Delta
    are built on the host either way, and both scans give the same list, so
    the result is the same."""
    if candidates_on is not None:
        with Adjacency(row, col, val, n_rows) as adj:
            src, dst, w = gpu_candidates(adj, alpha, candidates_on)
            return build_from_candidates(adj, src, dst, w, n_threads)
    row, col, val = _coo(row, col, val)
    lib = _lib.host()
    b = ctypes.c_void_p()
    rc = lib.cbm_build_create(row.data_ptr(), col.data_ptr(), val.data_ptr(), row.numel(), int(n_rows), int(alpha),
                              int(n_threads), ctypes.byref(b))
    _lib.check_host(rc, "init")
    try:
        return _read_build(lib, b)
    finally:
        lib.cbm_build_free(b)


@dataclass
class Plan:
    """Device layout of a CBM structure (see csrc/host/plan.cpp)."""
    n: int
    depth: int
    arrays: dict
    stats: dict = field(default_factory=dict)

    def __getattr__(self, name):
        arrays = self.__dict__.get("arrays")
        if arrays is not None and name in arrays:
            return arrays[name]
        raise AttributeError(name)


def plan(n_rows, crow, col, val, tree_src, tree_dst, n_threads=0):
    """Lays out Delta (CSR with +-scale values) and the tree for the kernels."""
    crow = _cpu_contig(crow, torch.int32, "crow_indices")
    col = _cpu_contig(col, torch.int32, "col_indices")
    val = _cpu_contig(val, torch.float32, "values")
    tree_src = _cpu_contig(tree_src, torch.int32, "mca_src_idx")
    tree_dst = _cpu_contig(tree_dst, torch.int32, "mca_dst_idx")
    n_rows = int(n_rows)
    if crow.numel() != n_rows + 1:
        raise ValueError(f"crow_indices has {crow.numel()} entries, expected {n_rows + 1}")
    if tree_src.numel() != tree_dst.numel():
        raise ValueError("mca_src_idx and mca_dst_idx differ in length")
    lib = _lib.host()
    p = ctypes.c_void_p()
    rc = lib.cbm_plan_create(n_rows, crow.data_ptr(), col.data_ptr(), val.data_ptr(), col.numel(),
                             tree_src.data_ptr(), tree_dst.data_ptr(), tree_src.numel(), int(n_threads),
                             ctypes.byref(p))
    _lib.check_host(rc, "plan")
    try:
        arrays = {name: _copy_out(lib.cbm_plan_array(p, i), lib.cbm_plan_size(p, i), torch.int32)
                  for i, name in enumerate(_PLAN_NAMES)}
        stats = {name: int(lib.cbm_plan_stat(p, i)) for i, name in enumerate(_PLAN_STATS)}
        return Plan(n=n_rows, depth=int(lib.cbm_plan_depth(p)), arrays=arrays, stats=stats)
    finally:
        lib.cbm_plan_free(p)


# Gather order of the Delta pass (cbm_partners, cbm_gather_order).
LONG_ROW = 256        # CBM_LONG_ROW
# Columns held by more rows than this are not counted by the partner scan
# (each adds one to all of them and separates little, at a quadratic cost).
PARTNER_CAP = 256
# Partners per row the forest is built from.
PARTNERS = 4


def device_partners(fwd_ptr, fwd_col, device, cap=PARTNER_CAP, k=PARTNERS, n_workgroups=0):
    """The k (1 to 4) best-overlapping partner rows of every short row of a
    forward row list (``cbm_partners``, HIP): ``(partner, overlap)``, int32
    ``[n, k]`` on ``device``; -1 / 0 where there is none.  Long rows are left
    out of the column lists, so they are nobody's partner."""
    dev = torch.device(device)
    if dev.type != "cuda":
        raise RuntimeError("tensor is not a CUDA tensor")
    hip = _lib.hip()
    with torch.cuda.device(dev):
        ptr = fwd_ptr.to(dev, torch.int32).contiguous()
        n = int(ptr.numel()) - 1
        col = fwd_col.to(dev, torch.int32).contiguous()
        lens = ptr.diff()
        rows = torch.repeat_interleave(torch.arange(n, dtype=torch.int32, device=dev), lens.long())
        keep = (lens <= LONG_ROW)[rows.long()]
        cols = (col[keep] & 0x7fffffff).long()
        order = torch.argsort(cols, stable=True)
        t_col = rows[keep][order].contiguous()
        t_ptr = torch.zeros(n + 1, dtype=torch.int32, device=dev)
        t_ptr[1:] = torch.cumsum(torch.bincount(cols, minlength=n), 0).to(torch.int32)
        del rows, keep, cols, order
        if col.numel() == 0:
            col = torch.zeros(1, dtype=torch.int32, device=dev)
        if t_col.numel() == 0:
            t_col = torch.zeros(1, dtype=torch.int32, device=dev)
        ws_bytes = int(hip.cbm_partners_workspace_bytes(n, int(n_workgroups)))
        if ws_bytes < 0:
            raise ValueError(f"no partner workspace for {n} rows")
        ws = torch.empty(ws_bytes + 256, dtype=torch.uint8, device=dev)
        ws_ptr = (ws.data_ptr() + 255) // 256 * 256
        partner = torch.empty((n, int(k)), dtype=torch.int32, device=dev)
        overlap = torch.empty((n, int(k)), dtype=torch.int32, device=dev)
        stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
        _lib.check_candidates(hip.cbm_partners(n, ptr.data_ptr(), col.data_ptr(), n, t_ptr.data_ptr(),
                                               t_col.data_ptr(), int(cap), int(k), int(n_workgroups), partner.data_ptr(),
                                               overlap.data_ptr(), ws_ptr, ws_bytes, stream), "partners")
        torch.cuda.current_stream(dev).synchronize()
        del ws
        return partner, overlap


def gather_order(fwd_ptr, partner, overlap):
    """The short rows of a forward row list in the gather order
    (``cbm_gather_order``): int32 positions, rows that share columns next to
    one another, from each row's partner rows (``[n, k]`` int32, -1 for none)."""
    ptr = _cpu_contig(fwd_ptr, torch.int32, "fwd_ptr")
    n = int(ptr.numel()) - 1
    partner = _cpu_contig(partner, torch.int32, "partner").reshape(n, -1)
    overlap = _cpu_contig(overlap, torch.int32, "overlap").reshape(n, -1)
    if partner.shape != overlap.shape:
        raise ValueError("partner and overlap differ in shape")
    out = torch.empty(max(n, 1), dtype=torch.int32)
    k = ctypes.c_int32()
    lib = _lib.host()
    rc = lib.cbm_gather_order(n, ptr.data_ptr(), partner.data_ptr(), overlap.data_ptr(), int(partner.shape[1]),
                              out.data_ptr(), ctypes.byref(k))
    _lib.check_host(rc, "gather order")
    return out[:k.value].clone()


def device_gather_order(fwd_ptr, fwd_col, device, cap=PARTNER_CAP, k=PARTNERS):
    """The gather order of a forward row list held on ``device``: the partner
    scan there, the forest walk on the host."""
    partner, overlap = device_partners(fwd_ptr, fwd_col, device, cap=cap, k=k)
    return gather_order(fwd_ptr.cpu(), partner.cpu(), overlap.cpu())
