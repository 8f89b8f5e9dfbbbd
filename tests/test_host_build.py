"""The product's host builder (libcbm_host.so) against the oracle, and the
device plan's invariants (CPU only)."""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp
import torch

from cbm import cbm_mkl_cpp, structure
from hubgraph import HUB_N, hub_graph, star_graph
from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "cbm_small.json")

# Workload generator parameters (bench.py) and the nnz each is meant to give.
WORKLOAD_GRAPHS = [
    (2708, 7, 0.8, 0.58, 0, 1001, 10556),
    (19717, 3, 0.85, 0.44, 0, 1002, 88648),
    (169343, 40, 0.8, 1.63, 0, 1003, 2315598),
]


def _build(ei, n, alpha):
    ei = torch.as_tensor(np.asarray(ei), dtype=torch.int32)
    return structure.build(ei[0], ei[1], torch.ones(ei.shape[1]), n, alpha)


def _same_structure(b, ref):
    for k in ("delta_row", "delta_col", "delta_val", "branches", "tree_src", "tree_dst"):
        assert np.array_equal(getattr(b, k).numpy(), ref[k]), k


def test_golden_fixtures_exactly():
    """Unique-optimum fixtures pin the parent array and every init array."""
    with open(GOLDEN) as fh:
        cases = json.load(fh)["cases"]
    for case in cases:
        n = case["n"]
        b = _build(case["edge_index"], n, case["alpha"])
        assert b.stats["candidates"] == len(case["candidates"])
        assert b.stats["msa_weight"] == case["weight"] == b.stats["nnz_delta"]
        assert b.parent.tolist() == case["parent"]
        a = O.csr_of(np.array(case["edge_index"]), n)
        _same_structure(b, O.delta_for_parent(a, b.parent.numpy()))


@pytest.mark.parametrize("alpha", [0, 1, 3])
def test_random_graphs_match_oracle(alpha):
    rng = np.random.default_rng(alpha)
    for trial in range(30):
        n = int(rng.integers(2, 120))
        dens = float(rng.uniform(0.02, 0.3))
        A = (rng.random((n, n)) < dens).astype(np.int64)
        if trial % 2:
            for i in range(1, n):  # copied rows make deeper trees
                j = int(rng.integers(0, i))
                A[i] = np.where(rng.random(n) < 0.9, A[j], A[i])
        A = np.triu(A, 1)
        A = A + A.T
        r, c = np.nonzero(A)
        if len(r) == 0:
            continue
        n_eff = int(max(r.max(), c.max())) + 1
        ei = np.stack([r, c])
        b = _build(ei, n_eff, alpha)
        ref = O.cbm_init(ei, n_eff, alpha)
        assert b.stats["candidates"] == len(ref["candidates"][0])
        assert b.stats["msa_weight"] == ref["weight"] == b.stats["nnz_delta"]
        # Delta and the branch arrays are a function of the parent array
        a = O.csr_of(ei, n_eff)
        _same_structure(b, O.delta_for_parent(a, b.parent.numpy()))
        # A = (I - P)^-1 Delta
        delta = sp.csr_matrix((b.csr_val.numpy().astype(float), b.csr_col.numpy(), b.csr_ptr.numpy()),
                              shape=(n_eff, n_eff))
        P = O.parent_matrix(b.parent.numpy(), n_eff)
        assert np.array_equal(((sp.identity(n_eff) - P) @ a).toarray(), delta.toarray())


def test_directed_graphs_match_oracle():
    """Row overlaps are |A_r & A_c| (A A^T, the reference's syrk) for a matrix
    that is not symmetric too: candidates, weight, Delta, A = (I - P)^-1 Delta."""
    rng = np.random.default_rng(11)
    for trial in range(12):
        n = int(rng.integers(5, 90))
        a = (rng.random((n, n)) < float(rng.uniform(0.03, 0.3))) & ~np.eye(n, dtype=bool)
        if trial % 2:
            for i in range(1, n):
                j = int(rng.integers(0, i))
                a[i] = np.where(rng.random(n) < 0.85, a[j], a[i])
        r, c = np.nonzero(a)
        if len(r) == 0:
            continue
        ei = np.stack([r, c])
        alpha = int(trial % 3)
        b = _build(ei, n, alpha)
        ref = O.cbm_init(ei, n, alpha)
        s_, d_, w_ = ref["candidates"]
        assert b.stats["candidates"] == len(s_)
        assert b.stats["msa_weight"] == ref["weight"] == b.stats["nnz_delta"]
        am = O.csr_of(ei, n)
        _same_structure(b, O.delta_for_parent(am, b.parent.numpy()))
        delta = sp.csr_matrix((b.csr_val.numpy().astype(float), b.csr_col.numpy(), b.csr_ptr.numpy()), shape=(n, n))
        P = O.parent_matrix(b.parent.numpy(), n)
        assert np.array_equal(((sp.identity(n) - P) @ am).toarray(), delta.toarray())


def test_workload_graph_matches_oracle():
    """The smallest workload graph: candidates, weight, parents and Delta."""
    n, nb, p, lam, sd, seed, _ = WORKLOAD_GRAPHS[0]
    ei = structure.copying_model(n, nb, p, lam, sd, seed)
    for alpha in (0, 4):
        b = _build(ei.numpy(), n, alpha)
        ref = O.cbm_init(ei.numpy(), n, alpha)
        s, d, w = ref["candidates"]
        assert b.stats["candidates"] == len(s)
        assert b.stats["msa_weight"] == ref["weight"]
        # Both implementations break ties toward the earliest candidate edge;
        # report, but do not require, parent agreement on tied instances.
        agree = float((b.parent.numpy() == ref["parent"]).mean())
        assert agree > 0.95
        _same_structure(b, O.delta_for_parent(O.csr_of(ei.numpy(), n), b.parent.numpy()))


def test_init_matches_reference_contract():
    """init(): six CPU tensors with the reference dtypes (cbm_extensions.cpp:300-344)
    and CHECK_DTYPE errors (:65-67)."""
    ei = structure.copying_model(300, 2, 0.8, 0.6, 0, 3)
    out = cbm_mkl_cpp.init(ei[0], ei[1], torch.ones(ei.shape[1]), 300, 0)
    assert len(out) == 6
    assert [t.dtype for t in out] == [torch.int32, torch.int32, torch.float32, torch.int32, torch.int32, torch.int32]
    assert all(t.device.type == "cpu" for t in out)
    assert len(out[0]) == len(out[1]) == len(out[2])
    assert len(out[4]) == len(out[5]) == int(out[3][-1])
    with pytest.raises(RuntimeError, match="int32"):
        cbm_mkl_cpp.init(ei[0].long(), ei[1], torch.ones(ei.shape[1]), 300, 0)
    with pytest.raises(RuntimeError, match="float32"):
        cbm_mkl_cpp.init(ei[0], ei[1], torch.ones(ei.shape[1], dtype=torch.float64), 300, 0)


def test_init_rejects_bad_input():
    ei = torch.tensor([[0, 1], [1, 0]], dtype=torch.int32)
    with pytest.raises(ValueError, match="binary"):
        structure.build(ei[0], ei[1], torch.tensor([1.0, 2.0]), 2, 0)
    with pytest.raises(ValueError, match="outside"):
        structure.build(ei[0], ei[1] + 5, torch.ones(2), 2, 0)


def test_duplicates_are_merged():
    ei = torch.tensor([[0, 0, 1, 1, 2], [1, 1, 0, 2, 1]], dtype=torch.int32)
    b = structure.build(ei[0], ei[1], torch.ones(5), 3, 0)
    assert b.stats["nnz_a"] == 4


@pytest.mark.parametrize("n,nb,p,lam,sd,seed,target", WORKLOAD_GRAPHS)
def test_generator_sizes_and_format(n, nb, p, lam, sd, seed, target):
    ei = structure.copying_model(n, nb, p, lam, sd, seed)
    assert abs(ei.shape[1] - target) <= 0.1 * target
    r, c = ei[0].long(), ei[1].long()
    key = r * n + c
    assert bool((key[1:] > key[:-1]).all())  # sorted, no duplicates
    assert not bool((r == c).any())  # no self-loops
    rev = torch.sort(c * n + r).values
    assert torch.equal(rev, key)  # symmetric
    assert int(max(r.max(), c.max())) == n - 1
    if n < 50000:
        assert torch.equal(structure.copying_model(n, nb, p, lam, sd, seed), ei)  # deterministic


def _plan_for(ei, n, alpha):
    b = _build(ei, n, alpha)
    return b, structure.plan(n, b.csr_ptr, b.csr_col, b.csr_val, b.tree_src, b.tree_dst)


def test_plan_invariants():
    ei = structure.copying_model(3000, 5, 0.85, 0.5, 0, 77).numpy()
    b, pl = _plan_for(ei, 3000, 0)
    n = 3000
    node = pl.lvl_node.numpy()
    par = pl.lvl_parent.numpy()
    lp = pl.lvl_ptr.numpy()
    assert sorted(node.tolist()) == list(range(n))
    level = np.empty(n, np.int64)
    for l in range(pl.depth):
        level[node[lp[l]:lp[l + 1]]] = l
    # level order is topological: a parent sits exactly one level up
    has = par >= 0
    assert np.all(level[par[has]] == level[node[has]] - 1)
    assert np.all(par[lp[0]:lp[1]] == -1)
    assert np.array_equal(par, b.parent.numpy()[node])
    # forward rows = Delta rows in level order, signs packed in bit 31
    ptr, col = b.csr_ptr.numpy(), b.csr_col.numpy()
    fp, fc = pl.fwd_ptr.numpy(), pl.fwd_col.numpy().view(np.uint32)
    vals = b.csr_val.numpy()
    for p in range(0, n, 7):
        v = node[p]
        got = fc[fp[p]:fp[p + 1]]
        assert np.array_equal(got & 0x7FFFFFFF, col[ptr[v]:ptr[v + 1]])
        assert np.array_equal((got >> 31).astype(bool), vals[ptr[v]:ptr[v + 1]] < 0)
    # backward = CSC of Delta, columns in the backward order
    bp, br = pl.bwd_ptr.numpy(), pl.bwd_row.numpy().view(np.uint32)
    bnode = pl.bwd_node.numpy()
    assert sorted(bnode.tolist()) == list(range(n))
    delta = sp.csr_matrix((vals.astype(float), col, ptr), shape=(n, n)).tocsc()
    delta.sort_indices()
    has_kids = np.bincount(b.parent.numpy()[b.parent.numpy() >= 0], minlength=n) > 0
    bmid = pl.bwd_mid.numpy()
    for p in range(0, n, 5):
        j = bnode[p]
        got = br[bp[p]:bp[p + 1]].astype(np.int64)
        rows = got & 0x7FFFFFFF
        signs = got >> 31
        want_rows = delta.indices[delta.indptr[j]:delta.indptr[j + 1]]
        want_neg = delta.data[delta.indptr[j]:delta.indptr[j + 1]] < 0
        # leaves first, then internal nodes, each part ascending
        m = bmid[p] - bp[p]
        assert not has_kids[rows[:m]].any() and has_kids[rows[m:]].all()
        assert np.all(np.diff(rows[:m]) > 0) and np.all(np.diff(rows[m:]) > 0)
        order = np.argsort(rows)
        assert np.array_equal(rows[order], want_rows)
        assert np.array_equal(signs[order].astype(bool), want_neg)
    # internal nodes by level and their children: leaf children first
    rp, rn = pl.rev_ptr.numpy(), pl.rev_node.numpy()
    assert sorted(rn.tolist()) == sorted(np.nonzero(has_kids)[0].tolist())
    for l in range(pl.depth):
        assert np.all(level[rn[rp[l]:rp[l + 1]]] == l)
    cp, cn, cm = pl.ch_ptr.numpy(), pl.ch_node.numpy(), pl.ch_mid.numpy()
    for i, u in enumerate(rn):
        kids = cn[cp[i]:cp[i + 1]]
        assert len(kids) > 0 and np.all(b.parent.numpy()[kids] == u)
        m = cm[i] - cp[i]
        assert not has_kids[kids[:m]].any() and has_kids[kids[m:]].all()
    assert cp[-1] == b.stats["tree_edges"]


def test_plan_rejects_cycles_and_double_parents():
    crow = torch.tensor([0, 1, 2, 3], dtype=torch.int32)
    col = torch.tensor([1, 2, 0], dtype=torch.int32)
    val = torch.ones(3)
    with pytest.raises(ValueError, match="cycle"):
        structure.plan(3, crow, col, val, torch.tensor([0, 1, 2], dtype=torch.int32),
                       torch.tensor([1, 2, 0], dtype=torch.int32))
    with pytest.raises(ValueError, match="more than one parent"):
        structure.plan(3, crow, col, val, torch.tensor([0, 1], dtype=torch.int32),
                       torch.tensor([2, 2], dtype=torch.int32))


def _check_split(ptr, first, n_long, ck_pos, ck_meta, ck_slot, red_pos, red_slot, red_cnt, slots, long_rows=None):
    """Chunks tile every long row exactly; slots are unique and in range.
    The long rows are [first, first + n_long) unless given as a set."""
    ptr = np.asarray(ptr, np.int64)
    lens = np.diff(ptr)
    if long_rows is None:
        long_rows = set(range(first, first + n_long))
    assert all(lens[p] > 256 for p in long_rows)
    seen = {}
    for pos, meta, slot in zip(ck_pos, ck_meta, ck_slot):
        idx, cnt = int(meta) & 0xFFFF, (int(meta) >> 16) & 0xFFFF
        assert pos in long_rows and 0 <= idx < cnt
        seen.setdefault(int(pos), []).append((idx, cnt, int(slot)))
    assert set(seen) == long_rows
    used = set()
    for pos, parts in seen.items():
        cnt = parts[0][1]
        assert sorted(i for i, _, _ in parts) == list(range(cnt))
        covered = sum(lens[pos] * (i + 1) // cnt - lens[pos] * i // cnt for i, _, _ in parts)
        assert covered == lens[pos]
        if cnt == 1:
            assert parts[0][2] == -1
        else:
            sl = sorted(s for _, _, s in parts)
            assert sl == list(range(sl[0], sl[0] + cnt)) and not (set(sl) & used)
            used |= set(sl)
            k = list(red_pos).index(pos)
            assert red_slot[k] == sl[0] and red_cnt[k] == cnt
    assert all(0 <= s < slots for s in used)


def test_plan_chunk_tables_on_long_rows():
    ei = hub_graph()
    b, pl = _plan_for(ei.numpy(), HUB_N, 0)
    st = pl.stats
    assert st["fwd_slots"] > 0 and st["bwd_slots"] > 0
    lp, nl = pl.lvl_ptr.numpy(), pl.lvl_nlong.numpy()
    fck, fred = pl.fck_ptr.numpy(), pl.fred_ptr.numpy()
    for l in range(pl.depth):
        c0, c1, r0, r1 = fck[l], fck[l + 1], fred[l], fred[l + 1]
        _check_split(pl.fwd_ptr.numpy(), lp[l], nl[l], pl.fck_pos.numpy()[c0:c1], pl.fck_meta.numpy()[c0:c1],
                     pl.fck_slot.numpy()[c0:c1], pl.fred_pos.numpy()[r0:r1], pl.fred_slot.numpy()[r0:r1],
                     pl.fred_cnt.numpy()[r0:r1], st["fwd_slots"])
        # short rows follow the long ones within the level
        lens = np.diff(pl.fwd_ptr.numpy())[lp[l]:lp[l + 1]]
        assert np.all(lens[:nl[l]] > 256) and np.all(lens[nl[l]:] <= 256)
    _check_split(pl.bwd_ptr.numpy(), 0, st["bwd_nlong"], pl.bck_pos.numpy(), pl.bck_meta.numpy(),
                 pl.bck_slot.numpy(), pl.bred_pos.numpy(), pl.bred_slot.numpy(), pl.bred_cnt.numpy(),
                 st["bwd_slots"])
    # The same forward chunks with slots numbered over all levels (one pass
    # over every level): unique over the whole table, same chunks otherwise.
    all_long = {int(p) for l in range(pl.depth) for p in range(lp[l], lp[l] + nl[l])}
    _check_split(pl.fwd_ptr.numpy(), 0, 0, pl.fck_pos.numpy(), pl.fck_meta.numpy(), pl.fck_gslot.numpy(),
                 pl.fred_pos.numpy(), pl.fred_gslot.numpy(), pl.fred_cnt.numpy(), st["fwd_gslots"],
                 long_rows=all_long)
    assert st["fwd_gslots"] == int(pl.fred_cnt.numpy().sum())
    assert np.array_equal(pl.fck_gslot.numpy() < 0, pl.fck_slot.numpy() < 0)


def test_reverse_chunk_tables_on_a_star():
    """A node with STAR_K children: the reverse pass's internal-node order puts
    long child lists first in each level, and their chunk tables tile them."""
    from hubgraph import STAR_K, star_graph
    ei, n = star_graph()
    b, pl = _plan_for(ei.numpy(), n, 0)
    parent = b.parent.numpy()
    assert np.count_nonzero(parent == 20) == STAR_K  # every copy hangs off the prototype
    st = pl.stats
    assert st["rev_slots"] >= (STAR_K + 1023) // 1024
    rp, rn = pl.rev_ptr.numpy(), pl.rev_nlong.numpy()
    cp = pl.ch_ptr.numpy()
    rck, rred = pl.rck_ptr.numpy(), pl.rred_ptr.numpy()
    for l in range(pl.depth):
        lens = np.diff(cp)[rp[l]:rp[l + 1]]
        assert np.all(lens[:rn[l]] > 256) and np.all(lens[rn[l]:] <= 256)
        c0, c1, r0, r1 = rck[l], rck[l + 1], rred[l], rred[l + 1]
        _check_split(cp, rp[l], rn[l], pl.rck_pos.numpy()[c0:c1], pl.rck_meta.numpy()[c0:c1],
                     pl.rck_slot.numpy()[c0:c1], pl.rred_pos.numpy()[r0:r1], pl.rred_slot.numpy()[r0:r1],
                     pl.rred_cnt.numpy()[r0:r1], st["rev_slots"])
    assert rn.sum() >= 1 and rred[-1] >= 1



def _check_pieces(n, ptr, idx, cut, b, e, mid, slot, rng, rs_slot, rs_cnt, order=None, internal=None,
                  n_ranges=structure.XCD_RANGES):
    """Pieces tile every row exactly, in slot order, each inside one column
    range and within the piece size; grouped by range.  Backward (cut given):
    the pieces index `order`, each row's entries rewritten range by range,
    with a piece's leaf rows before mid and its internal rows from mid."""
    width = (n + n_ranges - 1) // n_ranges
    n_pc = len(b)
    assert len(rng) == n_ranges + 1 and rng[0] == 0 and rng[-1] == n_pc
    assert np.all(np.diff(rng) >= 0)
    assert sorted(slot.tolist()) == list(range(n_pc))
    piece_of_slot = np.empty(n_pc, np.int64)
    piece_of_slot[slot] = np.arange(n_pc)
    range_of_piece = np.searchsorted(rng, np.arange(n_pc), side="right") - 1
    if cut is not None:
        for p in range(n):  # the rewritten row holds the same entries
            assert sorted(order[ptr[p]:ptr[p + 1]].tolist()) == sorted(idx[ptr[p]:ptr[p + 1]].tolist())
        idx = order
    col = (idx.view(np.uint32) & 0x7FFFFFFF).astype(np.int64)
    assert np.all(e - b >= 1) and np.all(e - b <= structure.PIECE)
    for p in range(n):
        pcs = piece_of_slot[rs_slot[p]:rs_slot[p] + rs_cnt[p]]
        if rs_cnt[p] == 0:
            assert ptr[p] == ptr[p + 1]
            continue
        covered = np.concatenate([np.arange(b[i], e[i]) for i in pcs])
        assert sorted(covered.tolist()) == list(range(ptr[p], ptr[p + 1]))
        ranges = [range_of_piece[i] for i in pcs]
        assert ranges == sorted(ranges)  # (range, part, piece) order
        for i in pcs:
            k = col[b[i]:e[i]] // width
            assert np.all(k == range_of_piece[i])
            if cut is not None:
                assert b[i] <= mid[i] <= e[i]
                assert not internal[col[b[i]:mid[i]]].any() and internal[col[mid[i]:e[i]]].all()


def test_range_pieces_tile_the_rows():
    def check(ei, n):
        b, pl = _plan_for(ei.numpy(), n, 0)
        assert pl.stats["fwd_pieces"] == len(pl.fpc_b) > 0 and pl.stats["bwd_pieces"] == len(pl.bpc_b) > 0
        par = b.parent.numpy()
        internal = np.bincount(par[par >= 0], minlength=n) > 0
        for r, sfx in ((structure.XCD_RANGES, ""), (structure.XCD_RANGES_WIDE, "4")):
            a = {k: pl.arrays[k.replace("_", sfx + "_", 1)].numpy() for k in (
                "fpc_b", "fpc_e", "fpc_slot", "fpc_range", "frs_slot", "frs_cnt", "bpc_b", "bpc_e", "bpc_mid",
                "bpc_slot", "bpc_range", "brs_slot", "brs_cnt", "bpc_row")}
            assert pl.stats["fwd_pieces" + sfx] == len(a["fpc_b"]) > 0
            assert pl.stats["bwd_pieces" + sfx] == len(a["bpc_b"]) > 0
            _check_pieces(n, pl.fwd_ptr.numpy(), pl.fwd_col.numpy(), None, a["fpc_b"], a["fpc_e"], None,
                          a["fpc_slot"], a["fpc_range"], a["frs_slot"], a["frs_cnt"], n_ranges=r)
            _check_pieces(n, pl.bwd_ptr.numpy(), pl.bwd_row.numpy(), pl.bwd_mid.numpy(), a["bpc_b"], a["bpc_e"],
                          a["bpc_mid"], a["bpc_slot"], a["bpc_range"], a["brs_slot"], a["brs_cnt"], a["bpc_row"],
                          internal, n_ranges=r)
        # fewer, wider ranges: no more pieces than the 8-range set
        assert pl.stats["fwd_pieces4"] <= pl.stats["fwd_pieces"]
        return pl

    check(hub_graph(1500), 1500)
    # the star graph's core rows hold more than one piece per range
    ei, n = star_graph()
    pl = check(ei, n)
    assert int((pl.fpc_e - pl.fpc_b).max()) == structure.PIECE or int((pl.bpc_e - pl.bpc_b).max()) == structure.PIECE


def test_no_range_pieces_above_the_size_limit():
    n = structure.RANGE_MAX_ROWS + 1
    ei = torch.tensor([[0, 1, n - 1, 0], [1, 0, 0, n - 1]], dtype=torch.int32)
    b = _build(ei.numpy(), n, 0)
    pl = structure.plan(n, b.csr_ptr, b.csr_col, b.csr_val, b.tree_src, b.tree_dst)
    assert pl.stats["fwd_pieces"] == 0 and pl.stats["bwd_pieces"] == 0
    assert pl.stats["fwd_pieces4"] == 0 and pl.stats["bwd_pieces4"] == 0
    assert pl.fpc_b.numel() == 0 and pl.fpc_range.numel() == 0
    assert pl.arrays["fpc4_b"].numel() == 0 and pl.arrays["bpc4_range"].numel() == 0


def _check_tail(pl, prefix, member_positions, parent_position, children_positions):
    """A tail's fragments partition the positions of levels >= its first
    level, one segment per level from the fragment's root down, ascending
    positions; the work of a segment depends only on its own fragment."""
    L = pl.depth
    a = pl.stats[prefix + "_level"]
    frag, seg, pos = (pl.arrays[prefix + k].numpy() for k in ("_frag", "_seg", "_pos"))
    if a >= L:
        assert frag.size == 0 and seg.size == 0 and pos.size == 0
        return None
    want = sorted(p for l in range(a, L) for p in member_positions(l))
    assert sorted(pos.tolist()) == want
    assert frag[0] == 0 and frag[-1] == seg.size - 1 and seg[0] == 0 and seg[-1] == pos.size
    where = {}  # position -> (fragment, depth in fragment)
    for f in range(frag.size - 1):
        for d, s in enumerate(range(frag[f], frag[f + 1])):
            chunk = pos[seg[s]:seg[s + 1]]
            assert chunk.size > 0 and np.all(np.diff(chunk) > 0)
            for p in chunk.tolist():
                where[p] = (f, d)
    return a, where


def test_tree_tails_are_closed_fragments():
    ei = torch.as_tensor(hub_graph(1500))
    b = _build(ei.numpy(), 1500, 0)
    for n, e in ((1500, ei),):
        pl = structure.plan(n, b.csr_ptr, b.csr_col, b.csr_val, b.tree_src, b.tree_dst)
        assert pl.stats["tf_level"] < pl.depth and pl.stats["tr_level"] < pl.depth
        lp = pl.lvl_ptr.numpy()
        ppos = pl.lvl_ppos.numpy()
        # forward: a position's parent is one segment up in its fragment, or above the tail
        a, where = _check_tail(pl, "tf", lambda l: range(lp[l], lp[l + 1]), None, None)
        for p, (f, d) in where.items():
            q = ppos[p]
            if d == 0:
                assert q < lp[a]
            else:
                assert where[q] == (f, d - 1)
        # reverse: a node's internal children are one segment down in its fragment
        rp = pl.rev_ptr.numpy()
        rnode = pl.rev_node.numpy()
        cptr, cnode, cmid = pl.ch_ptr.numpy(), pl.ch_node.numpy(), pl.ch_mid.numpy()
        rev_pos = {int(v): i for i, v in enumerate(rnode)}
        a, where = _check_tail(pl, "tr", lambda l: range(rp[l], rp[l + 1]), None, None)
        assert not pl.rev_nlong.numpy()[a:].any()
        for i, (f, d) in where.items():
            for k in range(cmid[i], cptr[i + 1]):
                child = int(cnode[k]) & 0x7FFFFFFF
                assert where[rev_pos[child]] == (f, d + 1)


def test_chain_and_star_tails():
    """A long chain goes to one fragment from level 1 (forward) or 0
    (reverse); a star has nothing worth a tail below its root."""
    n = 400
    ei = torch.tensor([list(range(n - 1)) + list(range(1, n)), list(range(1, n)) + list(range(n - 1))],
                      dtype=torch.int32)
    b = _build(ei.numpy(), n, 0)
    pl = structure.plan(n, b.csr_ptr, b.csr_col, b.csr_val, b.tree_src, b.tree_dst)
    if pl.depth > 8:
        assert pl.stats["tf_level"] == 1 and pl.tf_frag.numel() == 2
        assert pl.stats["tr_level"] < pl.depth


def test_build_reports_whether_a_is_symmetric():
    """The build's symmetric flag (exact A == A^T) is what lets the adjoint
    run through the forward structure: set for a symmetrised graph, clear
    for a directed one and for a row subset."""
    ei = structure.copying_model(500, 3, 0.8, 1.0, 0, 9)
    b = structure.build(ei[0], ei[1], torch.ones(ei.shape[1]), 500, 0)
    assert b.stats["symmetric"] == 1
    keep = ei[0] != 0  # drop row 0 only: A is no longer symmetric
    b = structure.build(ei[0][keep], ei[1][keep], torch.ones(int(keep.sum())), 500, 0)
    assert b.stats["symmetric"] == 0
    with structure.Adjacency(ei[0], ei[1], torch.ones(ei.shape[1]), 500) as adj:
        src, dst, w = structure.host_candidates(adj, 0)
        assert structure.build_from_candidates(adj, src, dst, w).stats["symmetric"] == 1


def _check_units(pl, n):
    """The unit tables: a partition of the positions into units of at most
    CBM_UNIT_ROWS rows and CBM_UNIT_PIECES pieces, each unit's rows in level
    order with level segments, local parents pointing at the parent's row, an
    outside parent only in a unit of lower index, and pieces of at most
    CBM_UNIT_PIECE entries tiling each row's Delta entries in order."""
    a = {k: pl.arrays[k].numpy() for k in ("un_ptr", "un_row", "un_lpar", "un_pc", "un_pcb", "un_pce", "un_seg_ptr",
                                            "un_seg", "un_of", "lvl_ptr", "lvl_ppos", "fwd_ptr")}
    up = a["un_ptr"]
    U = up.size - 1
    assert U == pl.stats["units"] > 0
    assert up[0] == 0 and up[-1] == n and np.all(np.diff(up) > 0)
    assert np.array_equal(np.sort(a["un_row"]), np.arange(n))
    level = np.zeros(n, np.int64)
    for l in range(pl.depth):
        level[a["lvl_ptr"][l]:a["lvl_ptr"][l + 1]] = l
    pc = a["un_pc"]
    assert pc[0] == 0 and pc[-1] == a["un_pcb"].size and np.all(np.diff(pc) >= 1)
    for u in range(U):
        r0, r1 = up[u], up[u + 1]
        rows = a["un_row"][r0:r1]
        assert r1 - r0 <= structure.UNIT_ROWS and pc[r1] - pc[r0] <= structure.UNIT_PIECES
        assert np.all(np.diff(rows) > 0) and np.all(a["un_of"][rows] == u)
        segs = a["un_seg"][a["un_seg_ptr"][u]:a["un_seg_ptr"][u + 1]]
        assert segs[0] == r0 and np.all(np.diff(segs) > 0)
        bounds = list(segs) + [r1]
        for k in range(len(segs)):
            assert np.unique(level[a["un_row"][bounds[k]:bounds[k + 1]]]).size == 1
        assert np.all(np.diff(level[a["un_row"][segs]]) > 0)
        for i in range(r0, r1):
            p = a["un_row"][i]
            pp = a["lvl_ppos"][p]
            lp = a["un_lpar"][i]
            if pp < 0:
                assert lp == -1
            elif a["un_of"][pp] == u:
                assert lp >= 0 and a["un_row"][r0 + lp] == pp
            else:
                assert lp == -2 and a["un_of"][pp] < u
            kb, ke = a["fwd_ptr"][p], a["fwd_ptr"][p + 1]
            b, e = a["un_pcb"][pc[i]:pc[i + 1]], a["un_pce"][pc[i]:pc[i + 1]]
            assert b[0] == kb and e[-1] == ke and np.all(b[1:] == e[:-1])
            assert np.all(e - b <= structure.UNIT_PIECE) and np.all(e >= b)
    return a


def _units_forward(pl, delta_rows, x):
    """The one-launch forward's arithmetic in FP64 from the unit tables alone:
    per unit, piece sums, then rows in order adding the parent's row (in the
    unit or already written by an earlier unit)."""
    a = {k: pl.arrays[k].numpy() for k in ("un_ptr", "un_row", "un_lpar", "un_pc", "un_pcb", "un_pce", "un_of",
                                            "lvl_node", "lvl_parent", "fwd_col")}
    col = a["fwd_col"].astype(np.int64)
    sign = np.where(col < 0, -1.0, 1.0)
    col = col & 0x7FFFFFFF
    y = np.full_like(x, np.nan, dtype=np.float64)
    for u in range(a["un_ptr"].size - 1):
        r0, r1 = a["un_ptr"][u], a["un_ptr"][u + 1]
        val = {}
        for i in range(r0, r1):
            s = np.zeros(x.shape[1])
            for q in range(a["un_pc"][i], a["un_pc"][i + 1]):
                e = slice(a["un_pcb"][q], a["un_pce"][q])
                s += (sign[e][:, None] * x[col[e]]).sum(axis=0)
            lp = a["un_lpar"][i]
            if lp >= 0:
                s += val[r0 + lp]
            elif lp == -2:
                par = a["lvl_parent"][a["un_row"][i]]
                assert not np.isnan(y[par]).any()  # written by an earlier unit
                s += y[par]
            val[i] = s
        for i in range(r0, r1):
            y[a["lvl_node"][a["un_row"][i]]] = val[i]
    return y


@pytest.mark.parametrize("which", ["c1", "hub", "star", "chain"])
def test_units_partition_and_compute_the_product(which):
    if which == "c1":
        n = 2708
        ei = structure.copying_model(n, 7, 0.8, 0.58, 0, 1001)
    elif which == "hub":
        n, ei = HUB_N, hub_graph()
    elif which == "star":
        ei, n = star_graph()
    else:  # a chain of depth 150: components stacked many levels deep
        m = 150
        pairs = sorted({(i, m + j) for i in range(m) for j in range(m - i)} |
                       {(m + j, i) for i in range(m) for j in range(m - i)})
        ei, n = torch.tensor(pairs, dtype=torch.int32).t().contiguous(), 2 * m
    b, pl = _plan_for(ei.numpy(), n, 0)
    a = _check_units(pl, n)
    if which == "chain":
        assert pl.stats["unit_levels"] >= 150 // structure.UNIT_ROWS
    x = np.random.default_rng(3).standard_normal((n, 5))
    y = _units_forward(pl, None, x)
    ref = O.forward(O.csr_of(ei.numpy(), n), x, "ax")
    assert np.allclose(y, ref, rtol=1e-12, atol=1e-9)


def test_no_units_when_a_row_exceeds_a_unit():
    """A Delta row of more than CBM_UNIT_PIECES pieces fits no unit: the plan
    then has none (the product runs the multi-launch forms)."""
    n = structure.UNIT_PIECES * structure.UNIT_PIECE + 10
    hub = n - 1
    r = np.concatenate([np.full(n - 1, hub), np.arange(n - 1)])
    c = np.concatenate([np.arange(n - 1), np.full(n - 1, hub)])
    ei = np.stack([r, c])
    b, pl = _plan_for(ei, n, 0)
    assert pl.stats["units"] == 0 and pl.arrays["un_ptr"].numel() == 0
