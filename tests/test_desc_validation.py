"""cbm_create's checks of a layout descriptor (include/cbm_hip.h): a complete
descriptor is accepted, and damaged level tables, range piece tables, modes
or missing arrays are refused with CBM_ERR_ARG and a message, before any
device work.  The plain product (mode AX) allocates and launches nothing at
creation, so these run on the CPU: the device-array fields hold host
addresses that cbm_create only checks for null (CPU only)."""
import ctypes

import pytest
import torch

from cbm import _lib, structure
from cbm._hip_sigs import DEVICE_ARRAYS, HOST_ARRAYS, CbmDesc
from hubgraph import HUB_N, hub_graph

MODE_AX, MODE_DADX = 0, 2


@pytest.fixture(scope="module")
def plan():
    ei = hub_graph()
    b = structure.build(ei[0], ei[1], torch.ones(ei.shape[1]), HUB_N, 0)
    pl = structure.plan(HUB_N, b.csr_ptr, b.csr_col, b.csr_val, b.tree_src, b.tree_dst)
    assert pl.stats["fwd_pieces"] > 0 and pl.stats["fwd_pieces4"] > 0
    return pl


def _desc(pl, host=None, **over):
    """A descriptor over the plan's host arrays (the device fields point at
    host memory, never read at creation in mode AX)."""
    host = {k: pl.arrays[k].clone().contiguous() for k in HOST_ARRAYS} if host is None else host
    keep = dict(host)
    fields = {k: (pl.arrays[k].data_ptr() if pl.arrays[k].numel() else None) for k in DEVICE_ARRAYS}
    fields.update({k + "_host": v.data_ptr() for k, v in host.items()})
    st = pl.stats
    kw = dict(n_rows=pl.n, n_levels=pl.depth, nnz=int(pl.fwd_col.numel()), n_edges=int(pl.ch_node.numel()),
              mode=MODE_AX, fwd_slots=st["fwd_slots"], bwd_nlong=st["bwd_nlong"], n_bck=int(pl.bck_pos.numel()),
              n_bred=int(pl.bred_pos.numel()), bwd_slots=st["bwd_slots"], fwd_gslots=st["fwd_gslots"],
              rev_slots=st["rev_slots"], fwd_pieces=st["fwd_pieces"], bwd_pieces=st["bwd_pieces"],
              fwd_pieces4=st["fwd_pieces4"], bwd_pieces4=st["bwd_pieces4"], tf_level=st["tf_level"],
              tr_level=st["tr_level"], tf_frags=max(0, pl.tf_frag.numel() - 1),
              tr_frags=max(0, pl.tr_frag.numel() - 1), d=None, **fields)
    kw.update(over)
    return CbmDesc(**kw), keep


def _create(desc):
    lib = _lib.hip()
    h = ctypes.c_void_p()
    rc = lib.cbm_create(ctypes.byref(desc), None, ctypes.byref(h))
    msg = lib.cbm_last_error().decode()
    if rc == _lib.CBM_OK:
        assert h.value
        assert lib.cbm_destroy(h) == _lib.CBM_OK
    return rc, msg


def test_a_complete_descriptor_is_accepted(plan):
    desc, _ = _desc(plan)
    rc, msg = _create(desc)
    assert rc == _lib.CBM_OK, msg


def test_damaged_range_tables_are_refused(plan):
    for name in ("fpc_range", "bpc_range", "fpc4_range", "bpc4_range"):
        host = {k: plan.arrays[k].clone().contiguous() for k in HOST_ARRAYS}
        host[name][-1] -= 1  # offsets no longer end at the piece count
        desc, keep = _desc(plan, host)
        rc, msg = _create(desc)
        assert rc == _lib.CBM_ERR_ARG and "range piece offsets" in msg, name
    for field in ("fpc4_b", "bpc4_row", "bpc_mid", "frs_cnt"):
        desc, keep = _desc(plan, **{field: None})
        rc, msg = _create(desc)
        assert rc == _lib.CBM_ERR_ARG and "incomplete range piece tables" in msg, field
    desc, keep = _desc(plan, fwd_pieces4=-1)
    assert _create(desc)[0] == _lib.CBM_ERR_ARG


def test_a_plan_without_pieces_is_accepted(plan):
    """Piece counts of 0 with null tables: the layout of a matrix above the
    range size limit."""
    over = {k: None for k in DEVICE_ARRAYS if k[:3] in ("fpc", "frs", "bpc", "brs")}
    host = {k: plan.arrays[k].clone().contiguous() for k in HOST_ARRAYS}
    desc, keep = _desc(plan, host, fwd_pieces=0, bwd_pieces=0, fwd_pieces4=0, bwd_pieces4=0, **over)
    for k in ("fpc_range", "bpc_range", "fpc4_range", "bpc4_range"):
        setattr(desc, k + "_host", None)
    assert _create(desc)[0] == _lib.CBM_OK


def test_damaged_level_tables_and_modes_are_refused(plan):
    host = {k: plan.arrays[k].clone().contiguous() for k in HOST_ARRAYS}
    host["lvl_ptr"][-1] += 1  # does not end at n_rows
    desc, keep = _desc(plan, host)
    rc, msg = _create(desc)
    assert rc == _lib.CBM_ERR_ARG and "level tables" in msg
    host = {k: plan.arrays[k].clone().contiguous() for k in HOST_ARRAYS}
    host["rev_nlong"][0] = int(host["rev_ptr"][1] - host["rev_ptr"][0]) + 1  # more long lists than nodes
    desc, keep = _desc(plan, host)
    rc, msg = _create(desc)
    assert rc == _lib.CBM_ERR_ARG and "inconsistent level tables" in msg
    desc, keep = _desc(plan, mode=7)
    rc, msg = _create(desc)
    assert rc == _lib.CBM_ERR_ARG and "unknown mode" in msg
    desc, keep = _desc(plan, mode=MODE_DADX)  # D A D needs the D^-1/2 vector
    rc, msg = _create(desc)
    assert rc == _lib.CBM_ERR_ARG and "D^-1/2" in msg
    desc, keep = _desc(plan, n_edges=plan.n)  # a forest over n nodes has fewer than n edges
    rc, msg = _create(desc)
    assert rc == _lib.CBM_ERR_ARG and "incomplete descriptor" in msg


def test_damaged_tree_tails_are_refused(plan):
    L = plan.depth
    assert plan.stats["tf_level"] < L or plan.stats["tr_level"] < L  # the hub graph has a tail
    for over in (dict(tf_level=0), dict(tf_level=L + 1), dict(tr_level=-1), dict(tr_level=L + 1)):
        desc, keep = _desc(plan, **over)
        rc, msg = _create(desc)
        assert rc == _lib.CBM_ERR_ARG and "tree tail" in msg, over
    if plan.stats["tr_level"] < L:
        desc, keep = _desc(plan, tr_frags=0)
        assert _create(desc)[0] == _lib.CBM_ERR_ARG
        desc, keep = _desc(plan, tr_pos=None)
        assert _create(desc)[0] == _lib.CBM_ERR_ARG
    # no tail at all (level = depth, no tables) is a valid layout
    desc, keep = _desc(plan, tf_level=L, tr_level=L, tf_frags=0, tr_frags=0, tf_frag=None, tf_seg=None, tf_pos=None,
                       tr_frag=None, tr_seg=None, tr_pos=None)
    assert _create(desc)[0] == _lib.CBM_OK
