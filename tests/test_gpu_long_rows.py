"""Rows longer than one workgroup chunk: the chunked long-row kernels and
their in-order reduction, forward and adjoint, on a dense graph with hubs
(rows up to ~2.4k entries, several multi-chunk rows in both directions).
These tests exercise the multi-launch forms, so the handles here run with
the one-launch forward off and the adjoint through the Delta^T pass
(test_gpu_fused.py covers those two)."""
import ctypes

import numpy as np
import pytest
import torch

from cbm.cbm4ad import cbm4ad
from cbm.cbm4dad import cbm4dad
from cbm.cbm4mm import cbm4mm
from hubgraph import HUB_N, hub_graph
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CLASSES = {"ax": cbm4mm, "adx": cbm4ad, "dadx": cbm4dad}


def _multi_launch(m):
    """The multi-launch forward and the Delta^T adjoint on this matrix."""
    m.cusparse_extension.set_fused(False)
    m.cusparse_extension.set_symmetric(False)
    return m


@pytest.fixture(scope="module")
def hub(cuda):
    ei = hub_graph()
    a = O.csr_of(ei.numpy(), HUB_N)
    mats = {mode: _multi_launch(cls(ei, torch.ones(ei.shape[1]), alpha=0, device=cuda))
            for mode, cls in CLASSES.items()}
    st = mats["ax"].cusparse_extension.plan_stats
    assert st["fwd_slots"] > 0 and st["bwd_slots"] > 0  # multi-chunk rows in both directions
    return a, mats


def _x(shape, seed):
    return np.random.default_rng(seed).standard_normal(shape).astype(np.float32)


@pytest.mark.parametrize("mode", ["ax", "adx", "dadx"])
@pytest.mark.parametrize("f", [1, 64, 256, 300])
def test_long_rows_fp32(hub, cuda, mode, f):
    a, mats = hub
    m = mats[mode]
    x = _x((HUB_N, f), f)
    y = torch.empty(HUB_N, f, device=cuda)
    m.matmul(torch.from_numpy(x).to(cuda), y)
    assert O.max_rel_error(y.cpu().numpy(), O.forward(a, x, mode)) <= 1e-5
    dx = torch.empty(HUB_N, f, device=cuda)
    m.matmul_t(torch.from_numpy(x).to(cuda), dx)
    assert O.max_rel_error(dx.cpu().numpy(), O.adjoint(a, x, mode)) <= 1e-5


@pytest.mark.parametrize("mode", ["ax", "dadx"])
def test_long_rows_bf16(hub, cuda, mode):
    a, mats = hub
    m = mats[mode]
    xb = torch.from_numpy(_x((HUB_N, 128), 5)).to(cuda).to(torch.bfloat16)
    xe = xb.float().cpu().numpy()
    y = torch.empty(HUB_N, 128, device=cuda)
    m.matmul(xb, y)
    assert O.max_rel_error(y.cpu().numpy(), O.forward(a, xe, mode)) <= 1e-5
    dx = torch.empty(HUB_N, 128, device=cuda)
    m.matmul_t(xb, dx)
    assert O.max_rel_error(dx.cpu().numpy(), O.adjoint(a, xe, mode)) <= 1e-5


def test_long_rows_integer_exact_and_deterministic(hub, cuda):
    a, mats = hub
    h = mats["ax"].cusparse_extension
    x = np.random.default_rng(3).integers(-3, 4, size=(HUB_N, 32)).astype(np.float32)
    xt = torch.from_numpy(x).to(cuda)
    y1 = torch.empty(HUB_N, 32, device=cuda)
    y2 = torch.empty(HUB_N, 32, device=cuda)
    h.forward(xt, y1)
    h.forward(xt, y2)
    assert np.array_equal(y1.cpu().numpy().astype(np.float64), O.forward(a, x, "ax"))
    assert torch.equal(y1, y2)
    d1 = torch.empty(HUB_N, 32, device=cuda)
    h.matmul_t(xt, d1)
    assert np.array_equal(d1.cpu().numpy().astype(np.float64), O.adjoint(a, x, "ax"))


@pytest.mark.parametrize("f", [16, 64, 256])
def test_every_row_long(cuda, f):
    """A complete graph on 300 nodes with alpha above every degree: no tree
    edge passes the rule, so Delta = A and every row (299 entries) is longer
    than the short-row limit in both directions.  The gather passes then have
    chunks and no short rows; with integer inputs both storage types and both
    grid forms give the oracle's values exactly."""
    n = 300
    r, c = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    keep = r != c
    ei = torch.from_numpy(np.stack([r[keep], c[keep]]).astype(np.int32))
    a = O.csr_of(ei.numpy(), n)
    h = _multi_launch(cbm4mm(ei, torch.ones(ei.shape[1]), alpha=n, device=cuda)).cusparse_extension
    assert h.plan_stats["depth"] == 1
    x = np.random.default_rng(17).integers(-3, 4, size=(n, f)).astype(np.float32)
    want_y, want_dx = O.forward(a, x, "ax"), O.adjoint(a, x, "ax")
    for dtype in (torch.float32, torch.bfloat16):
        xt = torch.from_numpy(x).to(cuda).to(dtype)
        for grid in (True, False, None):
            h.set_mixed_grid(grid)
            y = torch.full((n, f), float("nan"), device=cuda)
            dx = torch.full((n, f), float("nan"), device=cuda)
            h.forward(xt, y)
            h.matmul_t(xt, dx)
            assert np.array_equal(y.cpu().numpy().astype(np.float64), want_y), (dtype, grid)
            assert np.array_equal(dx.cpu().numpy().astype(np.float64), want_dx), (dtype, grid)
    h.set_mixed_grid(None)


@pytest.mark.parametrize("mode", ["ax", "dadx"])
@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16])
@pytest.mark.parametrize("f", [24, 64, 256])
def test_one_grid_and_two_launches_are_bit_identical(hub, cuda, mode, dtype, f):
    """The long-row chunks and the short rows of a pass run as one grid or as
    two launches (chunks first); every workgroup does the same work either
    way, so forward, Delta half and adjoint give the same bits in both forms,
    for every row mapping (a wavefront, 16 lanes or the row's own lanes per
    row) and both forward forms.  At F = 256 the rows take
    64 lanes and 1024-thread workgroups."""
    a, mats = hub
    h = mats[mode].cusparse_extension
    xt = torch.from_numpy(_x((HUB_N, f), 31 + f)).to(cuda).to(dtype)
    for wave, split in ((None, None), (True, True), (16, True), (False, True), (None, False)):
        h.set_wave_rows(wave)
        h.set_forward_split(split)
        res = {}
        for grid in (True, False):
            h.set_mixed_grid(grid)
            y = torch.full((HUB_N, f), float("nan"), device=cuda)
            yd = torch.full((HUB_N, f), float("nan"), device=cuda)
            dx = torch.full((HUB_N, f), float("nan"), device=cuda)
            h.forward(xt, y)
            h.delta_spmm(xt, yd)
            h.matmul_t(xt, dx)
            res[grid] = (y, yd, dx)
        for k in range(3):
            assert torch.equal(res[True][k], res[False][k]), (mode, f, wave, split, k)
        assert O.max_rel_error(res[True][0].cpu().numpy(), O.forward(a, xt.float().cpu().numpy(), mode)) <= 1e-5
        assert O.max_rel_error(res[True][2].cpu().numpy(), O.adjoint(a, xt.float().cpu().numpy(), mode)) <= 1e-5
    assert h._lib.cbm_set_mixed_grid(h._h, 2) != 0  # out of range: rejected
    h.set_mixed_grid(None)
    h.set_forward_split(None)
    h.set_wave_rows(None)


@pytest.mark.parametrize("mode", ["ax", "adx", "dadx"])
@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16])
@pytest.mark.parametrize("f", [8, 64, 256, 300])
def test_chunk_finish_in_pass_and_by_reduction_are_bit_identical(hub, cuda, mode, dtype, f):
    """Rows of several long-row chunks are finished either by their last
    chunk inside the gather pass (arrival counters) or by a reduction launch
    after it; both add the chunk sums in chunk order, so the forward, the
    Delta half, the adjoint's reverse and Delta^T passes give the same bits,
    call after call (the counters are left at zero by every launch), and
    the results match the oracle.  F = 300 takes two column passes per row."""
    a, mats = hub
    h = mats[mode].cusparse_extension
    xt = torch.from_numpy(_x((HUB_N, f), 41 + f)).to(cuda).to(dtype)
    res = {}
    for finish in (True, False, True):
        h.set_chunk_finish(finish)
        out = []
        for _ in range(2):
            y = torch.full((HUB_N, f), float("nan"), device=cuda)
            yd = torch.full((HUB_N, f), float("nan"), device=cuda)
            dx = torch.full((HUB_N, f), float("nan"), device=cuda)
            h.forward(xt, y)
            h.delta_spmm(xt, yd)
            h.matmul_t(xt, dx)
            out.append((y, yd, dx))
        for k in range(3):
            assert torch.equal(out[0][k], out[1][k]), (mode, f, finish, k)
        if finish in res:
            for k in range(3):
                assert torch.equal(res[finish][k], out[0][k]), (mode, f, finish, k)
        res[finish] = out[0]
    for k in range(3):
        assert torch.equal(res[True][k], res[False][k]), (mode, f, k)
    xf = xt.float().cpu().numpy()
    assert O.max_rel_error(res[True][0].cpu().numpy(), O.forward(a, xf, mode)) <= 1e-5
    assert O.max_rel_error(res[True][2].cpu().numpy(), O.adjoint(a, xf, mode)) <= 1e-5
    assert h._lib.cbm_set_chunk_finish(h._h, 2) != 0  # out of range: rejected
    h.set_chunk_finish(None)


@pytest.mark.parametrize("mode", ["ax", "adx", "dadx"])
@pytest.mark.parametrize("dtype,f", [(torch.float32, 1), (torch.float32, 3), (torch.float32, 4), (torch.float32, 8),
                                     (torch.float32, 64), (torch.bfloat16, 16), (torch.bfloat16, 128)])
def test_split_and_fused_forward_are_bit_identical(hub, cuda, mode, dtype, f):
    """The split forward (one Delta pass over all levels with chunk slots
    numbered over the levels, then per-level propagation) and the fused
    per-level forward give the same bits, plain and compensated, and
    delta_spmm + propagate gives them too.  The hub graph has multi-chunk
    rows at several levels, so the global slot numbering is exercised."""
    a, mats = hub
    h = mats[mode].cusparse_extension
    xt = torch.from_numpy(_x((HUB_N, f), 11 + f)).to(cuda).to(dtype)
    for ranges, wave, depth in ((False, False, None), (False, False, 10), (False, True, None), (False, True, 10),
                                (False, None, 1), (False, None, 0), (True, None, None), (True, None, 1)):
        h.set_xcd_ranges(ranges)
        h.set_wave_rows(wave)
        h.set_kahan_depth(depth)
        ys = []
        for split in (False, True):
            h.set_forward_split(split)
            y = torch.full((HUB_N, f), float("nan"), device=cuda)
            h.forward(xt, y)
            ys.append(y)
        h.set_forward_split(None)
        y3 = torch.full((HUB_N, f), float("nan"), device=cuda)
        h.delta_spmm(xt, y3)
        h.propagate(y3)
        assert torch.equal(ys[0], ys[1]), (depth, f)
        assert torch.equal(ys[0], y3), (depth, f)
        assert O.max_rel_error(ys[1].cpu().numpy(), O.forward(a, xt.float().cpu().numpy(), mode)) <= 1e-5
    h.set_kahan_depth(32)
    h.set_wave_rows(None)
    h.set_xcd_ranges(None)


def test_compensation_from_level_zero_equals_level_one(hub, cuda):
    """Level 0 has no parent term, so compensation from level 0 and from
    level 1 are the same computation; the level-0 rows carry no compensation
    row to read (a parent at level 0 contributes none)."""
    _, mats = hub
    h = mats["ax"].cusparse_extension
    xt = torch.from_numpy(_x((HUB_N, 32), 21)).to(cuda)
    out = []
    for depth in (1, 0):
        h.set_kahan_depth(depth)
        for split in (False, True):
            h.set_forward_split(split)
            y = torch.empty(HUB_N, 32, device=cuda)
            h.forward(xt, y)
            out.append(y)
    h.set_kahan_depth(32)
    h.set_forward_split(None)
    assert all(torch.equal(out[0], o) for o in out[1:])


@pytest.fixture(scope="module")
def star(cuda):
    from hubgraph import star_graph
    ei, n = star_graph()
    a = O.csr_of(ei.numpy(), n)
    mats = {mode: _multi_launch(cls(ei, torch.ones(ei.shape[1]), alpha=0, device=cuda))
            for mode, cls in CLASSES.items()}
    assert mats["ax"].cusparse_extension.plan_stats["rev_slots"] >= 3  # a multi-chunk child list
    return a, n, mats


@pytest.mark.parametrize("mode", ["ax", "adx", "dadx"])
@pytest.mark.parametrize("dtype,f", [(torch.float32, 1), (torch.float32, 64), (torch.float32, 260),
                                     (torch.bfloat16, 16), (torch.bfloat16, 256)])
def test_chunked_child_lists_in_the_reverse_pass(star, cuda, mode, dtype, f):
    """The adjoint over a node with 2500 children (its child list split into
    chunks and reduced in order) against the FP64 oracle, forward too."""
    a, n, mats = star
    m = mats[mode]
    xt = torch.from_numpy(_x((n, f), 40 + f)).to(cuda).to(dtype)
    xe = xt.float().cpu().numpy()
    dx = torch.empty(n, f, device=cuda)
    m.matmul_t(xt, dx)
    assert O.max_rel_error(dx.cpu().numpy(), O.adjoint(a, xe, mode)) <= 1e-5
    y = torch.empty(n, f, device=cuda)
    m.matmul(xt, y)
    assert O.max_rel_error(y.cpu().numpy(), O.forward(a, xe, mode)) <= 1e-5


def test_chunked_child_lists_integer_exact(star, cuda):
    a, n, mats = star
    h = mats["ax"].cusparse_extension
    x = np.random.default_rng(9).integers(-3, 4, size=(n, 24)).astype(np.float32)
    d1 = torch.empty(n, 24, device=cuda)
    d2 = torch.empty(n, 24, device=cuda)
    h.matmul_t(torch.from_numpy(x).to(cuda), d1)
    h.matmul_t(torch.from_numpy(x).to(cuda), d2)
    assert np.array_equal(d1.cpu().numpy().astype(np.float64), O.adjoint(a, x, "ax"))
    assert torch.equal(d1, d2)


@pytest.mark.parametrize("wave", [False, True])
@pytest.mark.parametrize("mode", ["ax", "dadx"])
def test_row_mappings_against_the_oracle(hub, star, cuda, wave, mode):
    """Both short-row mappings (rows packed per wavefront, a wavefront per row
    with its entries in sub-group slices) on the hub and star graphs, forward
    and adjoint, FP32 and BF16, with integer inputs bitwise exact."""
    for a, n, m in ((hub[0], HUB_N, hub[1][mode]), (star[0], star[1], star[2][mode])):
        h = m.cusparse_extension
        h.set_wave_rows(wave)
        for dtype, f in ((torch.float32, 8), (torch.float32, 64), (torch.bfloat16, 16), (torch.bfloat16, 128)):
            xt = torch.from_numpy(_x((n, f), 60 + f)).to(cuda).to(dtype)
            xe = xt.float().cpu().numpy()
            y = torch.empty(n, f, device=cuda)
            dx = torch.empty(n, f, device=cuda)
            h.forward(xt, y)
            h.matmul_t(xt, dx)
            assert O.max_rel_error(y.cpu().numpy(), O.forward(a, xe, mode)) <= 1e-5, (wave, dtype, f)
            assert O.max_rel_error(dx.cpu().numpy(), O.adjoint(a, xe, mode)) <= 1e-5, (wave, dtype, f)
        if mode == "ax":
            x = np.random.default_rng(5).integers(-3, 4, size=(n, 16)).astype(np.float32)
            y = torch.empty(n, 16, device=cuda)
            dx = torch.empty(n, 16, device=cuda)
            h.forward(torch.from_numpy(x).to(cuda), y)
            h.matmul_t(torch.from_numpy(x).to(cuda), dx)
            assert np.array_equal(y.cpu().numpy().astype(np.float64), O.forward(a, x, "ax"))
            assert np.array_equal(dx.cpu().numpy().astype(np.float64), O.adjoint(a, x, "ax"))
        h.set_wave_rows(None)


@pytest.mark.parametrize("mode", ["ax", "adx", "dadx"])
@pytest.mark.parametrize("dtype,f", [(torch.float32, 1), (torch.float32, 8), (torch.float32, 300),
                                     (torch.float32, 512), (torch.bfloat16, 64), (torch.bfloat16, 256),
                                     (torch.bfloat16, 520)])
def test_xcd_slabs_give_the_same_bits(hub, cuda, mode, dtype, f):
    """Feature slabs spread over the XCDs change which workgroup computes which
    columns, not the order in which a row's entries are added: forward (both
    forms) and adjoint are bit-identical to the unslabbed passes.  Narrow rows
    get fewer slabs than asked (F = 1, 8, 64 none; 256 BF16 two; 300 FP32,
    520 BF16 four; 512 FP32 eight) and long rows take the chunk kernel."""
    a, mats = hub
    h = mats[mode].cusparse_extension
    h.set_wave_rows(False)
    xt = torch.from_numpy(_x((HUB_N, f), 80 + f)).to(cuda).to(dtype)
    res = {}
    for slabs in (1, 8, 3):
        h.set_xcd_slabs(slabs)
        for split in (False, True):
            h.set_forward_split(split)
            y = torch.full((HUB_N, f), float("nan"), device=cuda)
            h.forward(xt, y)
            res[(slabs, split)] = y
        dx = torch.full((HUB_N, f), float("nan"), device=cuda)
        h.matmul_t(xt, dx)
        res[(slabs, "t")] = dx
    h.set_xcd_slabs(1)
    h.set_forward_split(None)
    h.set_wave_rows(None)
    for k, v in res.items():
        assert torch.equal(v, res[(1, k[1])]), k
    xe = xt.float().cpu().numpy()
    assert O.max_rel_error(res[(8, True)].cpu().numpy(), O.forward(a, xe, mode)) <= 1e-5
    assert O.max_rel_error(res[(8, "t")].cpu().numpy(), O.adjoint(a, xe, mode)) <= 1e-5


@pytest.mark.parametrize("mode", ["ax", "adx", "dadx"])
@pytest.mark.parametrize("dtype,f", [(torch.float32, 1), (torch.float32, 8), (torch.float32, 64),
                                     (torch.float32, 300), (torch.bfloat16, 20), (torch.bfloat16, 64),
                                     (torch.bfloat16, 256)])
def test_xcd_ranges_against_the_oracle(hub, star, cuda, mode, dtype, f):
    """Gathers by XCD column range (one partial row per range piece, reduced
    in order, fused with the level propagation), with the 8-range set (one
    range per XCD) and the 4-range set (each range on two XCDs, one column
    half each, the second half empty for F = 1): forward and adjoint against
    FP64 on the hub graph (hub rows of several pieces per range, deep tree)
    and the star graph (a 2500-child list); repeat calls give the same bits;
    with ranges off the results are within rounding of them."""
    for a, n, m in ((hub[0], HUB_N, hub[1][mode]), (star[0], star[1], star[2][mode])):
        h = m.cusparse_extension
        for k in ("fwd_pieces", "bwd_pieces", "fwd_pieces4", "bwd_pieces4"):
            assert h.plan_stats[k] > 0, k
        xt = torch.from_numpy(_x((n, f), 90 + f)).to(cuda).to(dtype)
        xe = xt.float().cpu().numpy()
        out = {}
        for ranges in (8, 4, False, 8, 4):
            h.set_xcd_ranges(ranges)
            assert h.xcd_ranges_active(dtype, f) == (ranges or 0)
            y = torch.full((n, f), float("nan"), device=cuda)
            dx = torch.full((n, f), float("nan"), device=cuda)
            h.forward(xt, y)
            h.matmul_t(xt, dx)
            if ranges in out:
                assert torch.equal(y, out[ranges][0]) and torch.equal(dx, out[ranges][1])
            out[ranges] = (y, dx)
        # column phases change which workgroup sums which columns, not the sums
        for ranges in (8, 4):
            h.set_xcd_ranges(ranges)
            for phases in (2, 4):
                h.set_xcd_phases(phases)
                y = torch.full((n, f), float("nan"), device=cuda)
                dx = torch.full((n, f), float("nan"), device=cuda)
                h.forward(xt, y)
                h.matmul_t(xt, dx)
                assert torch.equal(y, out[ranges][0]) and torch.equal(dx, out[ranges][1]), (ranges, phases)
            h.set_xcd_phases(None)
        h.set_xcd_ranges(None)
        for ranges in (8, 4):
            y, dx = out[ranges]
            assert O.max_rel_error(y.cpu().numpy(), O.forward(a, xe, mode)) <= 1e-5, (n, mode, f, ranges)
            assert O.max_rel_error(dx.cpu().numpy(), O.adjoint(a, xe, mode)) <= 1e-5, (n, mode, f, ranges)
            assert O.max_rel_error(y.cpu().numpy(), out[False][0].cpu().numpy()) <= 1e-5


def test_xcd_ranges_integer_exact_and_halves(hub, cuda):
    """Integer-valued inputs are exact by range too (both piece sets);
    delta_spmm + propagate give the product's bits, plain and compensated."""
    a, mats = hub
    h = mats["ax"].cusparse_extension
    x = np.random.default_rng(12).integers(-3, 4, size=(HUB_N, 24)).astype(np.float32)
    xt = torch.from_numpy(x).to(cuda)
    for ranges in (8, 4):
        h.set_xcd_ranges(ranges)
        y = torch.empty(HUB_N, 24, device=cuda)
        dx = torch.empty(HUB_N, 24, device=cuda)
        h.forward(xt, y)
        h.matmul_t(xt, dx)
        assert np.array_equal(y.cpu().numpy().astype(np.float64), O.forward(a, x, "ax")), ranges
        assert np.array_equal(dx.cpu().numpy().astype(np.float64), O.adjoint(a, x, "ax")), ranges
    for ranges in (8, 4):
        for depth in (None, 2):
            for name, m in mats.items():
                hm = m.cusparse_extension
                hm.set_xcd_ranges(ranges)
                hm.set_kahan_depth(depth)
                xf = torch.from_numpy(_x((HUB_N, 40), 7)).to(cuda)
                y1 = torch.empty(HUB_N, 40, device=cuda)
                y2 = torch.empty(HUB_N, 40, device=cuda)
                hm.forward(xf, y1)
                hm.delta_spmm(xf, y2)
                hm.propagate(y2)
                assert torch.equal(y1, y2), (name, depth, ranges)
                hm.set_kahan_depth(32)
                hm.set_xcd_ranges(None)
    h.set_kahan_depth(32)
    h.set_xcd_ranges(None)


def test_xcd_ranges_active_reports_the_choice(hub, cuda):
    """The query follows the setting: forced sets, forced on (the 4-range set
    once half a row is 256 bytes: F x element size >= 512) and off for both
    storage types; automatic off at this size (a gathered matrix under
    24 MiB), and the automatic choice is the one the forward takes (its
    bits: a range pass sums a row range by range, so on these random inputs
    it rounds differently from the row pass)."""
    _, mats = hub
    h = mats["ax"].cusparse_extension
    for mode, f, want32, want16 in ((True, 64, 8, 8), (True, 128, 4, 8), (True, 256, 4, 4), (8, 256, 8, 8),
                                    (4, 64, 4, 4), (False, 256, 0, 0), (None, 256, 0, 0)):
        h.set_xcd_ranges(mode)
        assert h.xcd_ranges_active(torch.float32, f) == want32, (mode, f)
        assert h.xcd_ranges_active(torch.bfloat16, f) == want16, (mode, f)
    with pytest.raises(ValueError):
        h.set_xcd_ranges(2)
    with pytest.raises(Exception):
        h.set_xcd_phases(3)
    x = torch.randn(HUB_N, 64, device=cuda, generator=torch.Generator(device=cuda).manual_seed(11))
    ys = {}
    for mode in (True, False, None):
        h.set_xcd_ranges(mode)
        ys[mode] = torch.empty(HUB_N, 64, device=cuda)
        h.forward(x, ys[mode])
    assert torch.equal(ys[None], ys[False])
    assert not torch.equal(ys[None], ys[True])
    with pytest.raises(Exception):
        h.xcd_ranges_active(torch.float16, 64)
    h.set_xcd_ranges(None)


@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16])
def test_xcd_ranges_with_column_slabs(hub, cuda, dtype):
    """Column slabs over the range path: each slab's pieces write partial rows
    of the slab's width and the reduction adds them per slab; forward and
    adjoint stay within FP32 rounding of FP64, and the halves of the forward
    still give the product's bits."""
    a, mats = hub
    for mode, m in mats.items():
        h = m.cusparse_extension
        h.set_xcd_ranges(True)
        elem = 2 if dtype == torch.bfloat16 else 4
        h.set_slab_bytes(HUB_N * 32 * elem)  # 96 columns in slabs of 32
        xt = torch.from_numpy(_x((HUB_N, 96), 41)).to(cuda).to(dtype)
        xe = xt.float().cpu().numpy()
        y = torch.full((HUB_N, 96), float("nan"), device=cuda)
        dx = torch.full((HUB_N, 96), float("nan"), device=cuda)
        h.forward(xt, y)
        h.matmul_t(xt, dx)
        y2 = torch.full((HUB_N, 96), float("nan"), device=cuda)
        h.delta_spmm(xt, y2)
        h.propagate(y2)
        h.set_slab_bytes(0)
        h.set_xcd_ranges(None)
        assert O.max_rel_error(y.cpu().numpy(), O.forward(a, xe, mode)) <= 1e-5, mode
        assert O.max_rel_error(dx.cpu().numpy(), O.adjoint(a, xe, mode)) <= 1e-5, mode
        assert torch.equal(y, y2), mode


@pytest.mark.parametrize("dtype,f", [(torch.float32, 1), (torch.float32, 64), (torch.float32, 300),
                                     (torch.bfloat16, 64), (torch.bfloat16, 520)])
def test_tree_tails(hub, cuda, dtype, f):
    """The deepest levels in one launch (one workgroup per fragment): the
    forward gives the per-level launches' bits, plain and compensated (Kahan
    from level 2, inside the tail), with the deep fragments staged in LDS
    (the default), none staged, or every one staged; the reverse pass (each
    row summed in entry order) stays within FP32 rounding of them and of
    FP64, and integer-valued inputs are exact."""
    a, mats = hub
    for mode, m in mats.items():
        h = m.cusparse_extension
        assert h.plan_stats["tf_level"] < h.depth and h.plan_stats["tr_level"] < h.depth
        xt = torch.from_numpy(_x((HUB_N, f), 60 + f)).to(cuda).to(dtype)
        xe = xt.float().cpu().numpy()
        for depth in (None, 2):
            h.set_kahan_depth(depth)
            res = {}
            for on in (True, "unstaged", "all_staged", False):
                h.set_tree_tails(on)
                y = torch.full((HUB_N, f), float("nan"), device=cuda)
                dx = torch.full((HUB_N, f), float("nan"), device=cuda)
                h.forward(xt, y)
                launches = h.last_launches()
                h.matmul_t(xt, dx)
                res[on] = (y, dx, launches)
            for on in (True, "unstaged", "all_staged"):
                assert torch.equal(res[on][0], res[False][0]), (mode, depth, on)
                assert res[on][2] < res[False][2]
            frag, seg = h._dev["tf_frag"].cpu().numpy(), h._dev["tf_seg"].cpu().numpy()
            rows = seg[frag[1:]] - seg[frag[:-1]]
            fit = int((rows <= 452).sum())  # 452 rows of 16 columns (+ compensation, 16 bytes each) fit 64 KiB
            assert 0 < fit
            # the staged fragments take a launch of their own, the others one more
            assert res["all_staged"][2] == res["unstaged"][2] + (1 if fit < len(rows) else 0)
            assert O.max_rel_error(res[True][1].cpu().numpy(), res[False][1].cpu().numpy()) <= 1e-6, (mode, depth)
            assert O.max_rel_error(res[True][0].cpu().numpy(), O.forward(a, xe, mode)) <= 1e-5, (mode, depth)
            assert O.max_rel_error(res[True][1].cpu().numpy(), O.adjoint(a, xe, mode)) <= 1e-5, (mode, depth)
        h.set_kahan_depth(32)
        h.set_tree_tails(True)
    h = mats["ax"].cusparse_extension
    xi = np.random.default_rng(5).integers(-3, 4, size=(HUB_N, 24)).astype(np.float32)
    xt = torch.from_numpy(xi).to(cuda)
    dx = torch.empty(HUB_N, 24, device=cuda)
    h.matmul_t(xt, dx)
    assert np.array_equal(dx.cpu().numpy().astype(np.float64), O.adjoint(a, xi, "ax"))


@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16])
def test_chunk_finish_replays_from_a_graph(hub, cuda, dtype):
    """The in-pass finish of multi-chunk rows leaves its arrival counters at
    zero after every launch, so a HIP graph of the forward and the adjoint
    replays correctly again and again (no memset node in the graph): each
    replay, with eager calls and synchronisations between replays, gives the
    eager bits."""
    a, mats = hub
    h = mats["dadx"].cusparse_extension
    h.set_chunk_finish(True)
    f = 64
    h.reserve(f)
    x = torch.from_numpy(_x((HUB_N, f), 91)).to(cuda).to(dtype)
    y = torch.empty(HUB_N, f, device=cuda)
    dx = torch.empty(HUB_N, f, device=cuda)
    z = torch.empty(HUB_N, f, device=cuda)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        h.forward(x, y)
        h.matmul_t(x, dx, workspace=z)
    torch.cuda.current_stream().wait_stream(s)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        h.forward(x, y)
        h.matmul_t(x, dx, workspace=z)
    for it in range(4):
        x.copy_(torch.from_numpy(_x((HUB_N, f), 92 + it)).to(cuda).to(dtype))
        y_ref = torch.empty_like(y)
        dx_ref = torch.empty_like(dx)
        h.forward(x, y_ref)
        h.matmul_t(x, dx_ref)
        torch.cuda.synchronize()
        y.fill_(float("nan"))
        dx.fill_(float("nan"))
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(y, y_ref), it
        assert torch.equal(dx, dx_ref), it
    xf = x.float().cpu().numpy()
    assert O.max_rel_error(y.cpu().numpy(), O.forward(a, xf, "dadx")) <= 1e-5
    assert O.max_rel_error(dx.cpu().numpy(), O.adjoint(a, xf, "dadx")) <= 1e-5
    h.set_chunk_finish(None)


def test_small_matrix_tables_and_plan_tables_agree(cuda):
    """Small matrices run the forward Delta pass with long-row tables of their
    own (rows above 64 entries in 256-entry chunks), built at the first
    reservation; before it, the allocation-free entry point runs the plan's
    tables (rows above 256 entries in 1024-entry chunks).  Integer inputs give
    the exact product either way; random inputs agree to FP32 rounding."""
    ei = hub_graph()
    a = O.csr_of(ei.numpy(), HUB_N)
    h = _multi_launch(cbm4mm(ei, torch.ones(ei.shape[1]), alpha=0, device=cuda)).cusparse_extension
    lens = np.diff(h._dev["fwd_ptr"].cpu().numpy())
    assert ((lens > 64) & (lens <= 256)).any() and (lens > 256).any()
    assert h._reserved == 0  # nothing has reserved this handle yet: the plan's tables serve
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def noalloc(x, y):
        code = 1 if x.dtype == torch.bfloat16 else 0
        return h._lib.cbm_matmul_noalloc(h._h, code, x.data_ptr(), x.stride(0), y.data_ptr(), y.stride(0),
                                         int(x.size(1)), st)

    f = 16
    xi = np.random.default_rng(5).integers(-3, 4, size=(HUB_N, f)).astype(np.float32)
    xr = _x((HUB_N, f), 6)
    plan_out = []
    for x in (xi, xr):
        y = torch.full((HUB_N, f), float("nan"), device=cuda)
        assert noalloc(torch.from_numpy(x).to(cuda), y) == 0, h._lib.cbm_last_error()
        plan_out.append(y)
    torch.cuda.synchronize()
    h.reserve(f)  # builds the small-matrix tables
    for x, y_plan in zip((xi, xr), plan_out):
        y = torch.full((HUB_N, f), float("nan"), device=cuda)
        h.forward(torch.from_numpy(x).to(cuda), y)
        if x is xi:
            assert torch.equal(y, y_plan)
            assert np.array_equal(y.cpu().numpy().astype(np.float64), O.forward(a, x, "ax"))
        else:
            assert O.max_rel_error(y.cpu().numpy(), y_plan.cpu().numpy().astype(np.float64)) <= 1e-6
            assert O.max_rel_error(y.cpu().numpy(), O.forward(a, x, "ax")) <= 1e-5
