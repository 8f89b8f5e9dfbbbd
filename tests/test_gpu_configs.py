"""GPU parity at the full sizes of the benchmark workloads (the graphs and
seeds of bench.py's WORKLOADS), so that the paths that only switch on at
those sizes run in a test: the automatic XCD column ranges (N = 132,534,
about 600 entries per row), the layout without range pieces above 524,288
rows (N = 2,449,029), and compensated tree levels inside the forward tree
tail (depth 52).  The reference validates on its full datasets too
(benchmark/compare_matmul_cuda.py:41-62).

Every product runs on the whole graph; the FP64 oracle (scipy over the
edge list) checks a sample of rows: random rows, the longest rows of M and
of M^T, and the rows of the deepest tree level.  Bounds, normwise over the
sampled rows (max |y - y*| / max |y*|):
  * FP32 X and G: <= 1e-5;
  * BF16 X and G: <= 1e-5 against FP64 of the BF16-rounded inputs, and
    <= 1e-2 against FP64 of the original FP32 inputs;
  * BF16 result against the FP32 CBM result: <= 1e-2;
  * integer X with |x| <= 8 (every partial sum an integer below 2^24):
    bitwise equal to the exact product, forward and adjoint.
"""
import numpy as np
import pytest
import torch

import bench
from cbm import structure
from cbm._hip_sigs import MODE_AX
from cbm.cbm4ad import cbm4ad
from cbm.cbm4dad import cbm4dad
from cbm.cbm4mm import cbm4mm
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CLASSES = {"mm": (cbm4mm, "ax"), "ad": (cbm4ad, "adx"), "dad": (cbm4dad, "dadx")}
N_SAMPLE = 3000


def _say(msg):
    print(f"[configs] {msg}", flush=True)


class Workload:
    """One benchmark graph, its FP64 oracle matrices and sampled rows."""

    def __init__(self, name, cuda):
        w = bench.WORKLOADS[name]
        self.name, self.w, self.n = name, w, w["n"]
        _say(f"{name}: generating N={self.n}")
        self.ei = structure.copying_model(w["n"], w["blocks"], w["p"], w["lam"], w["seed_deg"], w["seed"])
        self.a = O.csr_of_sorted(self.ei.numpy(), self.n)
        self.cuda = cuda
        self._m = {}

    def oracle(self, mode):
        """(M, M^T) in FP64, CSR."""
        if mode not in self._m:
            m = O.product_matrix(self.a, mode).tocsr()
            self._m[mode] = (m, m.T.tocsr())
        return self._m[mode]

    def sample_rows(self, mode, h, seed):
        m, mt = self.oracle(mode)
        rng = np.random.default_rng(seed)
        rows = [rng.choice(self.n, size=min(N_SAMPLE, self.n), replace=False)]
        rows.append(np.argsort(np.diff(m.indptr))[-64:])
        rows.append(np.argsort(np.diff(mt.indptr))[-64:])
        last_level = int(sum(h.level_sizes[:-1]))
        rows.append(h._dev["lvl_node"][last_level:].cpu().numpy()[:256])
        return np.unique(np.concatenate(rows).astype(np.int64))

    def build(self, product, alpha):
        cls = CLASSES[product][0]
        _say(f"{self.name}: building {cls.__name__} alpha={alpha}")
        return cls(self.ei, torch.ones(self.ei.shape[1]), alpha=alpha, device=self.cuda)


_CACHE = {}


def workload(name, cuda):
    if name not in _CACHE:
        _CACHE.clear()  # one full-size graph at a time
        _CACHE[name] = Workload(name, cuda)
    return _CACHE[name]


def _randn(n, f, seed):
    return np.random.default_rng(seed).standard_normal((n, f)).astype(np.float32)


def _rel(got, ref):
    return O.max_rel_error(got, ref)


def _check_products(wl, mat, mode, f, dtypes=("fp32", "bf16")):
    """Forward and adjoint at width f against FP64 on sampled rows; returns the
    errors."""
    h = mat.cusparse_extension
    m, mt = wl.oracle(mode)
    rows = wl.sample_rows(mode, h, 11 + f)
    x32 = _randn(wl.n, f, 2000 + f)
    g32 = _randn(wl.n, f, 3000 + f)
    cuda = wl.cuda
    errs = {}
    res = {}
    for tag in dtypes:
        xd = torch.from_numpy(x32).to(cuda)
        gd = torch.from_numpy(g32).to(cuda)
        if tag == "bf16":
            xd, gd = xd.to(torch.bfloat16), gd.to(torch.bfloat16)
        y = torch.empty(wl.n, f, device=cuda)
        dx = torch.empty(wl.n, f, device=cuda)
        mat.matmul(xd, y)
        h.matmul_t(gd, dx)
        torch.cuda.synchronize()
        ys, ds = y.cpu().numpy()[rows], dx.cpu().numpy()[rows]
        res[tag] = (ys, ds)
        xs, gs = (x32, g32) if tag == "fp32" else (xd.float().cpu().numpy(), gd.float().cpu().numpy())
        errs[tag] = (_rel(ys, m[rows] @ xs.astype(np.float64)), _rel(ds, mt[rows] @ gs.astype(np.float64)))
        if tag == "bf16":
            errs["bf16_end_to_end"] = (_rel(ys, m[rows] @ x32.astype(np.float64)),
                                       _rel(ds, mt[rows] @ g32.astype(np.float64)))
    if "fp32" in res and "bf16" in res:
        errs["bf16_vs_fp32_cbm"] = (_rel(res["bf16"][0], res["fp32"][0]), _rel(res["bf16"][1], res["fp32"][1]))
    _say(f"{wl.name} {mode} F={f}: {errs}")
    for tag in dtypes:
        assert errs[tag][0] <= 1e-5 and errs[tag][1] <= 1e-5, f"{tag}: {errs[tag]}"
    if "bf16" in dtypes:
        assert max(errs["bf16_end_to_end"]) <= 1e-2, errs
        if "fp32" in dtypes:
            assert max(errs["bf16_vs_fp32_cbm"]) <= 1e-2, errs
    return errs


def _check_order_bits(wl, mat, f):
    """The gather order changes which rows run together, never a bit: the
    forward and the adjoint equal those in position order."""
    h = mat.cusparse_extension
    x = torch.from_numpy(_randn(wl.n, f, 4000 + f)).to(wl.cuda).to(torch.bfloat16)
    outs = []
    try:
        for on in (True, False):
            h.set_order(on)
            y = torch.empty(wl.n, f, device=wl.cuda)
            dx = torch.empty(wl.n, f, device=wl.cuda)
            mat.matmul(x, y)
            h.matmul_t(x, dx)
            torch.cuda.synchronize()
            outs.append((y, dx))
    finally:
        h.set_order(None)
    assert torch.equal(outs[0][0], outs[1][0]) and torch.equal(outs[0][1], outs[1][1])


def _check_integer(wl, h, f, seed):
    """Integer X and G: the plain product A (handle h) bitwise exact on the
    sampled rows, forward and adjoint."""
    a = wl.a
    rows = wl.sample_rows("ax", h, seed)
    rng = np.random.default_rng(seed)
    x = rng.integers(-8, 9, size=(wl.n, f)).astype(np.float32)
    xd = torch.from_numpy(x).to(wl.cuda)
    y = torch.empty(wl.n, f, device=wl.cuda)
    dx = torch.empty(wl.n, f, device=wl.cuda)
    h.forward(xd, y)
    h.matmul_t(xd, dx)
    torch.cuda.synchronize()
    exact = a[rows] @ x.astype(np.float64)
    exact_t = a.T.tocsr()[rows] @ x.astype(np.float64)
    assert np.array_equal(y.cpu().numpy()[rows].astype(np.float64), exact)
    assert np.array_equal(dx.cpu().numpy()[rows].astype(np.float64), exact_t)


# --------------------------------------------------------------------- C2

@pytest.mark.parametrize("alpha", [0, 8])
def test_c2_dad_full_graph(cuda, alpha):
    """cbm4dad, N = 19,717, F in {16, 128}: the whole result against FP64."""
    wl = workload("c2", cuda)
    mat = wl.build("dad", alpha)
    h = mat.cusparse_extension
    _say(f"c2 alpha={alpha}: nnz(Delta)={h.nnz} depth={h.depth}")
    for f in (16, 128):
        _check_products(wl, mat, "dadx", f)


# --------------------------------------------------------------------- C3

@pytest.mark.parametrize("over_a", [False, True])
def test_c3_dad_full_graph(cuda, over_a):
    """cbm4dad, N = 169,343, alpha = 16 (about 14 entries per row: the
    forward Delta pass takes the small-matrix long-row tables, rows above 64
    entries in 256-entry chunks, and the gather order), F in {128, 256},
    compressed and over A (the form the cost model picks at this size):
    forward and adjoint against FP64 on sampled rows, integer X exact over
    A, and the gather order changes no bit."""
    wl = workload("c3", cuda)
    mat = CLASSES["dad"][0](wl.ei, torch.ones(wl.ei.shape[1]), alpha=16, device=cuda, uncompressed=over_a)
    h = mat.cusparse_extension
    if over_a:
        h.attach_uncompressed(h._plain, True)
        assert h.uncompressed_active(torch.float32, 128) == 1
    _say(f"c3 over_a={over_a}: nnz(Delta)={h.nnz} depth={h.depth}")
    for f in (128, 256):
        _check_products(wl, mat, "dadx", f)
    _check_order_bits(wl, mat, 128)
    if over_a:
        plain = h._plain.with_scaling(MODE_AX)
        plain.reserve(128)
        _check_integer(wl, plain, 128, 43)


# --------------------------------------------------------------------- C4

@pytest.fixture(scope="module")
def c4(cuda):
    wl = workload("c4", cuda)
    return wl, wl.build("ad", 32)


def test_c4_layout_reaches_the_size_dependent_paths(c4):
    wl, mat = c4
    h = mat.cusparse_extension
    assert h.depth > 32, "compensated levels (from 32 on) must exist"
    assert h.plan_stats["tf_level"] < h.depth, "the forward tree tail must exist"
    assert h.plan_stats["fwd_pieces"] > 0 and h.plan_stats["fwd_pieces4"] > 0
    # automatic XCD ranges (about 600 entries per row, gathered matrix 24-256 MiB)
    assert h.xcd_ranges_active(torch.bfloat16, 256) == 4
    assert h.xcd_ranges_active(torch.float32, 256) == 4
    assert h.xcd_ranges_active(torch.float32, 64) == 8
    assert h.xcd_ranges_active(torch.bfloat16, 64) == 0


@pytest.mark.parametrize("form", ["forward_structure", "delta_t"])
@pytest.mark.parametrize("f", [64, 256])
def test_c4_ad_full_graph(c4, f, form):
    """cbm4ad, N = 132,534, alpha = 32, BF16 and FP32 with the automatic XCD
    ranges; Kahan levels 32..51 lie inside the forward tail.  The adjoint
    (A D)^T = D A through the forward structure (the default: A is
    symmetric) and through the Delta^T pass."""
    wl, mat = c4
    h = mat.cusparse_extension
    assert h.symmetric
    h.set_symmetric(form == "forward_structure")
    try:
        _check_products(wl, mat, "adx", f)
    finally:
        h.set_symmetric(True)


@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16])
def test_c4_tail_staging_same_bits(c4, dtype):
    """C4's forward tail (fragments of up to 41 levels and 372 rows, Kahan
    levels inside it) staged in LDS for its deep fragments (the default), for
    every fragment that fits, and for none, and per level: one set of bits."""
    wl, mat = c4
    h = mat.cusparse_extension
    f = 256
    x = torch.from_numpy(_randn(wl.n, f, 77)).to(wl.cuda).to(dtype)
    out = {}
    try:
        for mode in (True, "all_staged", "unstaged", False):
            h.set_tree_tails(mode)
            y = torch.full((wl.n, f), float("nan"), device=x.device)
            h.forward(x, y)
            out[mode] = y
    finally:
        h.set_tree_tails(True)
    for mode in (True, "all_staged", "unstaged"):
        assert torch.equal(out[mode], out[False]), mode


def test_c4_integer_exact(c4):
    """The plain product A over C4's structure (rows of up to thousands of
    entries, depth 52, with XCD ranges at FP32 F = 64): integer X exact."""
    wl, mat = c4
    h = mat.cusparse_extension.with_scaling(MODE_AX)
    h.reserve(64)
    assert h.xcd_ranges_active(torch.float32, 64) == 8
    _check_integer(wl, h, 64, 41)


# --------------------------------------------------------------------- C5

@pytest.mark.parametrize("alpha", [8, 32])
def test_c5_mm_full_graph(cuda, alpha):
    """cbm4mm, N = 2,449,029 (no range pieces above 524,288 rows), F in
    {32, 128}, FP32 and BF16, plus integer X exact."""
    wl = workload("c5", cuda)
    mat = wl.build("mm", alpha)
    h = mat.cusparse_extension
    assert h.plan_stats["fwd_pieces"] == 0 and h.plan_stats["bwd_pieces"] == 0
    _say(f"c5 alpha={alpha}: nnz(Delta)={h.nnz} depth={h.depth}")
    assert h.order_active(torch.bfloat16, 128) and h.order_active(torch.float32, 32)
    for f in (32, 128):
        _check_products(wl, mat, "ax", f)
    _check_integer(wl, h, 32, 51 + alpha)
    _check_order_bits(wl, mat, 128)
    del mat, h
    torch.cuda.empty_cache()
