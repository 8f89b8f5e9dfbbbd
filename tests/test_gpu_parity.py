"""GPU parity of the CBM product against the oracle (FP64) and the rocSPARSE
CSR product on the same GPU.

Tolerances:
  * FP32 forward / adjoint: normwise relative error max|y - y*| / max|y*| <= 1e-5
    against the FP64 oracle, for N(0,1) inputs;
  * integer-valued X (partial sums < 2^24): bitwise equal to the exact result;
  * X ~ U(0,1): elementwise rtol = 1e-5, atol = 0 against rocSPARSE CSR, the
    protocol of benchmark/compare_matmul_cuda.py:17-18,43,58-59;
  * BF16 storage: <= 1e-5 against FP64 of the BF16-rounded inputs, and
    <= 1e-2 against FP64 of the original FP32 inputs.
"""
import numpy as np
import pytest
import torch

from cbm import structure
from cbm.cbm4ad import cbm4ad
from cbm.cbm4dad import cbm4dad
from cbm.cbm4mm import cbm4mm
from cbm.csr4ad import csr4ad
from cbm.csr4dad import csr4dad
from cbm.csr4mm import csr4mm
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CLASSES = {"ax": (cbm4mm, csr4mm), "adx": (cbm4ad, csr4ad), "dadx": (cbm4dad, csr4dad)}
N1 = 2708  # copying model of the smallest workload: 7 blocks, p = 0.8


@pytest.fixture(scope="module")
def graph1():
    ei = structure.copying_model(N1, 7, 0.8, 0.58, 0, 1001)
    return ei, O.csr_of(ei.numpy(), N1)


@pytest.fixture(scope="module")
def mats(graph1, cuda):
    ei, _ = graph1
    out = {}
    for mode, (cbm_cls, csr_cls) in CLASSES.items():
        out[mode] = (cbm_cls(ei, torch.ones(ei.shape[1]), alpha=0, device=cuda), csr_cls(ei, device=cuda))
    return out


def _rand(shape, seed, dist="normal"):
    rng = np.random.default_rng(seed)
    if dist == "normal":
        return rng.standard_normal(shape).astype(np.float32)
    return rng.random(shape, dtype=np.float32)


@pytest.mark.parametrize("mode", ["ax", "adx", "dadx"])
@pytest.mark.parametrize("f", [1, 3, 16, 64, 128, 256, 520])
def test_forward_fp32_vs_oracle(mats, graph1, cuda, mode, f):
    _, a = graph1
    m, _ = mats[mode]
    x = _rand((N1, f), 2000 + f)
    y = torch.empty(N1, f, device=cuda)
    m.matmul(torch.from_numpy(x).to(cuda), y)
    err = O.max_rel_error(y.cpu().numpy(), O.forward(a, x, mode))
    assert err <= 1e-5, f"{mode} F={f}: rel err {err:.3e}"


@pytest.mark.parametrize("form", ["forward_structure", "delta_t"])
@pytest.mark.parametrize("mode", ["ax", "adx", "dadx"])
@pytest.mark.parametrize("f", [1, 8, 64, 256])
def test_adjoint_fp32_vs_oracle(mats, graph1, cuda, mode, f, form):
    """M^T G both ways: through the forward structure (A = A^T, the default
    for these symmetric graphs) and through the Delta^T pass."""
    _, a = graph1
    m, _ = mats[mode]
    h = m.cusparse_extension
    assert h.symmetric  # the build found A = A^T
    h.set_symmetric(form == "forward_structure")
    g = _rand((N1, f), 3000 + f)
    dx = torch.empty(N1, f, device=cuda)
    m.matmul_t(torch.from_numpy(g).to(cuda), dx)
    h.set_symmetric(True)
    err = O.max_rel_error(dx.cpu().numpy(), O.adjoint(a, g, mode))
    assert err <= 1e-5, f"{mode} F={f}: rel err {err:.3e}"


@pytest.mark.parametrize("f", [1, 5, 64, 128])
def test_integer_x_is_bitwise_exact(mats, graph1, cuda, f):
    """Small-integer X makes every partial sum exact in FP32, so any correct
    evaluation order returns the exact integer result (the adjoint both
    through the forward structure and through the Delta^T pass)."""
    _, a = graph1
    m, _ = mats["ax"]
    rng = np.random.default_rng(77 + f)
    x = rng.integers(-8, 9, size=(N1, f)).astype(np.float32)
    y = torch.empty(N1, f, device=cuda)
    m.matmul(torch.from_numpy(x).to(cuda), y)
    exact = O.forward(a, x, "ax")
    assert np.array_equal(y.cpu().numpy().astype(np.float64), exact)
    for sym in (True, False):
        m.cusparse_extension.set_symmetric(sym)
        dx = torch.empty(N1, f, device=cuda)
        m.matmul_t(torch.from_numpy(x).to(cuda), dx)
        assert np.array_equal(dx.cpu().numpy().astype(np.float64), O.adjoint(a, x, "ax")), sym
    m.cusparse_extension.set_symmetric(True)


@pytest.mark.parametrize("mode", ["ax", "adx", "dadx"])
def test_reference_protocol_vs_rocsparse(mats, cuda, mode):
    """benchmark/compare_matmul_cuda.py:41-62: X ~ U(0,1); CBM, rocSPARSE CSR
    and torch.sparse CSR (``csr.a @ x``) compared pairwise with
    assert_close(rtol=1e-5, atol=0); every failure is counted (the reference
    adds 0 on failure, :62)."""
    m, c = mats[mode]
    failed = 0
    for it in range(5):
        x = torch.from_numpy(_rand((N1, 64), 500 + it, "uniform")).to(cuda)
        y_cbm = torch.empty(N1, 64, device=cuda)
        y_csr = torch.empty(N1, 64, device=cuda)
        m.matmul(x, y_cbm)
        c.matmul(x, y_csr)
        y_torch = c.a @ x  # the torch.sparse leg (compare_matmul_cuda.py:53)
        torch.cuda.synchronize()
        for ref, got in ((y_csr, y_cbm), (y_torch, y_cbm), (y_torch, y_csr)):
            try:
                torch.testing.assert_close(ref, got, rtol=1e-5, atol=0)
            except AssertionError:
                failed += 1
    assert failed == 0, f"{failed} of 15 comparisons failed"


@pytest.mark.parametrize("mode", ["ax", "adx", "dadx"])
def test_adjoint_vs_rocsparse_transpose(mats, cuda, mode):
    """M^T G against rocSPARSE's transposed product.  Even with G >= 0 the
    Delta^T sums mix signs, so an element far below the row's scale can lose
    relative accuracy to cancellation: the bound is rtol = 1e-5 plus an
    absolute floor of 1e-5 * max|M^T G|."""
    m, c = mats[mode]
    g = torch.from_numpy(_rand((N1, 32), 91, "uniform")).to(cuda)
    d_csr = torch.empty(N1, 32, device=cuda)
    c.matmul_t(g, d_csr)
    for sym in (True, False):
        m.cusparse_extension.set_symmetric(sym)
        d_cbm = torch.empty(N1, 32, device=cuda)
        m.matmul_t(g, d_cbm)
        torch.testing.assert_close(d_csr, d_cbm, rtol=1e-5, atol=1e-5 * float(d_csr.abs().max()))
    m.cusparse_extension.set_symmetric(True)


@pytest.mark.parametrize("mode", ["ax", "adx", "dadx"])
@pytest.mark.parametrize("f", [8, 64, 256])
def test_bf16_storage(mats, graph1, cuda, mode, f):
    _, a = graph1
    m, _ = mats[mode]
    x32 = _rand((N1, f), 4000 + f)
    xb = torch.from_numpy(x32).to(cuda).to(torch.bfloat16)
    xb_exact = xb.float().cpu().numpy()
    y = torch.empty(N1, f, device=cuda)
    m.matmul(xb, y)
    assert O.max_rel_error(y.cpu().numpy(), O.forward(a, xb_exact, mode)) <= 1e-5
    assert O.max_rel_error(y.cpu().numpy(), O.forward(a, x32, mode)) <= 1e-2
    dx = torch.empty(N1, f, device=cuda)
    m.matmul_t(xb, dx)
    assert O.max_rel_error(dx.cpu().numpy(), O.adjoint(a, xb_exact, mode)) <= 1e-5


def test_split_halves_match_fused(mats, cuda):
    """delta_spmm followed by propagate equals the multi-launch forward bit
    for bit, and the one-launch forward within FP32 rounding (it sums long
    rows in pieces)."""
    for mode in ("ax", "dadx"):
        h = mats[mode][0].cusparse_extension
        x = torch.from_numpy(_rand((N1, 64), 5)).to(cuda)
        y1 = torch.empty(N1, 64, device=cuda)
        y2 = torch.empty(N1, 64, device=cuda)
        y3 = torch.empty(N1, 64, device=cuda)
        h.set_fused(False)
        h.forward(x, y1)
        h.set_fused(None)
        h.forward(x, y3)
        h.delta_spmm(x, y2)
        h.propagate(y2)
        assert torch.equal(y1, y2)
        assert float((y3 - y1).abs().max()) <= 1e-5 * float(y1.abs().max())


def test_strided_operands(mats, graph1, cuda):
    """Row strides that are not multiples of 16 bytes take the scalar path."""
    _, a = graph1
    m, _ = mats["ax"]
    big = torch.from_numpy(_rand((N1, 70), 8)).to(cuda)
    x = big[:, 3:67]  # ld = 70, misaligned start
    yb = torch.zeros(N1, 66, device=cuda)
    y = yb[:, 1:65]
    m.matmul(x, y)
    err = O.max_rel_error(y.cpu().numpy(), O.forward(a, x.cpu().numpy(), "ax"))
    assert err <= 1e-5
    assert torch.all(yb[:, 0] == 0) and torch.all(yb[:, 65] == 0)


def test_autograd_matches_adjoint_and_dot_product(mats, graph1, cuda):
    _, a = graph1
    m, _ = mats["dadx"]
    x = torch.from_numpy(_rand((N1, 16), 1)).to(cuda).requires_grad_(True)
    g = torch.from_numpy(_rand((N1, 16), 2)).to(cuda)
    y = m.apply(x)
    (y * g).sum().backward()
    ref = O.adjoint(a, g.cpu().numpy(), "dadx")
    assert O.max_rel_error(x.grad.cpu().numpy(), ref) <= 1e-5
    lhs = float((y.double() * g.double()).sum())
    rhs = float((x.detach().double() * x.grad.double()).sum())
    assert abs(lhs - rhs) <= 1e-5 * max(1.0, abs(lhs))


def test_finite_differences_fp32(mats, cuda):
    """Central differences of f(x) = <G, M x> along random directions
    (eps = 1e-3) agree with <M^T G, d> to the FP32 cancellation limit."""
    m, _ = mats["dadx"]
    x = torch.from_numpy(_rand((N1, 8), 3)).to(cuda)
    g = torch.from_numpy(_rand((N1, 8), 4)).to(cuda)
    grad = torch.empty(N1, 8, device=cuda)
    m.matmul_t(g, grad)
    eps = 1e-3
    y = torch.empty(N1, 8, device=cuda)
    for k in range(3):
        d = torch.from_numpy(_rand((N1, 8), 10 + k)).to(cuda)
        m.matmul(x + eps * d, y)
        fp = float((g.double() * y.double()).sum())
        m.matmul(x - eps * d, y)
        fm = float((g.double() * y.double()).sum())
        fd = (fp - fm) / (2 * eps)
        an = float((grad.double() * d.double()).sum())
        assert abs(fd - an) <= 1e-3 * max(1.0, abs(an)), (fd, an)


def test_bf16_autograd(mats, graph1, cuda):
    _, a = graph1
    m, _ = mats["adx"]
    x = torch.from_numpy(_rand((N1, 64), 21)).to(cuda).to(torch.bfloat16).requires_grad_(True)
    y = m.apply(x)
    assert y.dtype == torch.float32
    g = torch.from_numpy(_rand((N1, 64), 22)).to(cuda)
    y.backward(g)
    assert x.grad.dtype == torch.bfloat16
    ref = O.adjoint(a, g.to(torch.bfloat16).float().cpu().numpy(), "adx")
    assert O.max_rel_error(x.grad.float().cpu().numpy(), ref) <= 1e-2


@pytest.mark.parametrize("mode", ["ax", "adx", "dadx"])
def test_degenerate_graphs(cuda, mode):
    """Isolated nodes (D = 0), identical rows (empty Delta rows) and a nested
    chain of rows (one tree level per row) on a small matrix."""
    cbm_cls = CLASSES[mode][0]
    n = 260
    rows, cols = [], []
    # nested rows: row i = {200 + j : j < 40 - i} for i < 40 gives a deep chain
    for i in range(40):
        for j in range(40 - i):
            rows += [i, 200 + j]
            cols += [200 + j, i]
    # identical rows 50..59 all linked to 100..104
    for i in range(50, 60):
        for j in range(100, 105):
            rows += [i, j]
            cols += [j, i]
    # nodes 150..199 are isolated; 259 links to 258 so N is inferred as 260
    rows += [258, 259]
    cols += [259, 258]
    key = sorted(set(zip(rows, cols)))
    ei = torch.tensor(key, dtype=torch.int32).t().contiguous()
    m = cbm_cls(ei, torch.ones(ei.shape[1]), alpha=0, device=cuda)
    assert m.num_nodes == n
    a = O.csr_of(ei.numpy(), n)
    x = _rand((n, 12), 9)
    y = torch.empty(n, 12, device=cuda)
    m.matmul(torch.from_numpy(x).to(cuda), y)
    assert O.max_rel_error(y.cpu().numpy(), O.forward(a, x, mode)) <= 1e-5
    assert m.cusparse_extension.depth >= 30
    dx = torch.empty(n, 12, device=cuda)
    m.matmul_t(torch.from_numpy(x).to(cuda), dx)
    assert O.max_rel_error(dx.cpu().numpy(), O.adjoint(a, x, mode)) <= 1e-5
    assert torch.all(y[150:200] == 0)


def test_reference_style_construction(graph1, cuda):
    """cusparseCBM built exactly as cbm/cbm4dad.py does it: COO -> coalesce ->
    CSR with values scaled by D[row] D[col] and multipliers = D."""
    from cbm import cbm_cusparse, cbm_mkl_cpp
    ei, a = graph1
    data = cbm_mkl_cpp.init(ei[0], ei[1], torch.ones(ei.shape[1]), N1, 0)
    deltas = torch.sparse_coo_tensor(torch.stack([data[0], data[1]]), data[2], (N1, N1)).coalesce().to_sparse_csr()
    deltas = torch.sparse_csr_tensor(deltas.crow_indices().int(), deltas.col_indices().int(), deltas.values(),
                                     deltas.shape)
    d = torch.from_numpy(O.degree_scale(a).astype(np.float32))
    rows = torch.repeat_interleave(torch.arange(N1), deltas.crow_indices().diff().long())
    vals = deltas.values() * d[deltas.col_indices().long()] * d[rows]
    dad = torch.sparse_csr_tensor(deltas.crow_indices(), deltas.col_indices(), vals, deltas.shape).to(cuda)
    h = cbm_cusparse.cusparseCBM(dad, data[3].to(cuda), data[4].to(cuda), data[5].to(cuda), d.to(cuda))
    x = _rand((N1, 32), 12)
    y = torch.empty(N1, 32, device=cuda)
    assert h.matmul_DAX_DADX(torch.from_numpy(x).to(cuda), y) == 0
    assert O.max_rel_error(y.cpu().numpy(), O.forward(a, x, "dadx")) <= 1e-5
    with pytest.raises(ValueError):
        h.matmul_AX_ADX(torch.from_numpy(x).to(cuda), y)


@pytest.mark.parametrize("ranges", [None, 8, 4])
def test_graph_capture_replays(mats, cuda, ranges):
    """Forward + adjoint captured in a HIP graph replay correctly several times,
    with eager work and synchronisations between the replays (with the XCD
    column ranges as chosen by size, and forced on with either piece set)."""
    m, _ = mats["dadx"]
    h = m.cusparse_extension
    h.set_xcd_ranges(ranges)
    h.reserve(64)
    x = torch.from_numpy(_rand((N1, 64), 31)).to(cuda)
    y = torch.empty(N1, 64, device=cuda)
    dx = torch.empty(N1, 64, device=cuda)
    z = torch.empty(N1, 64, device=cuda)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        h.forward(x, y)
        h.matmul_t(y, dx, workspace=z)
    torch.cuda.current_stream().wait_stream(s)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        h.forward(x, y)
        h.matmul_t(y, dx, workspace=z)
    for it in range(3):
        x.copy_(torch.from_numpy(_rand((N1, 64), 40 + it)).to(cuda))
        y_ref = torch.empty_like(y)
        dx_ref = torch.empty_like(dx)
        h.forward(x, y_ref)
        h.matmul_t(y_ref, dx_ref)
        torch.cuda.synchronize()
        y.zero_()
        dx.zero_()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(y, y_ref) and torch.equal(dx, dx_ref)
    h.set_xcd_ranges(None)


def test_product_calls_never_allocate(graph1, cuda):
    """The native product calls refuse a width above the reserved one (they
    would have to allocate); cbm_reserve is what sizes the scratch."""
    import ctypes
    from cbm import _lib
    ei, a = graph1
    m = cbm4mm(ei, torch.ones(ei.shape[1]), alpha=0, device=cuda)
    h = m.cusparse_extension
    lib = _lib.hip()
    x = torch.from_numpy(_rand((N1, 48), 7)).to(cuda)
    y = torch.empty(N1, 48, device=cuda)
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = lib.cbm_matmul(h._h, 0, x.data_ptr(), 48, y.data_ptr(), 48, 48, st)
    assert rc == _lib.CBM_ERR_ARG
    assert b"reserved width" in lib.cbm_last_error()
    z = torch.empty(N1, 48, device=cuda)
    assert lib.cbm_matmul_t(h._h, 0, x.data_ptr(), 48, y.data_ptr(), 48, z.data_ptr(), 48, 48, st) == _lib.CBM_ERR_ARG
    assert lib.cbm_reserve(h._h, 48) == 0
    assert lib.cbm_matmul(h._h, 0, x.data_ptr(), 48, y.data_ptr(), 48, 48, st) == 0
    assert lib.cbm_matmul(h._h, 0, x.data_ptr(), 48, y.data_ptr(), 48, 32, st) == 0  # narrower: fine
    torch.cuda.synchronize()
    assert O.max_rel_error(y[:, :32].cpu().numpy(), O.forward(a, x[:, :32].cpu().numpy(), "ax")) <= 1e-5


def test_autograd_forward_and_backward_capture(graph1, cuda):
    """The autograd product's forward and backward inside one HIP graph,
    replayed three times with eager work between replays; each replay equals
    the eager product and adjoint."""
    ei, a = graph1
    m = cbm4dad(ei, torch.ones(ei.shape[1]), alpha=0, device=cuda)
    m.cusparse_extension.reserve(32)
    x = torch.from_numpy(_rand((N1, 32), 21)).to(cuda).requires_grad_(True)
    g = torch.from_numpy(_rand((N1, 32), 22)).to(cuda)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):  # warm-up outside capture
        (m.apply(x) * g).sum().backward()
    torch.cuda.current_stream().wait_stream(s)
    x.grad = None
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        y = m.apply(x)
        (y * g).sum().backward()
    for it in range(3):
        x.data.copy_(torch.from_numpy(_rand((N1, 32), 60 + it)).to(cuda))
        g.copy_(torch.from_numpy(_rand((N1, 32), 70 + it)).to(cuda))
        y_ref = torch.empty(N1, 32, device=cuda)
        dx_ref = torch.empty(N1, 32, device=cuda)
        m.matmul(x.detach(), y_ref)
        m.matmul_t(g, dx_ref)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(y.detach(), y_ref)
        assert torch.equal(x.grad, dx_ref)
        assert O.max_rel_error(x.grad.cpu().numpy(), O.adjoint(a, g.cpu().numpy(), "dadx")) <= 1e-5
