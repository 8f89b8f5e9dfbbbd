"""The Delta pass in the gather order (csrc/host/order.cpp, RowPass::order)
against the pass in position order and the oracle.

The order only changes which rows run together: every row is summed by the
same kernel, in its own entry order, so the two orders must agree bit for
bit (both with rows packed per wavefront: on graphs this small the plain
pass would otherwise give each row a whole wavefront, whose sub-group sums
add in another order).  Against the FP64 oracle: max-norm relative error
<= 1e-5 (FP32 inputs), and integer-valued X gives the exact result bit for
bit.
"""
import numpy as np
import pytest
import torch

from cbm import structure
from cbm.cbm4ad import cbm4ad
from cbm.cbm4dad import cbm4dad
from cbm.cbm4mm import cbm4mm
from hubgraph import hub_graph
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CLASSES = {"ax": cbm4mm, "adx": cbm4ad, "dadx": cbm4dad}
N = 3000


@pytest.fixture(scope="module")
def graph():
    ei = structure.copying_model(N, 7, 0.8, 2.0, 0, 4242)
    return ei, O.csr_of(ei.numpy(), N)


@pytest.fixture(scope="module")
def mats(graph, cuda):
    ei, _ = graph
    out = {}
    for mode, cls in CLASSES.items():
        m = cls(ei, torch.ones(ei.shape[1]), alpha=0, device=cuda, gather_order=True)
        m.cusparse_extension.set_wave_rows(False)
        out[mode] = m
    return out


def _x(n, f, dtype, seed, cuda, integer=False):
    rng = np.random.default_rng(seed)
    if integer:
        x = rng.integers(-8, 9, size=(n, f)).astype(np.float32)
    else:
        x = rng.standard_normal((n, f)).astype(np.float32)
    t = torch.from_numpy(x).to(cuda)
    return t.to(dtype) if dtype is not None else t


def _both(ext, fn):
    """fn() in the gather order, then in position order."""
    ext.set_order(True)
    a = fn()
    ext.set_order(False)
    try:
        b = fn()
    finally:
        ext.set_order(True)
    return a, b


def test_order_lists_every_short_row_once(mats):
    ext = mats["ax"].cusparse_extension
    order = ext._order
    short = torch.nonzero(ext._dev["fwd_ptr"].diff().cpu() <= structure.LONG_ROW).view(-1)
    assert torch.equal(torch.sort(order.long()).values, short)
    # the walk puts rows with partners next to each other: not position order
    assert not torch.equal(order.long(), short)


@pytest.mark.parametrize("mode", ["ax", "adx", "dadx"])
@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16])
@pytest.mark.parametrize("f", [3, 16, 64, 128, 520])
def test_forward_in_gather_order_equals_position_order(mats, graph, cuda, mode, dtype, f):
    m = mats[mode]
    ext = m.cusparse_extension
    x = _x(N, f, dtype, 100 + f, cuda)
    assert ext.order_active(dtype, f)

    def run():
        y = torch.empty(N, f, device=cuda)
        m.matmul(x, y)
        torch.cuda.synchronize()
        return y

    a, b = _both(ext, run)
    assert torch.equal(a, b)
    if dtype == torch.float32:
        _, csr = graph
        assert O.max_rel_error(a.cpu().numpy(), O.forward(csr, x.cpu().numpy(), mode)) <= 1e-5


@pytest.mark.parametrize("mode", ["ax", "dadx"])
@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16])
def test_adjoint_in_gather_order_equals_position_order(mats, graph, cuda, mode, dtype):
    """A symmetric A: the adjoint runs the forward structure, in the order too."""
    m = mats[mode]
    ext = m.cusparse_extension
    assert ext.symmetric
    f = 128
    g = _x(N, f, dtype, 7, cuda)

    def run():
        out = torch.empty(N, f, device=cuda)
        m.matmul_t(g, out)
        torch.cuda.synchronize()
        return out

    a, b = _both(ext, run)
    assert torch.equal(a, b)
    if dtype == torch.float32:
        _, csr = graph
        assert O.max_rel_error(a.cpu().numpy(), O.adjoint(csr, g.cpu().numpy(), mode)) <= 1e-5


def test_delta_half_in_gather_order(mats, cuda):
    ext = mats["dadx"].cusparse_extension
    x = _x(N, 64, torch.bfloat16, 3, cuda)

    def run():
        y = torch.empty(N, 64, device=cuda)
        ext.delta_spmm(x, y)
        torch.cuda.synchronize()
        return y

    a, b = _both(ext, run)
    assert torch.equal(a, b)


def test_integer_inputs_are_exact(mats, graph, cuda):
    _, csr = graph
    x = _x(N, 128, None, 11, cuda, integer=True)
    y = torch.empty(N, 128, device=cuda)
    mats["ax"].matmul(x, y)
    torch.cuda.synchronize()
    exact = O.forward(csr, x.cpu().numpy().astype(np.float64), "ax")
    assert np.array_equal(y.cpu().numpy().astype(np.float64), exact)


def test_wave_per_row_in_gather_order(graph, cuda):
    """The order also maps the pass that gives each row a whole wavefront
    (the default on small graphs): same bits as that pass in position order."""
    ei, csr = graph
    m = cbm4mm(ei, torch.ones(ei.shape[1]), alpha=0, device=cuda, gather_order=True)
    ext = m.cusparse_extension
    ext.set_wave_rows(True)
    x = _x(N, 128, torch.float32, 21, cuda)

    def run():
        y = torch.empty(N, 128, device=cuda)
        m.matmul(x, y)
        torch.cuda.synchronize()
        return y

    a, b = _both(ext, run)
    assert torch.equal(a, b)
    assert O.max_rel_error(a.cpu().numpy(), O.forward(csr, x.cpu().numpy(), "ax")) <= 1e-5


@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16])
def test_long_rows_keep_their_chunks(cuda, dtype):
    """Hub rows longer than CBM_LONG_ROW entries are summed by their chunks;
    every other row in the order."""
    ei = hub_graph()
    n = int(ei.max()) + 1
    m = cbm4mm(ei, torch.ones(ei.shape[1]), alpha=4, device=cuda, gather_order=True)
    ext = m.cusparse_extension
    ext.set_wave_rows(False)
    assert int((ext._dev["fwd_ptr"].diff() > structure.LONG_ROW).sum()) > 0
    f = 128
    x = _x(n, f, dtype, 9, cuda)

    def run():
        y = torch.empty(n, f, device=cuda)
        m.matmul(x, y)
        torch.cuda.synchronize()
        return y

    a, b = _both(ext, run)
    assert torch.equal(a, b)
    if dtype == torch.float32:
        csr = O.csr_of(ei.numpy(), n)
        assert O.max_rel_error(a.cpu().numpy(), O.forward(csr, x.cpu().numpy(), "ax")) <= 1e-5


def test_given_order_is_checked(mats):
    ext = mats["ax"].cusparse_extension
    bad = ext._order.clone()
    bad[0] = bad[1]
    with pytest.raises(ValueError):
        ext.attach_order(bad)
    ext.attach_order(ext._order)  # the order itself is accepted again


def test_automatic_choice_orders_only_large_matrices(graph, cuda):
    ei, _ = graph
    m = cbm4mm(ei, torch.ones(ei.shape[1]), alpha=0, device=cuda)
    assert getattr(m.cusparse_extension, "_order", None) is None
    assert not m.cusparse_extension.order_active(torch.bfloat16, 128)


def test_saved_order_loads_back(mats, cuda, tmp_path):
    m = mats["ax"]
    path = tmp_path / "m.cbm.npz"
    m.save(path)
    m2 = cbm4mm.load(path, device=cuda)
    assert torch.equal(m2.cusparse_extension._order, m.cusparse_extension._order)
    m2.cusparse_extension.set_wave_rows(False)
    x = _x(N, 64, torch.bfloat16, 31, cuda)
    y1 = torch.empty(N, 64, device=cuda)
    y2 = torch.empty(N, 64, device=cuda)
    m.matmul(x, y1)
    m2.matmul(x, y2)
    torch.cuda.synchronize()
    assert torch.equal(y1, y2)
    m3 = cbm4mm.load(path, device=cuda, gather_order=False)
    assert getattr(m3.cusparse_extension, "_order", None) is None


def test_product_over_a_in_gather_order(graph, cuda):
    """The handle over A itself (small and mid-size graphs) gets an order of
    its own rows: same bits as in position order."""
    ei, csr = graph
    m = cbm4dad(ei, torch.ones(ei.shape[1]), alpha=0, device=cuda, uncompressed=True, gather_order=True)
    ext = m.cusparse_extension
    plain = ext._plain
    assert plain is not None and getattr(plain, "_order", None) is not None
    ext.attach_uncompressed(plain, True)
    plain.set_wave_rows(False)
    assert ext.order_active(torch.float32, 128)
    x = _x(N, 128, torch.float32, 41, cuda)

    def run():
        y = torch.empty(N, 128, device=cuda)
        dx = torch.empty(N, 128, device=cuda)
        m.matmul(x, y)
        m.matmul_t(x, dx)
        torch.cuda.synchronize()
        return torch.cat([y, dx])

    plain.set_order(True)
    a = run()
    plain.set_order(False)
    b = run()
    plain.set_order(True)
    assert torch.equal(a, b)
    assert O.max_rel_error(a[:N].cpu().numpy(), O.forward(csr, x.cpu().numpy(), "dadx")) <= 1e-5
