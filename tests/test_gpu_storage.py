"""Structure files through the classes (cbm4*.save / cbm4*.load): a loaded
matrix computes the same bits as the one that was saved, forward and adjoint,
FP32 and BF16, with long rows; a file loads into any product class."""
import numpy as np
import pytest
import torch

from cbm import structure
from cbm.cbm4ad import cbm4ad
from cbm.cbm4dad import cbm4dad
from cbm.cbm4mm import cbm4mm
from hubgraph import HUB_N, hub_graph
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _products(m, x, cuda):
    n, f = x.shape
    y = torch.empty(n, f, device=cuda)
    dx = torch.empty(n, f, device=cuda)
    m.cusparse_extension.forward(x, y)
    m.cusparse_extension.matmul_t(x, dx)
    return y, dx


@pytest.mark.parametrize("cls", [cbm4mm, cbm4ad, cbm4dad])
def test_saved_matrix_loads_bit_identical(cls, cuda, tmp_path):
    ei = structure.copying_model(2708, 7, 0.8, 0.58, 0, 1001)
    m = cls(ei, torch.ones(ei.shape[1]), alpha=3, device=cuda)
    path = tmp_path / "c1.npz"
    m.save(path)
    back = cls.load(path, device=cuda)
    assert back.num_nodes == m.num_nodes and back.alpha == 3 and back.build_stats == m.build_stats
    assert torch.equal(back.deltas.values(), m.deltas.values())
    for dtype, f in ((torch.float32, 64), (torch.bfloat16, 128), (torch.float32, 3)):
        x = torch.randn(2708, f, device=cuda, generator=torch.Generator(device=cuda).manual_seed(f)).to(dtype)
        for a, b in zip(_products(m, x, cuda), _products(back, x, cuda)):
            assert torch.equal(a, b)


def test_long_rows_round_trip(cuda, tmp_path):
    ei = hub_graph()
    m = cbm4dad(ei, torch.ones(ei.shape[1]), alpha=0, device=cuda)
    path = tmp_path / "hub.npz"
    m.save(path)
    back = cbm4dad.load(path, device=cuda)
    x = torch.randn(HUB_N, 96, device=cuda, generator=torch.Generator(device=cuda).manual_seed(3))
    for a, b in zip(_products(m, x, cuda), _products(back, x, cuda)):
        assert torch.equal(a, b)
    a = O.csr_of(ei.numpy(), HUB_N)
    y, _ = _products(back, x, cuda)
    assert O.max_rel_error(y.cpu().numpy(), O.forward(a, x.cpu().numpy(), "dadx")) <= 1e-5


def test_file_loads_into_another_product(cuda, tmp_path):
    """The structure does not depend on the product: a cbm4dad file loaded as
    cbm4mm computes A x, bit for bit as a cbm4mm built from the edges."""
    ei = structure.copying_model(2708, 7, 0.8, 0.58, 0, 1001)
    path = tmp_path / "dad.npz"
    cbm4dad(ei, torch.ones(ei.shape[1]), alpha=1, device=cuda).save(path)
    as_mm = cbm4mm.load(path, device=cuda)
    fresh = cbm4mm(ei, torch.ones(ei.shape[1]), alpha=1, device=cuda)
    x = torch.randn(2708, 32, device=cuda, generator=torch.Generator(device=cuda).manual_seed(9))
    for a, b in zip(_products(as_mm, x, cuda), _products(fresh, x, cuda)):
        assert torch.equal(a, b)
    a = O.csr_of(ei.numpy(), 2708)
    assert np.allclose(_products(as_mm, x, cuda)[0].cpu().numpy(), a @ x.cpu().double().numpy(), rtol=1e-5,
                       atol=1e-4)
