"""Training step of a 2-layer GCN (hidden 256) over cbm4dad on the third
workload graph (N = 169,343, ~2.3M nnz, alpha = 16), X N x 128, G N x 256:
gradients against the rocSPARSE CSR-transpose path and against central
finite differences of the loss evaluated in FP64 by the oracle."""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

from cbm import structure
from cbm.cbm4dad import cbm4dad
from cbm.csr4dad import csr4dad
from cbm.gcn import GCN
from oracle import oracle as O

pytestmark = pytest.mark.gpu

N = 169343


@pytest.fixture(scope="module")
def setup(cuda):
    ei = structure.copying_model(N, 40, 0.8, 1.63, 0, 1003)
    rng = np.random.default_rng(2003)
    x = rng.standard_normal((N, 128)).astype(np.float32)
    g = rng.standard_normal((N, 256)).astype(np.float32)
    cbm_adj = cbm4dad(ei, torch.ones(ei.shape[1]), alpha=16, device=cuda)
    csr_adj = csr4dad(ei, num_nodes=N, device=cuda)
    return ei, x, g, cbm_adj, csr_adj


def _grads(adj, x, g, cuda):
    model = GCN(adj, 128, 256, 256, seed=5, device=cuda)
    xt = torch.from_numpy(x).to(cuda)
    gt = torch.from_numpy(g).to(cuda)
    loss = (model(xt) * gt).sum()
    loss.backward()
    return model, float(loss), model.w1.grad.clone(), model.w2.grad.clone()


def test_gcn_gradients_match_csr_transpose(setup, cuda):
    _, x, g, cbm_adj, csr_adj = setup
    _, l1, a1, b1 = _grads(cbm_adj, x, g, cuda)
    _, l2, a2, b2 = _grads(csr_adj, x, g, cuda)
    assert abs(l1 - l2) <= 1e-5 * max(1.0, abs(l2))
    assert O.max_rel_error(a1.cpu().numpy(), a2.cpu().numpy()) <= 1e-4
    assert O.max_rel_error(b1.cpu().numpy(), b2.cpu().numpy()) <= 1e-4


def test_gcn_training_step_and_fd(setup, cuda):
    ei, x, g, cbm_adj, _ = setup
    model, loss, gw1, gw2 = _grads(cbm_adj, x, g, cuda)
    a = O.csr_of(ei.numpy(), N)
    ahat = O.product_matrix(a, "dadx")
    w1 = model.w1.detach().double().cpu().numpy()
    w2 = model.w2.detach().double().cpu().numpy()
    x64, g64 = x.astype(np.float64), g.astype(np.float64)

    def loss64(w1_, w2_):
        h = np.maximum(ahat @ (x64 @ w1_), 0.0)
        return float(np.sum((ahat @ (h @ w2_)) * g64))

    assert abs(loss64(w1, w2) - loss) <= 1e-4 * max(1.0, abs(loss))
    rng = np.random.default_rng(8)
    eps = 1e-3
    for which, w_grad in (("w1", gw1), ("w2", gw2)):
        d = rng.standard_normal(w1.shape if which == "w1" else w2.shape)
        if which == "w1":
            fd = (loss64(w1 + eps * d, w2) - loss64(w1 - eps * d, w2)) / (2 * eps)
        else:
            fd = (loss64(w1, w2 + eps * d) - loss64(w1, w2 - eps * d)) / (2 * eps)
        an = float(np.sum(w_grad.double().cpu().numpy() * d))
        assert abs(fd - an) <= 1e-3 * max(1.0, abs(fd)), (which, fd, an)
    # one SGD step lowers a quadratic loss built on the same model
    opt = torch.optim.SGD(model.parameters(), lr=1e-3)
    xt = torch.from_numpy(x).to(cuda)
    target = torch.zeros(N, 256, device=cuda)
    l_before = float(((model(xt) - target) ** 2).mean())
    for _ in range(3):
        opt.zero_grad()
        lq = ((model(xt) - target) ** 2).mean()
        lq.backward()
        opt.step()
    assert float(((model(xt) - target) ** 2).mean()) < l_before
