"""Training step of a 2-layer GCN (hidden 256) over cbm4dad on the third
workload graph (N = 169,343, ~2.3M nnz, alpha = 16), X N x 128, G N x 256:
gradients against the rocSPARSE CSR-transpose path and against central
finite differences of the loss evaluated in FP64 by the oracle."""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

from cbm import cbm_mkl_cpp, structure
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
    return model, float(loss.detach()), model.w1.grad.clone(), model.w2.grad.clone()


def test_gcn_gradients_match_csr_transpose(setup, cuda):
    _, x, g, cbm_adj, csr_adj = setup
    _, l1, a1, b1 = _grads(cbm_adj, x, g, cuda)
    _, l2, a2, b2 = _grads(csr_adj, x, g, cuda)
    assert abs(l1 - l2) <= 1e-5 * max(1.0, abs(l2))
    assert O.max_rel_error(a1.cpu().numpy(), a2.cpu().numpy()) <= 1e-4
    assert O.max_rel_error(b1.cpu().numpy(), b2.cpu().numpy()) <= 1e-4


def test_gcn_gradients_match_fp64_reference(setup, cuda):
    """Weight gradients of the GPU step against FP64 autograd on the host
    (torch.sparse in double precision), and central differences of the FP64
    loss along random weight directions (eps = 1e-6, so that almost no ReLU
    input crosses zero inside the stencil).  Bound 2e-3 (normwise): the FP32
    and FP64 forwards disagree on the ReLU state of the few pre-activations
    that lie within rounding of zero, which moves the weight gradient by a
    few 1e-4 whatever the sparse products do; the sparse adjoint alone is
    held to 1e-5 in test_adjoint_finite_differences_fp64."""
    ei, x, g, cbm_adj, _ = setup
    model, loss, gw1, gw2 = _grads(cbm_adj, x, g, cuda)
    a = O.csr_of(ei.numpy(), N)
    ahat = O.product_matrix(a, "dadx").tocoo()
    at = torch.sparse_coo_tensor(np.stack([ahat.row, ahat.col]), ahat.data, ahat.shape).coalesce()
    w1 = model.w1.detach().double().cpu().requires_grad_(True)
    w2 = model.w2.detach().double().cpu().requires_grad_(True)
    x64, g64 = torch.from_numpy(x).double(), torch.from_numpy(g).double()
    out = torch.sparse.mm(at, torch.relu(torch.sparse.mm(at, x64 @ w1)) @ w2)
    ref_loss = (out * g64).sum()
    ref_loss.backward()
    assert abs(float(ref_loss.detach()) - loss) <= 1e-4 * max(1.0, abs(float(ref_loss.detach())))
    assert O.max_rel_error(gw1.cpu().numpy(), w1.grad.numpy()) <= 2e-3
    assert O.max_rel_error(gw2.cpu().numpy(), w2.grad.numpy()) <= 2e-3

    ahat_csr = ahat.tocsr()

    def loss64(w1_, w2_):
        h = np.maximum(ahat_csr @ (x.astype(np.float64) @ w1_), 0.0)
        return float(np.sum((ahat_csr @ (h @ w2_)) * g.astype(np.float64)))

    rng = np.random.default_rng(8)
    eps = 1e-6
    w1n, w2n = w1.detach().numpy(), w2.detach().numpy()
    for which, w_grad in (("w1", gw1), ("w2", gw2)):
        d = rng.standard_normal(w1n.shape if which == "w1" else w2n.shape)
        if which == "w1":
            fd = (loss64(w1n + eps * d, w2n) - loss64(w1n - eps * d, w2n)) / (2 * eps)
        else:
            fd = (loss64(w1n, w2n + eps * d) - loss64(w1n, w2n - eps * d)) / (2 * eps)
        an = float(np.sum(w_grad.double().cpu().numpy() * d))
        assert abs(fd - an) <= 2e-3 * max(1.0, abs(fd)), (which, fd, an)


def test_adjoint_finite_differences_fp64(setup, cuda):
    """The adjoint against central differences (eps = 1e-3) of
    f(X) = <G, M X> with M X evaluated in FP64 by the oracle's CBM form
    (Delta' X, then the reference propagation): <M^T G, D> = (f(X+eps D) - f(X-eps D)) / 2 eps."""
    ei, x, g, cbm_adj, _ = setup
    a = O.csr_of(ei.numpy(), N)
    built = cbm_adj.build_stats  # the structure the GPU handle was made from
    assert built["n"] == N
    data = cbm_mkl_cpp.init(ei[0], ei[1], torch.ones(ei.shape[1]), N, 16)
    struct = {"delta_row": data[0].numpy(), "delta_col": data[1].numpy(), "delta_val": data[2].numpy(),
              "branches": data[3].numpy(), "tree_src": data[4].numpy(), "tree_dst": data[5].numpy()}
    f = 32
    g32 = g[:, :f].copy()
    dx = torch.empty(N, f, device=cuda)
    cbm_adj.cusparse_extension.matmul_t(torch.from_numpy(g32).to(cuda), dx)
    rng = np.random.default_rng(44)
    x0 = x[:, :f].astype(np.float64)
    eps = 1e-3
    for _ in range(2):
        d = rng.standard_normal((N, f))
        fp = float(np.sum(g32 * O.cbm_forward(a, struct, x0 + eps * d, "dadx")))
        fm = float(np.sum(g32 * O.cbm_forward(a, struct, x0 - eps * d, "dadx")))
        fd = (fp - fm) / (2 * eps)
        an = float(np.sum(dx.double().cpu().numpy() * d))
        assert abs(fd - an) <= 1e-5 * max(1.0, abs(fd)), (fd, an)


def test_gcn_training_step(setup, cuda):
    """A few SGD steps on a quadratic loss lower it (forward + both adjoints + update)."""
    _, x, _, cbm_adj, _ = setup
    model = GCN(cbm_adj, 128, 256, 256, seed=6, device=cuda)
    opt = torch.optim.SGD(model.parameters(), lr=1e-2)
    xt = torch.from_numpy(x).to(cuda)
    target = torch.zeros(N, 256, device=cuda)
    with torch.no_grad():
        l_before = float(((model(xt) - target) ** 2).mean())
    for _ in range(3):
        opt.zero_grad()
        lq = ((model(xt) - target) ** 2).mean()
        lq.backward()
        opt.step()
    with torch.no_grad():
        assert float(((model(xt) - target) ** 2).mean()) < l_before
