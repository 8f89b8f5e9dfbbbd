"""The eight s_*_csr_int32 operators (cbm/cbm_extensions.cpp:351-907) on the
GPU against the FP64 oracle, fed exactly as the reference would feed them:
Delta from ``init`` as a 4-array CSR and the branch / tree-edge arrays."""
import numpy as np
import pytest
import torch

from cbm import cbm_mkl_cpp, structure
from oracle import oracle as O

pytestmark = pytest.mark.gpu

N = 1500


@pytest.fixture(scope="module")
def setup(cuda):
    ei = structure.copying_model(N, 4, 0.85, 0.6, 0, 4242)
    a = O.csr_of(ei.numpy(), N)
    data = cbm_mkl_cpp.init(ei[0], ei[1], torch.ones(ei.shape[1]), N, 0)
    coo = torch.sparse_coo_tensor(torch.stack([data[0], data[1]]).long(), data[2], (N, N)).coalesce()
    csr = coo.to_sparse_csr()
    crow = csr.crow_indices().int()
    dev = {
        "row_b": crow[:-1].contiguous().to(cuda), "row_e": crow[1:].contiguous().to(cuda),
        "col": csr.col_indices().int().to(cuda), "val": csr.values().float().to(cuda),
        "branches": data[3].to(cuda), "src": data[4].to(cuda), "dst": data[5].to(cuda),
    }
    struct = {"delta_row": data[0].numpy(), "delta_col": data[1].numpy(), "delta_val": data[2].numpy(),
              "branches": data[3].numpy(), "tree_src": data[4].numpy(), "tree_dst": data[5].numpy()}
    # A itself as a 4-array CSR with values
    ar = torch.from_numpy(a.indptr.astype(np.int32))
    dev_a = {"row_b": ar[:-1].contiguous().to(cuda), "row_e": ar[1:].contiguous().to(cuda),
             "col": torch.from_numpy(a.indices.astype(np.int32)).to(cuda),
             "val": torch.ones(a.nnz, dtype=torch.float32, device=cuda)}
    return a, dev, dev_a, struct


def _x(f, seed, cuda):
    return torch.from_numpy(np.random.default_rng(seed).standard_normal((N, f)).astype(np.float32)).to(cuda)


def test_spmm_update_is_a_x(setup, cuda):
    a, d, _, _ = setup
    x = _x(48, 1, cuda)
    y = torch.empty(N, 48, device=cuda)
    cbm_mkl_cpp.s_spmm_update_csr_int32(d["row_b"], d["row_e"], d["col"], d["val"], x, d["branches"], d["src"],
                                        d["dst"], y)
    assert O.max_rel_error(y.cpu().numpy(), O.forward(a, x.cpu().numpy(), "ax")) <= 1e-5


def test_spmm_fused_update_is_dad_x(setup, cuda):
    a, d, _, _ = setup
    dv = torch.from_numpy(O.degree_scale(a).astype(np.float32)).to(cuda)
    rows = torch.repeat_interleave(torch.arange(N, device=cuda), (d["row_e"] - d["row_b"]).long())
    val = d["val"] * dv[d["col"].long()] * dv[rows]
    x = _x(16, 2, cuda)
    y = torch.empty(N, 16, device=cuda)
    cbm_mkl_cpp.s_spmm_fused_update_csr_int32(d["row_b"], d["row_e"], d["col"], val, x, d["branches"], d["src"],
                                              d["dst"], dv, y)
    assert O.max_rel_error(y.cpu().numpy(), O.forward(a, x.cpu().numpy(), "dadx")) <= 1e-5


def test_spmm_add_and_update_add(setup, cuda):
    a, d, da, _ = setup
    x = _x(32, 3, cuda)
    eps = 0.25
    ref = O.forward(a, x.cpu().numpy(), "ax") + (1 + eps) * x.cpu().numpy().astype(np.float64)
    y = torch.empty(N, 32, device=cuda)
    cbm_mkl_cpp.s_spmm_add_csr_int32(da["row_b"], da["row_e"], da["col"], da["val"], x, eps, y)
    assert O.max_rel_error(y.cpu().numpy(), ref) <= 1e-5
    y2 = torch.empty(N, 32, device=cuda)
    cbm_mkl_cpp.s_spmm_update_add_csr_int32(d["row_b"], d["row_e"], d["col"], d["val"], x, eps, d["branches"],
                                            d["src"], d["dst"], y2)
    assert O.max_rel_error(y2.cpu().numpy(), ref) <= 1e-5


def test_plain_and_transposed_spmm(setup, cuda):
    a, _, da, _ = setup
    x = _x(24, 4, cuda)
    y = torch.empty(N, 24, device=cuda)
    cbm_mkl_cpp.s_spmm_csr_int32(da["row_b"], da["row_e"], da["col"], da["val"], x, y)
    assert O.max_rel_error(y.cpu().numpy(), O.forward(a, x.cpu().numpy(), "ax")) <= 1e-5
    # a non-symmetric operand: A D (values scaled by column)
    dv = torch.from_numpy(O.degree_scale(a).astype(np.float32)).to(cuda)
    vals = da["val"] * dv[da["col"].long()]
    yt = torch.empty(N, 24, device=cuda)
    cbm_mkl_cpp.s_spmm_csr_t_int32(da["row_b"], da["row_e"], da["col"], vals, x, yt)
    assert O.max_rel_error(yt.cpu().numpy(), O.adjoint(a, x.cpu().numpy(), "adx")) <= 1e-5


def test_update_ops_follow_reference_semantics(setup, cuda):
    a, d, _, struct = setup
    x = _x(40, 5, cuda)
    base = torch.empty(N, 40, device=cuda)
    cbm_mkl_cpp.s_spmm_csr_int32(d["row_b"], d["row_e"], d["col"], d["val"], x, base)
    y = base.clone()
    cbm_mkl_cpp.s_update_csr_int32(d["branches"], d["src"], d["dst"], y)
    ref = O.cbm_forward(a, struct, x.cpu().numpy(), "ax")
    assert O.max_rel_error(y.cpu().numpy(), ref) <= 1e-5
    m = torch.rand(N, device=cuda) + 0.5
    y2 = base.clone()
    cbm_mkl_cpp.s_fused_update_csr_int32(d["branches"], d["src"], d["dst"], m, y2)
    y_ref = base.double().cpu().numpy().copy()
    mm = m.double().cpu().numpy()
    br, s, t = struct["branches"], struct["tree_src"], struct["tree_dst"]
    for b in range(len(br) - 1):
        for k in range(br[b], br[b + 1]):
            y_ref[t[k]] += y_ref[s[k]] * mm[t[k]] / mm[s[k]]
    assert O.max_rel_error(y2.cpu().numpy(), y_ref) <= 1e-5


def test_ops_reject_host_tensors_and_wrong_dtypes(setup, cuda):
    _, d, _, _ = setup
    x = _x(4, 6, cuda)
    y = torch.empty(N, 4, device=cuda)
    with pytest.raises(RuntimeError, match="not a CUDA tensor"):
        cbm_mkl_cpp.s_spmm_csr_int32(d["row_b"].cpu(), d["row_e"], d["col"], d["val"], x, y)
    with pytest.raises(RuntimeError, match="float32"):
        cbm_mkl_cpp.s_spmm_csr_int32(d["row_b"], d["row_e"], d["col"], d["val"], x.double(), y)


def test_transposed_spmm_rectangular_gapped_and_deterministic(cuda):
    """s_spmm_csr_t_int32 on a rectangular operand given as a 4-array CSR with
    gaps between rows and an empty row: equal to FP64 A^T rhs, the same bits
    on every call (the transpose is a stable sort), and the MKL-style column
    count N = dst.size(0)."""
    rng = np.random.default_rng(9)
    n_rows, n_cols, f = 700, 450, 20
    lens = rng.integers(0, 30, size=n_rows)
    lens[5] = 0
    gaps = rng.integers(0, 4, size=n_rows)
    row_b = np.zeros(n_rows, np.int64)
    pos = 0
    cols, vals = [], []
    for r in range(n_rows):
        pos += gaps[r]
        row_b[r] = pos
        cols.append(np.sort(rng.choice(n_cols, size=lens[r], replace=False)))
        vals.append(rng.standard_normal(lens[r]))
        pos += lens[r]
    total = pos
    col = np.zeros(total, np.int32)
    val = np.zeros(total, np.float32)
    for r in range(n_rows):
        col[row_b[r]:row_b[r] + lens[r]] = cols[r]
        val[row_b[r]:row_b[r] + lens[r]] = vals[r]
    row_e = row_b + lens
    a = np.zeros((n_rows, n_cols))
    for r in range(n_rows):
        a[r, cols[r]] = vals[r].astype(np.float32)
    x = rng.standard_normal((n_rows, f)).astype(np.float32)
    t = lambda v: torch.from_numpy(np.ascontiguousarray(v)).to(cuda)
    dst = torch.empty(n_cols, f, device=cuda)
    cbm_mkl_cpp.s_spmm_csr_t_int32(t(row_b.astype(np.int32)), t(row_e.astype(np.int32)), t(col), t(val), t(x), dst)
    ref = a.T @ x.astype(np.float64)
    assert O.max_rel_error(dst.cpu().numpy(), ref) <= 1e-5
    again = torch.empty_like(dst)
    cbm_mkl_cpp.s_spmm_csr_t_int32(t(row_b.astype(np.int32)), t(row_e.astype(np.int32)), t(col), t(val), t(x), again)
    assert torch.equal(dst, again)
