"""``A @ X`` in CBM form: drop-in for cbm/cbm4mm.py:7-54."""
import torch

from . import cbm_cusparse as cusparse
from . import cbm_mkl_cpp as cbm_mkl
from . import storage, structure
from ._hip_sigs import MODE_AX
from .autograd import cbm_apply


# The uncompressed form is kept beside the CBM structure for matrices up to
# this size (its layout costs about as much memory as A's CSR), when the
# constructor's `uncompressed` is True or None with this default True.
UNCOMPRESSED = True
UNCOMPRESSED_MAX_ROWS = 1 << 18
UNCOMPRESSED_MAX_NNZ = 1 << 24


# The gather order of the Delta pass is computed for matrices of more than
# this many rows when the constructor's `gather_order` is None: there the
# gathered matrix is larger than the L2s, so every gather an L2 serves is a
# fabric request saved (N = 169,343, F = 256 FP32: step 0.60 -> 0.51 ms over
# A; N = 2,449,029, F = 128 BF16: 8.1 -> 6.7 ms; the partner scan and the
# forest walk cost about 0.3 s per million rows, once).
ORDER_MIN_ROWS = 1 << 17


def _wants_uncompressed(num_rows, edge_index):
    return num_rows <= UNCOMPRESSED_MAX_ROWS and int(edge_index.shape[1]) <= UNCOMPRESSED_MAX_NNZ


def _num_rows(edge_index, num_nodes):
    if num_nodes is not None:
        return int(num_nodes)
    # N = max index + 1 (cbm/cbm4mm.py:12): trailing isolated nodes are dropped.
    return int(max(int(edge_index[0].max()), int(edge_index[1].max()))) + 1


class cbm4mm:
    """CBM form of the binary matrix given by ``edge_index`` (``[2, nnz]``),
    ``edge_values`` (all ones) and the pruning parameter ``alpha``.

    Attributes as in the reference: ``deltas`` (CUDA sparse CSR, int32/float32),
    ``num_nodes``, ``mca_offset``, ``mca_src_idx``, ``mca_dst_idx`` and
    ``cusparse_extension``.  ``matmul(x, y)`` writes ``y = A x`` in place.
    Additions: ``x @``-style autograd product via :meth:`apply`, BF16 ``x``,
    ``num_nodes=``, the candidate-edge scan on the GPU (``candidates_on``),
    the build statistics in ``build_stats``, :meth:`save` / :meth:`load`
    of the built structure, and ``uncompressed``: for matrices of up to
    UNCOMPRESSED_MAX_ROWS rows the same product over A itself is kept beside
    the CBM form and the handle runs it where a cost model predicts it
    faster (small graphs, whose tree levels are dependent launches);
    False keeps the CBM form only.  Loaded structures have no A, so they
    run the CBM form.  ``gather_order``: the Delta pass sums the short rows
    in an order that puts rows sharing columns together, each XCD taking one
    stretch of it (same results, more gathers served by the L2s): None for
    matrices of more than ORDER_MIN_ROWS rows, True always, False never.
    """

    def __init__(self, edge_index, edge_values, alpha=0, *, num_nodes=None, device="cuda", n_threads=0,
                 candidates_on="device", uncompressed=None, gather_order=None):
        num_rows = _num_rows(edge_index, num_nodes)
        # The candidate-edge scan of the build runs on the target GPU
        # ("device"), or on the host (None); the structure is the same.
        if candidates_on == "device":
            candidates_on = device if torch.device(device).type == "cuda" else None
        built = cbm_mkl.build(edge_index[0], edge_index[1], edge_values, num_rows, alpha, n_threads=n_threads,
                              candidates_on=candidates_on)
        self.alpha = int(alpha)
        if uncompressed is None:
            uncompressed = UNCOMPRESSED
        self._plain_edges = edge_index if uncompressed and _wants_uncompressed(num_rows, edge_index) else None
        self._order_wanted = num_rows > ORDER_MIN_ROWS if gather_order is None else bool(gather_order)
        self._assemble(num_rows, built.csr_ptr, built.csr_col, built.csr_val, built.branches, built.tree_src,
                       built.tree_dst, built.stats, device, n_threads)
        self._plain_edges = None

    def _assemble(self, num_rows, crow, col, val, branches, tree_src, tree_dst, stats, device, n_threads, plan=None):
        self.build_stats = stats
        # Delta as CSR sorted by (row, col): what coalesce().to_sparse_csr()
        # gives in the reference (cbm/cbm4mm.py:26-42), built directly.
        self.deltas = torch.sparse_csr_tensor(crow, col, val, (num_rows, num_rows)).to(device=device)
        self.num_nodes = num_rows
        self.mca_offset = branches.to(dtype=torch.int32, device=device)
        self.mca_src_idx = tree_src.to(dtype=torch.int32, device=device)
        self.mca_dst_idx = tree_dst.to(dtype=torch.int32, device=device)
        dummy = torch.empty(0, device=device)
        self.cusparse_extension = cusparse.cusparseCBM(self.deltas, self.mca_offset, self.mca_src_idx,
                                                       self.mca_dst_idx, dummy, mode=MODE_AX, n_threads=n_threads,
                                                       plan=plan)
        # A = A^T (the build compares them exactly): the adjoint then runs
        # through the forward structure.
        self.cusparse_extension.set_symmetric(bool(stats.get("symmetric", 0)))
        order = getattr(self, "_stored_order", None)
        if order is not None or getattr(self, "_order_wanted", False):
            if torch.device(device).type != "cuda":
                raise RuntimeError("tensor is not a CUDA tensor")
            self.cusparse_extension.attach_order(order)
        edges = getattr(self, "_plain_edges", None)
        if edges is not None:
            # The same product over A itself, which the handle runs where the
            # tree's dependent launches would cost more than Delta saves.
            from .csr4mm import csr_from_edge_index
            a = csr_from_edge_index(edges, num_rows, device)
            e0 = torch.zeros(0, dtype=torch.int32, device=device)
            plain = cusparse.cusparseCBM(a, torch.zeros(1, dtype=torch.int32, device=device), e0, e0, dummy,
                                         mode=MODE_AX, n_threads=n_threads)
            if getattr(self, "_order_wanted", False):
                plain.attach_order()
            self.cusparse_extension.attach_uncompressed(plain)
        self._scale()

    def _scale(self):
        """The product's scalings (none for A; see cbm4ad / cbm4dad)."""

    def save(self, path):
        """Writes the structure (Delta, tree, device layout) to a file that
        :meth:`load` reads back in place of a rebuild (cbm/storage.py)."""
        ext = self.cusparse_extension
        plan = structure.Plan(n=self.num_nodes, depth=ext.depth,
                              arrays={**{k: v.cpu() for k, v in ext._dev.items()}, **ext._host},
                              stats=dict(ext.plan_stats))
        storage.save(path, n=self.num_nodes, alpha=self.alpha, delta_crow=self.deltas.crow_indices(),
                     delta_col=self.deltas.col_indices(), delta_sign=torch.sign(self.deltas.values()),
                     mca_offset=self.mca_offset, mca_src_idx=self.mca_src_idx, mca_dst_idx=self.mca_dst_idx,
                     plan=plan, build_stats=self.build_stats, saved_by=type(self).__name__,
                     gather_order=getattr(ext, "_order", None))

    @classmethod
    def load(cls, path, *, device="cuda", gather_order=None):
        """A matrix of this class from a structure file (any ``cbm4*`` class's
        file: the structure does not depend on the product).  The file's
        gather order is used when it has one; otherwise ``gather_order``
        decides as in the constructor (False: none either way)."""
        st = storage.load(path)
        obj = cls.__new__(cls)
        obj.alpha = st.alpha
        if gather_order is not False:
            obj._stored_order = st.gather_order
            obj._order_wanted = st.n > ORDER_MIN_ROWS if gather_order is None else True
        obj._assemble(st.n, st.delta_crow, st.delta_col, st.delta_values(), st.mca_offset, st.mca_src_idx,
                      st.mca_dst_idx, st.build_stats, device, 0, plan=st.plan)
        return obj

    def matmul(self, x, y):
        self.cusparse_extension.matmul_AX_ADX(x, y)

    def apply(self, x):
        """``M @ x`` as a new FP32 tensor, differentiable w.r.t. ``x``."""
        return cbm_apply(self.cusparse_extension, x)

    __matmul__ = apply

    def matmul_t(self, g, out):
        """``out = M^T g`` (the adjoint)."""
        self.cusparse_extension.matmul_t(g, out)

    def _degree_scale(self):
        """D^-1/2 from the row degrees d = A 1, computed with the CBM product
        itself at F = 1 as in cbm/cbm4ad.py:12-19 (exact integers in FP32);
        the power is taken in FP64 and rounded once to FP32.  Degree-0 rows get
        0 instead of the reference's inf, so fused scalings never form 0 * inf."""
        n = self.num_nodes
        dev = self.deltas.device
        d = torch.zeros(n, 1, dtype=torch.float32, device=dev)
        x = torch.ones(n, 1, dtype=torch.float32, device=dev)
        cbm4mm.matmul(self, x, d)
        d = d.view(-1).double()
        return torch.where(d > 0, d.clamp_min(1.0).pow(-0.5), torch.zeros_like(d)).float()
