"""MI355X counterpart of the reference extension module ``cbm_cusparse``
(cbm/cbm.cu:302-310): class :class:`cusparseCBM` with the same constructor
and ``matmul_AX_ADX`` / ``matmul_DAX_DADX`` methods, backed by the HIP kernels
of libcbm_hip.so instead of cuSPARSE + the ``update`` kernels.

Additions that leave the reference surface intact: BF16 right-hand sides,
the adjoint :meth:`cusparseCBM.matmul_t`, and the two halves of the forward
product (:meth:`delta_spmm`, :meth:`propagate`) for separate timing.
"""
import ctypes

import torch

from . import _lib, structure
from ._hip_sigs import BF16, DEVICE_ARRAYS, F32, HOST_ARRAYS, MODE_AX, MODE_ADX, MODE_DADX, CbmDesc

_MODE_NAMES = {MODE_AX: "A", MODE_ADX: "A D^-1/2", MODE_DADX: "D^-1/2 A D^-1/2"}


def _require_cuda(t):
    # Same message as the reference's TORCH_CHECK (cbm/cbm.cu:101-108,151-152).
    if not t.is_cuda:
        raise RuntimeError("tensor is not a CUDA tensor")


_raw_stream = getattr(torch._C, "_cuda_getCurrentRawStream", None)


def current_stream_handle(device=None):
    """torch's current stream on ``device`` as a HIP stream handle.  Where
    torch offers the raw handle it is read directly (about 0.1 us instead of
    2 us for building a Stream object: the product calls are a few
    microseconds each on small graphs)."""
    if _raw_stream is not None:
        if isinstance(device, torch.device) and device.index is not None:
            index = device.index
        elif isinstance(device, int):
            index = device
        else:
            index = torch.cuda.current_device()
        return ctypes.c_void_p(_raw_stream(index))
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _dtype_code(t):
    if t.dtype == torch.float32:
        return F32
    if t.dtype == torch.bfloat16:
        return BF16
    raise TypeError(f"right-hand side must be float32 or bfloat16, got {t.dtype}")


def infer_scaling(n, crow, col, val, multipliers):
    """Reads the product a reference-style ``deltas`` tensor encodes.

    * ``multipliers`` non-empty: values are +-D[row]*D[col]   -> D A D  (cbm/cbm4dad.py:25-34)
    * all values +-1:                                          -> A      (cbm/cbm4mm.py)
    * values +-c[col], one magnitude per column:               -> A D    (cbm/cbm4ad.py:25)
    Anything else is not a CBM product and raises ``ValueError``.
    """
    mag = val.abs()
    if multipliers is not None and multipliers.numel():
        d = multipliers.reshape(-1).to(torch.float32)
        if d.numel() != n:
            raise ValueError(f"multipliers has {d.numel()} entries, expected {n}")
        rows = torch.repeat_interleave(torch.arange(n, device=crow.device), crow.diff().long())
        expect = d[rows] * d[col.long()]
        if mag.numel() and not torch.allclose(mag, expect, rtol=1e-5, atol=0.0):
            raise ValueError("deltas values are not +-multipliers[row]*multipliers[col]")
        return MODE_DADX, d
    if mag.numel() == 0 or bool((mag == 1).all()):
        return MODE_AX, None
    colv = col.long()
    hi = torch.zeros(n, dtype=torch.float32, device=mag.device).scatter_reduce(0, colv, mag, "amax",
                                                                               include_self=False)
    lo = torch.full((n,), float("inf"), dtype=torch.float32, device=mag.device).scatter_reduce(
        0, colv, mag, "amin", include_self=False)
    used = torch.bincount(colv, minlength=n) > 0
    if not torch.equal(hi[used], lo[used]):
        raise ValueError("deltas values are neither +-1, +-scale[col] nor +-multipliers[row]*multipliers[col]")
    return MODE_ADX, hi


class cusparseCBM:
    """CBM matrix on the GPU.  Same constructor as cbm/cbm.cu:95-142:
    ``cusparseCBM(deltas, mca_branches, mca_src_node_idx, mca_dst_node_idx, multipliers)``
    where ``deltas`` is a CUDA sparse CSR tensor (int32 indices, float32 values)
    and ``multipliers`` is D^-1/2 for the D A D product or an empty tensor.

    Keyword-only ``mode`` (0: A, 1: A D, 2: D A D) and ``scale`` (D^-1/2) skip
    the inference of the product from the values; ``plan`` (a
    :class:`cbm.structure.Plan` of this matrix, e.g. from a structure file)
    skips the layout build.  The handle owns device copies of its layout;
    nothing is allocated per call.
    """

    def __init__(self, deltas, mca_branches, mca_src_node_idx, mca_dst_node_idx, multipliers, *, mode=None,
                 scale=None, n_threads=0, plan=None):
        for t in (deltas, mca_branches, mca_src_node_idx, mca_dst_node_idx):
            _require_cuda(t)
        if multipliers is not None and multipliers.numel():
            _require_cuda(multipliers)
        if deltas.layout != torch.sparse_csr:
            raise TypeError("deltas must be a sparse CSR tensor")
        n = int(deltas.size(0))
        if deltas.dim() != 2 or int(deltas.size(1)) != n:
            raise ValueError(f"deltas must be square, got {tuple(deltas.shape)}")
        crow, col, val = deltas.crow_indices(), deltas.col_indices(), deltas.values()
        for t, dt, name in ((crow, torch.int32, "crow_indices"), (col, torch.int32, "col_indices"),
                            (val, torch.float32, "values"), (mca_src_node_idx, torch.int32, "mca_src_node_idx"),
                            (mca_dst_node_idx, torch.int32, "mca_dst_node_idx")):
            if t.dtype != dt:
                raise TypeError(f'"{name}" is not a tensor of type "{dt}"')
        if mode is None:
            mode, scale = infer_scaling(n, crow, col, val, multipliers)
        elif mode != MODE_AX and scale is None:
            raise ValueError(f"mode {mode} needs the D^-1/2 vector as `scale`")
        self.device = deltas.device
        if plan is None:
            plan = structure.plan(n, crow.cpu(), col.cpu(), val.cpu(), mca_src_node_idx.cpu(),
                                  mca_dst_node_idx.cpu(), n_threads=n_threads)
        elif (plan.n != n or int(plan.lvl_node.numel()) != n or int(plan.fwd_col.numel()) != int(col.numel())
              or int(plan.ch_node.numel()) != int(mca_src_node_idx.numel())):
            raise ValueError("the given layout does not belong to this matrix")
        self.num_rows = n
        self.depth = plan.depth
        self.nnz = int(col.numel())
        self.n_edges = int(mca_src_node_idx.numel())
        self.level_sizes = plan.lvl_ptr.diff().tolist()
        self.plan_stats = dict(plan.stats)
        self._host = {k: plan.arrays[k].contiguous() for k in HOST_ARRAYS}
        self._dev = {k: plan.arrays[k].to(self.device) for k in DEVICE_ARRAYS}
        self._lib = _lib.hip()
        self._symmetric = False
        self._plain = None
        self._plain_mode = -1
        self._bind(mode, scale)

    def _bind(self, mode, scale):
        """Creates the native handle for product `mode` over this layout."""
        self.mode = int(mode)
        self._d = None if self.mode == MODE_AX else scale.to(self.device, torch.float32).contiguous()
        ptrs = {k: (v.data_ptr() or None) for k, v in self._dev.items()}
        ptrs.update({k + "_host": v.data_ptr() for k, v in self._host.items()})
        st = self.plan_stats
        desc = CbmDesc(n_rows=self.num_rows, n_levels=self.depth, nnz=self.nnz, n_edges=self.n_edges,
                       mode=self.mode, fwd_slots=st["fwd_slots"], bwd_nlong=st["bwd_nlong"],
                       n_bck=int(self._dev["bck_pos"].numel()), n_bred=int(self._dev["bred_pos"].numel()),
                       bwd_slots=st["bwd_slots"], fwd_gslots=st["fwd_gslots"], rev_slots=st["rev_slots"],
                       fwd_pieces=st.get("fwd_pieces", 0), bwd_pieces=st.get("bwd_pieces", 0),
                       fwd_pieces4=st.get("fwd_pieces4", 0), bwd_pieces4=st.get("bwd_pieces4", 0),
                       tf_level=st.get("tf_level", self.depth), tr_level=st.get("tr_level", self.depth),
                       tf_frags=max(0, int(self._dev["tf_frag"].numel()) - 1),
                       tr_frags=max(0, int(self._dev["tr_frag"].numel()) - 1),
                       n_units=max(0, int(self._dev["un_ptr"].numel()) - 1),
                       unit_levels=int(st.get("unit_levels", 0)),
                       d=self._d.data_ptr() if self._d is not None else None, **ptrs)
        self._reserved = 0
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            rc = self._lib.cbm_create(ctypes.byref(desc), current_stream_handle(self.device), ctypes.byref(h))
        _lib.check_hip(rc, "cusparseCBM")
        self._h = h
        if self._symmetric:
            _lib.check_hip(self._lib.cbm_set_symmetric(self._h, 1), "set_symmetric")
        self._bind_order()

    def with_scaling(self, mode, scale=None):
        """A handle for another product (A, A D or D A D) over the same device
        layout; the layout arrays are shared, not copied."""
        other = object.__new__(type(self))
        for k in ("device", "num_rows", "depth", "nnz", "n_edges", "level_sizes", "plan_stats", "_host", "_dev",
                  "_lib", "_symmetric", "_plain_mode"):
            setattr(other, k, getattr(self, k))
        for k in ("_order", "_order_dev", "_order_mode"):
            if hasattr(self, k):
                setattr(other, k, getattr(self, k))
        if mode != MODE_AX and scale is None:
            raise ValueError(f"mode {mode} needs the D^-1/2 vector as `scale`")
        other._plain = None
        other._bind(mode, scale)
        if self._plain is not None:
            other.attach_uncompressed(self._plain.with_scaling(mode, scale), self._plain_mode)
        return other

    def attach_uncompressed(self, plain, mode=None):
        """Attaches ``plain``, a :class:`cusparseCBM` of the same product over
        A itself (no tree), and lets the product calls run over it: ``mode``
        None (automatic, by the launch/bytes cost model of
        cbm_attach_uncompressed), True always, False never."""
        code = -1 if mode is None or mode == -1 else int(bool(mode))
        if plain is not None and (plain.num_rows != self.num_rows or plain.mode != self.mode):
            raise ValueError("the uncompressed handle must have the same rows and product")
        _lib.check_hip(self._lib.cbm_attach_uncompressed(self._h, plain._h if plain is not None else None, code),
                       "attach_uncompressed")
        self._plain, self._plain_mode = plain, code
        if plain is not None:
            plain._symmetric = self._symmetric
            plain._reserved = max(plain._reserved, self._reserved)

    def attach_order(self, order=None, mode=None):
        """Runs the Delta pass over the short rows in a gather order
        (``cbm_attach_order``): ``order`` the positions of this layout's short
        rows (e.g. from a structure file), or None to compute it (the partner
        scan on this GPU, then the forest walk on the host).  ``mode``: None
        (automatic), True, False.  Returns the order (int32, host)."""
        short = self._dev["fwd_ptr"].diff() <= structure.LONG_ROW
        if order is None:
            order = structure.device_gather_order(self._dev["fwd_ptr"], self._dev["fwd_col"], self.device)
        else:
            order = order.to("cpu", torch.int32).contiguous()
            n_short = int(short.sum())
            if (order.numel() != n_short or (n_short and (int(order.min()) < 0 or int(order.max()) >= self.num_rows))
                    or not bool(short.cpu()[order.long()].all()) or int(torch.unique(order).numel()) != n_short):
                raise ValueError("the given order is not an order of this matrix's short rows")
        self._order = order
        self._order_dev = (order if order.numel() else torch.zeros(1, dtype=torch.int32)).to(self.device)
        self._order_mode = -1 if mode is None else int(bool(mode))
        self._bind_order()
        return order

    def _bind_order(self):
        order = getattr(self, "_order", None)
        if order is None:
            return
        with torch.cuda.device(self.device):
            _lib.check_hip(self._lib.cbm_attach_order(self._h, int(order.numel()), self._order_dev.data_ptr(),
                                                      self._order_mode, current_stream_handle(self.device)),
                           "attach_order")

    def set_order(self, mode):
        """Switches the gather order of the Delta pass: True, False or None
        (automatic); an attached uncompressed handle's own order too."""
        self._order_mode = -1 if mode is None else int(bool(mode))
        _lib.check_hip(self._lib.cbm_set_order(self._h, self._order_mode), "set_order")
        plain = getattr(self, "_plain", None)
        if plain is not None and getattr(plain, "_order", None) is not None:
            plain.set_order(mode)

    def order_active(self, dtype, n_cols):
        """Whether a forward product with dtype storage and n_cols columns
        runs its Delta pass in the gather order now."""
        code = self._lib.cbm_order_active(self._h, BF16 if dtype == torch.bfloat16 else F32, int(n_cols))
        _lib.check_hip(min(code, 0), "order_active")
        return bool(code)

    def uncompressed_active(self, dtype, n_cols):
        """Whether a product with dtype storage and n_cols columns runs over
        the attached uncompressed handle now."""
        code = self._lib.cbm_uncompressed_active(self._h, BF16 if dtype == torch.bfloat16 else F32, int(n_cols))
        _lib.check_hip(min(code, 0), "uncompressed_active")
        return bool(code)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.cbm_destroy(h)  # before the attached uncompressed handle it borrows
            self._h = None

    @property
    def D(self):
        """D^-1/2 (FP32, zero for degree-0 nodes) or None for the plain product."""
        return self._d

    def describe(self):
        return {"n": self.num_rows, "nnz_delta": self.nnz, "tree_edges": self.n_edges, "depth": self.depth,
                "product": _MODE_NAMES[self.mode]}

    # ---------------------------------------------------------------- checks
    def _check_in(self, x):
        _require_cuda(x)
        if x.dim() != 2 or int(x.size(0)) != self.num_rows:
            raise ValueError(f"operand must be {self.num_rows} x F, got {tuple(x.shape)}")
        if x.get_device() != self.device.index:
            raise ValueError(f"operand is on {x.device}, the matrix on {self.device}")
        code = _dtype_code(x)
        if x.stride(1) != 1 or x.stride(0) < x.size(1):
            x = x.contiguous()
        return x, code

    def _check_out(self, y, n_cols, name="res"):
        _require_cuda(y)
        if y.dtype != torch.float32:
            raise TypeError(f'"{name}" is not a tensor of type "torch.float32"')
        if y.dim() != 2 or tuple(y.shape) != (self.num_rows, n_cols):
            raise ValueError(f"{name} must be {self.num_rows} x {n_cols}, got {tuple(y.shape)}")
        if y.get_device() != self.device.index:
            raise ValueError(f"{name} is on {y.device}, the matrix on {self.device}")
        if y.stride(1) != 1 or y.stride(0) < n_cols:
            raise ValueError(f"{name} must have unit column stride")

    def _operands(self, x, y, name):
        """(x, dtype code, F, ld of x, ld of y) after the checks of _check_in
        and _check_out, read in one pass over the tensors' attributes when
        they hold (a product call on a small graph takes a few microseconds,
        so its host-side checks are a large part of it); otherwise the full
        checks, which raise or make x contiguous."""
        xs, ys = x.shape, y.shape
        xst, yst = x.stride(), y.stride()
        dt = x.dtype
        idx = self.device.index
        if (x.is_cuda and y.is_cuda and len(xs) == 2 and xs[0] == self.num_rows and ys == xs
                and (dt is torch.float32 or dt is torch.bfloat16) and y.dtype is torch.float32
                and xst[1] == 1 and xst[0] >= xs[1] and yst[1] == 1 and yst[0] >= ys[1]
                and x.get_device() == idx and y.get_device() == idx):
            return x, (F32 if dt is torch.float32 else BF16), xs[1], xst[0], yst[0]
        x, code = self._check_in(x)
        f = int(x.size(1))
        self._check_out(y, f, name)
        return x, code, f, x.stride(0), y.stride(0)

    def _stream(self):
        """torch's current stream on this matrix's device, as a HIP stream
        handle (an int, which ctypes passes as a pointer)."""
        if _raw_stream is not None:
            return _raw_stream(self.device.index)
        return current_stream_handle(self.device)

    def _ensure_width(self, n_cols):
        """Reserves the native scratch for n_cols columns if the handle serves
        fewer (the native product calls never allocate).  Under HIP-graph
        capture that allocation is not allowed: reserve before capturing."""
        if n_cols <= self._reserved:
            return
        if torch.cuda.is_current_stream_capturing():
            raise RuntimeError(f"CBM product of {n_cols} columns under graph capture: call reserve({n_cols}) "
                               f"before capturing (reserved: {self._reserved})")
        self.reserve(n_cols)

    # --------------------------------------------------------------- products
    def forward(self, rhs, res):
        """res = M rhs for this handle's product M (A, A D or D A D)."""
        x, code, f, ldx, ldy = self._operands(rhs, res, "res")
        if f == 0:
            return 0
        if f > self._reserved:
            self._ensure_width(f)
        rc = self._lib.cbm_matmul(self._h, code, x.data_ptr(), ldx, res.data_ptr(), ldy, f, self._stream())
        if rc != _lib.CBM_OK:
            _lib.check_hip(rc, "matmul")
        return 0

    def matmul_AX_ADX(self, rhs, res):
        """res = A rhs or A D^-1/2 rhs (cbm/cbm.cu:150-207)."""
        _require_cuda(rhs)
        _require_cuda(res)
        if self.mode == MODE_DADX:
            raise ValueError("this matrix carries D^-1/2 A D^-1/2 scalings; use matmul_DAX_DADX")
        return self.forward(rhs, res)

    def matmul_DAX_DADX(self, rhs, res):
        """res = D^-1/2 A D^-1/2 rhs (cbm/cbm.cu:210-268)."""
        _require_cuda(rhs)
        _require_cuda(res)
        if self.mode != MODE_DADX:
            raise ValueError("this matrix was built without multipliers; use matmul_AX_ADX")
        return self.forward(rhs, res)

    def matmul_t(self, grad, out, workspace=None):
        """out = M^T grad.  ``workspace`` (FP32, N x F) is allocated if not given."""
        g, code, f, ldg, ldo = self._operands(grad, out, "out")
        if f == 0:
            return 0
        if workspace is not None:
            self._check_out(workspace, f, "workspace")
        elif not self._symmetric:  # the forward-structure adjoint needs none
            workspace = torch.empty((self.num_rows, f), dtype=torch.float32, device=self.device)
        if f > self._reserved:
            self._ensure_width(f)
        ws_ptr, ws_ld = (workspace.data_ptr(), workspace.stride(0)) if workspace is not None else (None, f)
        rc = self._lib.cbm_matmul_t(self._h, code, g.data_ptr(), ldg, out.data_ptr(), ldo, ws_ptr, ws_ld, f,
                                    self._stream())
        if rc != _lib.CBM_OK:
            _lib.check_hip(rc, "matmul_t")
        return 0

    def adjoint_halves(self, grad, out, workspace):
        """The adjoint in its two halves (reverse gather into `workspace`,
        then the Delta^T pass into `out`), for separate timing."""
        g, code = self._check_in(grad)
        f = int(g.size(1))
        self._check_out(out, f, "out")
        self._check_out(workspace, f, "workspace")
        self._ensure_width(f)
        st = current_stream_handle(self.device)
        _lib.check_hip(self._lib.cbm_adjoint_reverse(self._h, code, g.data_ptr(), g.stride(0), workspace.data_ptr(),
                                                     workspace.stride(0), f, st), "adjoint_reverse")
        _lib.check_hip(self._lib.cbm_adjoint_delta_t(self._h, code, g.data_ptr(), g.stride(0), workspace.data_ptr(),
                                                     workspace.stride(0), out.data_ptr(), out.stride(0), f, st),
                       "adjoint_delta_t")
        return 0

    def adjoint_reverse(self, grad, workspace):
        g, code = self._check_in(grad)
        f = int(g.size(1))
        self._check_out(workspace, f, "workspace")
        self._ensure_width(f)
        _lib.check_hip(self._lib.cbm_adjoint_reverse(self._h, code, g.data_ptr(), g.stride(0), workspace.data_ptr(),
                                                     workspace.stride(0), f, current_stream_handle(self.device)),
                       "adjoint_reverse")

    def adjoint_delta_t(self, grad, workspace, out):
        g, code = self._check_in(grad)
        f = int(g.size(1))
        self._check_out(out, f, "out")
        self._ensure_width(f)
        _lib.check_hip(self._lib.cbm_adjoint_delta_t(self._h, code, g.data_ptr(), g.stride(0), workspace.data_ptr(),
                                                     workspace.stride(0), out.data_ptr(), out.stride(0), f,
                                                     current_stream_handle(self.device)), "adjoint_delta_t")

    def delta_spmm(self, rhs, res):
        """res = Delta' rhs without the tree propagation."""
        x, code = self._check_in(rhs)
        f = int(x.size(1))
        self._check_out(res, f)
        if f == 0:
            return 0
        self._ensure_width(f)
        rc = self._lib.cbm_delta_spmm(self._h, code, x.data_ptr(), x.stride(0), res.data_ptr(), res.stride(0), f,
                                      current_stream_handle(self.device))
        _lib.check_hip(rc, "delta_spmm")
        return 0

    def propagate(self, res):
        """In place, level by level: res_v += s_v res_parent(v) (cbm/cbm.cu:25-91)."""
        f = int(res.size(1)) if res.dim() == 2 else -1
        self._check_out(res, f)
        if f == 0:
            return 0
        self._ensure_width(f)
        rc = self._lib.cbm_propagate(self._h, res.data_ptr(), res.stride(0), f, current_stream_handle(self.device))
        _lib.check_hip(rc, "propagate")
        return 0

    def reserve(self, n_cols):
        """Sizes the native scratch (long-row partial sums, compensation rows,
        XCD-range partial rows) for products of up to n_cols columns.  The
        product methods reserve a wider width by themselves outside graph
        capture; under capture the width must already be reserved."""
        _lib.check_hip(self._lib.cbm_reserve(self._h, int(n_cols)), "reserve")
        self._reserved = max(self._reserved, int(n_cols))
        if self._plain is not None:
            self._plain._reserved = max(self._plain._reserved, int(n_cols))

    @property
    def symmetric(self):
        """Whether A is declared symmetric (then matmul_t runs through the
        forward structure)."""
        return self._symmetric

    def set_symmetric(self, symmetric):
        """Declares A = A^T (True) or not (False): with True, ``matmul_t``
        computes M^T G through the forward structure (M^T = M for A and
        D A D, D A for A D) instead of the Delta^T pass.  The cbm4* classes
        set it from the structure build's exact comparison of A with A^T."""
        _lib.check_hip(self._lib.cbm_set_symmetric(self._h, int(bool(symmetric))), "set_symmetric")
        self._symmetric = bool(symmetric)
        if self._plain is not None:
            self._plain._symmetric = bool(symmetric)

    def set_fused(self, mode):
        """The forward product in one launch (units of the plan, propagation
        in LDS): True on where possible, False or None (automatic, the
        default) off -- the multi-launch forms measured faster.  Rounds each
        row in pieces, so bits differ from the multi-launch forms."""
        m = -1 if mode is None else int(bool(mode))
        _lib.check_hip(self._lib.cbm_set_fused(self._h, m), "set_fused")

    def fused_active(self, dtype, n_cols):
        """Whether a product with dtype storage and n_cols columns runs in
        one launch now (the width must be reserved)."""
        if dtype not in (torch.float32, torch.bfloat16):
            raise TypeError(f"dtype must be float32 or bfloat16, got {dtype}")
        code = self._lib.cbm_fused_active(self._h, BF16 if dtype == torch.bfloat16 else F32, int(n_cols))
        _lib.check_hip(min(code, 0), "fused_active")
        return bool(code)

    def fused_errors(self):
        """Dependency waits of the one-launch forward that hit their poll
        limit since creation (0 in a correct run); synchronises the device."""
        n = int(self._lib.cbm_fused_errors(self._h))
        _lib.check_hip(min(n, 0), "fused_errors")
        return n

    def set_kahan_depth(self, level):
        """Tree levels from `level` on propagate with an FP32 compensation
        term (default 32; None or a negative value switches it off)."""
        lvl = -1 if level is None or level < 0 else int(level)
        _lib.check_hip(self._lib.cbm_set_kahan_depth(self._h, lvl), "set_kahan_depth")

    def set_forward_split(self, mode):
        """Form of the forward product: True split (Delta pass over all levels,
        then per-level propagation), False fused per level, None automatic.
        Both give identical results."""
        m = -1 if mode is None else int(bool(mode))
        _lib.check_hip(self._lib.cbm_set_forward_split(self._h, m), "set_forward_split")

    def set_wave_rows(self, mode):
        """Row mapping of the short-row passes: True a whole wavefront per row,
        16 sixteen lanes per row (for rows narrower than that), False rows
        packed per wavefront, None by matrix size (the default)."""
        m = -1 if mode is None else 2 if (mode == 16 and not isinstance(mode, bool)) else int(bool(mode))
        _lib.check_hip(self._lib.cbm_set_wave_rows(self._h, m), "set_wave_rows")

    def set_mixed_grid(self, mode):
        """Long-row chunks and short rows of a gather pass: True one grid,
        False two launches, None by rule (the default).  Same results."""
        m = -1 if mode is None else int(bool(mode))
        _lib.check_hip(self._lib.cbm_set_mixed_grid(self._h, m), "set_mixed_grid")

    def set_chunk_finish(self, mode):
        """Rows of several long-row chunks: True finished by their last chunk
        inside the gather pass, False by a reduction launch after it, None by
        matrix size (the default).  Same results."""
        m = -1 if mode is None else int(bool(mode))
        _lib.check_hip(self._lib.cbm_set_chunk_finish(self._h, m), "set_chunk_finish")

    def set_xcd_ranges(self, mode):
        """XCD column ranges for the gather passes: None automatic (by size),
        False off, True on (when the layout has range pieces; the piece set
        by row width), 8 or 4 on with that many ranges."""
        if mode is None:
            code = -1
        elif isinstance(mode, bool):
            code = int(mode)
        elif mode in (4, 8):
            code = int(mode)
        else:
            raise ValueError(f"set_xcd_ranges: None, a bool, 4 or 8, got {mode!r}")
        _lib.check_hip(self._lib.cbm_set_xcd_ranges(self._h, code), "set_xcd_ranges")

    def set_tree_tails(self, on):
        """Tree tails in one launch each (True, the default) or per level
        (False); ``"unstaged"`` and ``"all_staged"`` keep the tails with no
        forward fragment, or every one that fits, staged in LDS (the default
        stages those of at least 8 levels).  The forward's bits do not
        change."""
        codes = {"unstaged": 2, "all_staged": 3}
        code = codes[on] if isinstance(on, str) and on in codes else None
        if code is None:
            if isinstance(on, str):
                raise ValueError(f"set_tree_tails: a bool, 'unstaged' or 'all_staged', got {on!r}")
            code = int(bool(on))
        _lib.check_hip(self._lib.cbm_set_tree_tails(self._h, code), "set_tree_tails")

    def set_xcd_phases(self, phases):
        """Column phases of the XCD range passes: None automatic, 1, 2 or 4
        (each XCD gathers its range one column slab after the other); the
        results do not change."""
        code = -1 if phases is None else int(phases)
        _lib.check_hip(self._lib.cbm_set_xcd_phases(self._h, code), "set_xcd_phases")

    def xcd_ranges_active(self, dtype, n_cols):
        """The number of XCD column ranges (8 or 4) a call with dtype
        (torch.float32 or torch.bfloat16) storage and n_cols columns runs its
        gathers with now, 0 for none."""
        if dtype not in (torch.float32, torch.bfloat16):
            raise TypeError(f"dtype must be float32 or bfloat16, got {dtype}")
        code = self._lib.cbm_xcd_ranges_active(self._h, BF16 if dtype == torch.bfloat16 else F32, int(n_cols))
        _lib.check_hip(min(code, 0), "xcd_ranges_active")
        return int(code)

    def set_xcd_slabs(self, slabs):
        """Feature slabs spread over the XCDs (1 = off); same results."""
        _lib.check_hip(self._lib.cbm_set_xcd_slabs(self._h, int(slabs)), "set_xcd_slabs")

    def set_slab_bytes(self, nbytes):
        """Column-slab size for cache residency of the gathers (0: whole rows)."""
        _lib.check_hip(self._lib.cbm_set_slab_bytes(self._h, int(nbytes)), "set_slab_bytes")

    def last_launches(self):
        return int(self._lib.cbm_last_launches(self._h))
