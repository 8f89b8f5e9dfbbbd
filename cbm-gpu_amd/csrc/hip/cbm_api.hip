// C ABI of the CBM product (include/cbm_hip.h).  Replaces class cusparseCBM
// of cbm/cbm.cu:93-300: no cuSPARSE handle, no per-call descriptors and no
// per-call workspace (cbm/cbm.cu:157-181,203-205), and every launch goes to
// the caller's stream instead of the legacy default stream (:196,256).
#include "cbm_kernels.hpp"
#include "../../../include/cbm_hip.h"

#include <cstdarg>
#include <cstdio>
#include <new>
#include <string>
#include <vector>

struct cbm_handle {
    cbm_desc d{};
    std::vector<int32_t> lvl_ptr;  // host copy of the level offsets
    float *rs_pos = nullptr;       // D A D only: d[node] per level position
    float *ps_pos = nullptr;       // D A D only: d[node] / d[parent] per level position
    int last_launches = 0;
};

namespace {

thread_local std::string g_error;

void set_error(const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_error = buf;
}

int device_error(hipError_t e, const char *what) {
    set_error("%s: %s", what, hipGetErrorString(e));
    return CBM_ERR_DEVICE;
}

bool overlaps(const void *a, int64_t a_bytes, const void *b, int64_t b_bytes) {
    const auto pa = reinterpret_cast<uintptr_t>(a), pb = reinterpret_cast<uintptr_t>(b);
    return pa < pb + uintptr_t(b_bytes) && pb < pa + uintptr_t(a_bytes);
}

int64_t span_bytes(int32_t n_rows, int64_t ld, int32_t n_cols, int elem) {
    return n_rows > 0 ? ((int64_t(n_rows) - 1) * ld + n_cols) * elem : 0;
}

int check_dense(const cbm_handle *h, int32_t dtype, const void *in, int64_t ldin, const float *out, int64_t ldout,
                int32_t n_cols, const char *what) {
    if (!h) {
        set_error("%s: null handle", what);
        return CBM_ERR_ARG;
    }
    if (dtype != CBM_F32 && dtype != CBM_BF16) {
        set_error("%s: unknown dtype %d", what, dtype);
        return CBM_ERR_ARG;
    }
    if (n_cols < 1 || ldin < n_cols || ldout < n_cols || !in || !out) {
        set_error("%s: invalid dense operands (n_cols=%d ldin=%lld ldout=%lld)", what, n_cols, (long long)ldin,
                  (long long)ldout);
        return CBM_ERR_ARG;
    }
    const int in_elem = dtype == CBM_BF16 ? 2 : 4;
    if (overlaps(in, span_bytes(h->d.n_rows, ldin, n_cols, in_elem), out, span_bytes(h->d.n_rows, ldout, n_cols, 4))) {
        set_error("%s: input and output overlap", what);
        return CBM_ERR_ARG;
    }
    return CBM_OK;
}

// Scale vectors of the forward Delta pass by mode.
const float *fwd_colscale(const cbm_handle *h) { return h->d.mode == CBM_MODE_AX ? nullptr : h->d.d; }

}  // namespace

extern "C" const char *cbm_last_error(void) { return g_error.c_str(); }

extern "C" int cbm_create(const cbm_desc *desc, hipStream_t stream, cbm_handle **out) {
    if (!desc || !out) {
        set_error("cbm_create: null argument");
        return CBM_ERR_ARG;
    }
    *out = nullptr;
    const cbm_desc &d = *desc;
    if (d.n_rows < 1 || d.n_levels < 1 || d.nnz < 0 || d.n_edges < 0 || d.n_edges >= d.n_rows || !d.lvl_ptr_host ||
        !d.lvl_node || !d.lvl_parent || !d.fwd_ptr || !d.bwd_ptr || !d.ch_ptr || (d.nnz > 0 && (!d.fwd_col || !d.bwd_row)) ||
        (d.n_edges > 0 && !d.ch_node)) {
        set_error("cbm_create: incomplete descriptor (n_rows=%d levels=%d nnz=%lld edges=%lld)", d.n_rows, d.n_levels,
                  (long long)d.nnz, (long long)d.n_edges);
        return CBM_ERR_ARG;
    }
    if (d.mode != CBM_MODE_AX && d.mode != CBM_MODE_ADX && d.mode != CBM_MODE_DADX) {
        set_error("cbm_create: unknown mode %d", d.mode);
        return CBM_ERR_ARG;
    }
    if (d.mode != CBM_MODE_AX && !d.d) {
        set_error("cbm_create: mode %d needs the D^-1/2 vector", d.mode);
        return CBM_ERR_ARG;
    }
    if (d.lvl_ptr_host[0] != 0 || d.lvl_ptr_host[d.n_levels] != d.n_rows) {
        set_error("cbm_create: level offsets run from %d to %d, expected 0 to %d", d.lvl_ptr_host[0],
                  d.lvl_ptr_host[d.n_levels], d.n_rows);
        return CBM_ERR_ARG;
    }
    for (int32_t l = 0; l < d.n_levels; ++l)
        if (d.lvl_ptr_host[l + 1] <= d.lvl_ptr_host[l]) {
            set_error("cbm_create: level %d is empty", l);
            return CBM_ERR_ARG;
        }
    cbm_handle *h = new (std::nothrow) cbm_handle;
    if (!h) {
        set_error("cbm_create: out of host memory");
        return CBM_ERR_NOMEM;
    }
    h->d = d;
    h->lvl_ptr.assign(d.lvl_ptr_host, d.lvl_ptr_host + d.n_levels + 1);
    h->d.lvl_ptr_host = h->lvl_ptr.data();
    if (d.mode == CBM_MODE_DADX) {
        hipError_t e = hipMalloc(&h->rs_pos, sizeof(float) * size_t(d.n_rows));
        if (e == hipSuccess) e = hipMalloc(&h->ps_pos, sizeof(float) * size_t(d.n_rows));
        if (e == hipSuccess)
            e = cbmk::launch_dad_scales(d.n_rows, d.lvl_node, d.lvl_parent, d.d, h->rs_pos, h->ps_pos, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        if (e != hipSuccess) {
            cbm_destroy(h);
            return device_error(e, "cbm_create");
        }
    }
    *out = h;
    return CBM_OK;
}

extern "C" int cbm_destroy(cbm_handle *h) {
    if (!h) return CBM_OK;
    hipError_t e1 = h->rs_pos ? hipFree(h->rs_pos) : hipSuccess;
    hipError_t e2 = h->ps_pos ? hipFree(h->ps_pos) : hipSuccess;
    delete h;
    if (e1 != hipSuccess) return device_error(e1, "cbm_destroy");
    if (e2 != hipSuccess) return device_error(e2, "cbm_destroy");
    return CBM_OK;
}

extern "C" int cbm_matmul(cbm_handle *h, int32_t dtype, const void *x, int64_t ldx, float *y, int64_t ldy,
                          int32_t n_cols, hipStream_t stream) {
    int rc = check_dense(h, dtype, x, ldx, y, ldy, n_cols, "cbm_matmul");
    if (rc != CBM_OK) return rc;
    cbmk::RowPass rp;
    rp.node_of = h->d.lvl_node;
    rp.parent_of = h->d.lvl_parent;
    rp.ptr = h->d.fwd_ptr;
    rp.idx = h->d.fwd_col;
    rp.colscale = fwd_colscale(h);
    rp.rowscale = h->d.mode == CBM_MODE_DADX ? h->rs_pos : nullptr;
    rp.escale = h->d.mode == CBM_MODE_DADX ? h->ps_pos : nullptr;
    rp.g = x;
    rp.ldg = ldx;
    rp.g_bf16 = dtype == CBM_BF16;
    rp.e = y;
    rp.lde = ldy;
    rp.e_bf16 = false;
    rp.y = y;
    rp.ldy = ldy;
    rp.ncols = n_cols;
    h->last_launches = 0;
    for (int32_t l = 0; l < h->d.n_levels; ++l) {
        rp.p_begin = h->lvl_ptr[size_t(l)];
        rp.p_end = h->lvl_ptr[size_t(l) + 1];
        rp.extra = l == 0 ? cbmk::kExtraNone : cbmk::kExtraParent;
        const hipError_t e = cbmk::launch_row_pass(rp, stream);
        if (e != hipSuccess) return device_error(e, "cbm_matmul");
        ++h->last_launches;
    }
    return CBM_OK;
}

extern "C" int cbm_matmul_t(cbm_handle *h, int32_t dtype, const void *g, int64_t ldg, float *dx, int64_t lddx,
                            float *z, int64_t ldz, int32_t n_cols, hipStream_t stream) {
    int rc = check_dense(h, dtype, g, ldg, dx, lddx, n_cols, "cbm_matmul_t");
    if (rc != CBM_OK) return rc;
    if (!z || ldz < n_cols) {
        set_error("cbm_matmul_t: invalid workspace (ldz=%lld)", (long long)ldz);
        return CBM_ERR_ARG;
    }
    const int64_t n = h->d.n_rows;
    if (overlaps(z, span_bytes(int32_t(n), ldz, n_cols, 4), dx, span_bytes(int32_t(n), lddx, n_cols, 4)) ||
        overlaps(z, span_bytes(int32_t(n), ldz, n_cols, 4), g,
                 span_bytes(int32_t(n), ldg, n_cols, dtype == CBM_BF16 ? 2 : 4))) {
        set_error("cbm_matmul_t: workspace overlaps an operand");
        return CBM_ERR_ARG;
    }
    h->last_launches = 0;
    // Reverse pass, deepest level first: z_u = s_u g_u + sum_{children v} z_v
    // with s = D^-1/2 for D A D (the adjoint's (I-P)^-T applied to D G).
    cbmk::RowPass rv;
    rv.node_of = h->d.lvl_node;
    rv.ptr = h->d.ch_ptr;
    rv.idx = h->d.ch_node;
    rv.g = z;
    rv.ldg = ldz;
    rv.g_bf16 = false;
    rv.e = g;
    rv.lde = ldg;
    rv.e_bf16 = dtype == CBM_BF16;
    rv.escale = h->d.mode == CBM_MODE_DADX ? h->d.d : nullptr;
    rv.extra = cbmk::kExtraSelf;
    rv.y = z;
    rv.ldy = ldz;
    rv.ncols = n_cols;
    for (int32_t l = h->d.n_levels - 1; l >= 0; --l) {
        rv.p_begin = h->lvl_ptr[size_t(l)];
        rv.p_end = h->lvl_ptr[size_t(l) + 1];
        const hipError_t e = cbmk::launch_row_pass(rv, stream);
        if (e != hipSuccess) return device_error(e, "cbm_matmul_t (reverse pass)");
        ++h->last_launches;
    }
    // Delta^T pass: dx_j = s_j * sum_v sign * z_v over column j of Delta.
    cbmk::RowPass tp;
    tp.p_begin = 0;
    tp.p_end = int32_t(n);
    tp.ptr = h->d.bwd_ptr;
    tp.idx = h->d.bwd_row;
    tp.rowscale = h->d.mode == CBM_MODE_AX ? nullptr : h->d.d;
    tp.g = z;
    tp.ldg = ldz;
    tp.y = dx;
    tp.ldy = lddx;
    tp.ncols = n_cols;
    const hipError_t e = cbmk::launch_row_pass(tp, stream);
    if (e != hipSuccess) return device_error(e, "cbm_matmul_t (Delta^T pass)");
    ++h->last_launches;
    return CBM_OK;
}

extern "C" int cbm_delta_spmm(cbm_handle *h, int32_t dtype, const void *x, int64_t ldx, float *y, int64_t ldy,
                              int32_t n_cols, hipStream_t stream) {
    int rc = check_dense(h, dtype, x, ldx, y, ldy, n_cols, "cbm_delta_spmm");
    if (rc != CBM_OK) return rc;
    cbmk::RowPass rp;
    rp.p_begin = 0;
    rp.p_end = h->d.n_rows;
    rp.node_of = h->d.lvl_node;
    rp.ptr = h->d.fwd_ptr;
    rp.idx = h->d.fwd_col;
    rp.colscale = fwd_colscale(h);
    rp.rowscale = h->d.mode == CBM_MODE_DADX ? h->rs_pos : nullptr;
    rp.g = x;
    rp.ldg = ldx;
    rp.g_bf16 = dtype == CBM_BF16;
    rp.y = y;
    rp.ldy = ldy;
    rp.ncols = n_cols;
    const hipError_t e = cbmk::launch_row_pass(rp, stream);
    if (e != hipSuccess) return device_error(e, "cbm_delta_spmm");
    h->last_launches = 1;
    return CBM_OK;
}

extern "C" int cbm_propagate(cbm_handle *h, float *y, int64_t ldy, int32_t n_cols, hipStream_t stream) {
    if (!h || !y || n_cols < 1 || ldy < n_cols) {
        set_error("cbm_propagate: invalid arguments");
        return CBM_ERR_ARG;
    }
    h->last_launches = 0;
    const float *ps = h->d.mode == CBM_MODE_DADX ? h->ps_pos : nullptr;
    for (int32_t l = 1; l < h->d.n_levels; ++l) {
        const hipError_t e = cbmk::launch_propagate(h->lvl_ptr[size_t(l)], h->lvl_ptr[size_t(l) + 1], h->d.lvl_node,
                                                    h->d.lvl_parent, ps, y, ldy, n_cols, stream);
        if (e != hipSuccess) return device_error(e, "cbm_propagate");
        ++h->last_launches;
    }
    return CBM_OK;
}

extern "C" int cbm_last_launches(const cbm_handle *h) { return h ? h->last_launches : -1; }
