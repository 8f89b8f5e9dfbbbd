// CSR comparator on the same GPU: rocSPARSE SpMM (include/cbm_hip.h).
// Replaces class cusparseCSR (cbm/csr.cu:24-126).  Unlike the reference,
// the work buffer and the preprocessing are kept across calls (the reference
// mallocs and frees a buffer on every call, cbm/csr.cu:76-101).
#include "../../../include/cbm_hip.h"

#include <rocsparse/rocsparse.h>

#include <cstdarg>
#include <cstdio>
#include <new>
#include <string>

struct csr_handle {
    int32_t n_rows = 0, n_cols = 0;
    int64_t nnz = 0;
    rocsparse_handle sp = nullptr;
    rocsparse_spmat_descr a = nullptr;
    rocsparse_spmm_alg alg = rocsparse_spmm_alg_default;
    void *buffer = nullptr;
    size_t buffer_bytes = 0;
    // Shape the current preprocessing was done for.
    int32_t pre_trans = -1, pre_cols = -1;
    int64_t pre_ldx = -1, pre_ldy = -1;
    int32_t pre_alg = -1;
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

int sparse_error(rocsparse_status s, const char *what) {
    set_error("%s: rocSPARSE status %d", what, int(s));
    return CBM_ERR_DEVICE;
}

}  // namespace

extern "C" const char *csr_last_error(void) { return g_error.c_str(); }

extern "C" int csr_create(int32_t n_rows, int32_t n_cols, int64_t nnz, const int32_t *crow, const int32_t *col,
                          const float *val, csr_handle **out) {
    if (!out || n_rows < 1 || n_cols < 1 || nnz < 0 || !crow || (nnz > 0 && (!col || !val))) {
        set_error("csr_create: invalid arguments");
        return CBM_ERR_ARG;
    }
    *out = nullptr;
    csr_handle *h = new (std::nothrow) csr_handle;
    if (!h) {
        set_error("csr_create: out of host memory");
        return CBM_ERR_NOMEM;
    }
    h->n_rows = n_rows;
    h->n_cols = n_cols;
    h->nnz = nnz;
    rocsparse_status s = rocsparse_create_handle(&h->sp);
    if (s == rocsparse_status_success)
        s = rocsparse_create_csr_descr(&h->a, n_rows, n_cols, nnz, const_cast<int32_t *>(crow),
                                       const_cast<int32_t *>(col), const_cast<float *>(val), rocsparse_indextype_i32,
                                       rocsparse_indextype_i32, rocsparse_index_base_zero, rocsparse_datatype_f32_r);
    if (s != rocsparse_status_success) {
        csr_destroy(h);
        return sparse_error(s, "csr_create");
    }
    *out = h;
    return CBM_OK;
}

extern "C" int csr_destroy(csr_handle *h) {
    if (!h) return CBM_OK;
    if (h->buffer) (void)hipFree(h->buffer);
    if (h->a) rocsparse_destroy_spmat_descr(h->a);
    if (h->sp) rocsparse_destroy_handle(h->sp);
    delete h;
    return CBM_OK;
}

// Selects the rocSPARSE SpMM algorithm (a rocsparse_spmm_alg value).
extern "C" int csr_set_algorithm(csr_handle *h, int32_t alg) {
    if (!h) {
        set_error("csr_set_algorithm: null handle");
        return CBM_ERR_ARG;
    }
    h->alg = rocsparse_spmm_alg(alg);
    h->pre_alg = -1;
    return CBM_OK;
}

extern "C" int csr_matmul(csr_handle *h, int32_t trans, const float *x, int64_t ldx, float *y, int64_t ldy,
                          int32_t n_cols, hipStream_t stream) {
    if (!h || !x || !y || n_cols < 1 || ldx < n_cols || ldy < n_cols || (trans != 0 && trans != 1)) {
        set_error("csr_matmul: invalid arguments");
        return CBM_ERR_ARG;
    }
    const int64_t rows_x = trans ? h->n_rows : h->n_cols;
    const int64_t rows_y = trans ? h->n_cols : h->n_rows;
    const rocsparse_operation op = trans ? rocsparse_operation_transpose : rocsparse_operation_none;
    rocsparse_dnmat_descr bx = nullptr, cy = nullptr;
    rocsparse_status s = rocsparse_set_stream(h->sp, stream);
    if (s == rocsparse_status_success)
        s = rocsparse_create_dnmat_descr(&bx, rows_x, n_cols, ldx, const_cast<float *>(x), rocsparse_datatype_f32_r,
                                         rocsparse_order_row);
    if (s == rocsparse_status_success)
        s = rocsparse_create_dnmat_descr(&cy, rows_y, n_cols, ldy, y, rocsparse_datatype_f32_r, rocsparse_order_row);
    const float one = 1.0f, zero = 0.0f;
    int rc = CBM_OK;
    if (s == rocsparse_status_success &&
        (h->pre_trans != trans || h->pre_cols != n_cols || h->pre_ldx != ldx || h->pre_ldy != ldy ||
         h->pre_alg != int32_t(h->alg))) {
        size_t bytes = 0;
        s = rocsparse_spmm(h->sp, op, rocsparse_operation_none, &one, h->a, bx, &zero, cy, rocsparse_datatype_f32_r,
                           h->alg, rocsparse_spmm_stage_buffer_size, &bytes, nullptr);
        if (s == rocsparse_status_success && bytes > h->buffer_bytes) {
            if (h->buffer) (void)hipFree(h->buffer);
            h->buffer = nullptr;
            h->buffer_bytes = 0;
            const hipError_t e = hipMalloc(&h->buffer, bytes);
            if (e != hipSuccess) {
                set_error("csr_matmul: hipMalloc of %zu bytes: %s", bytes, hipGetErrorString(e));
                rc = CBM_ERR_DEVICE;
            } else {
                h->buffer_bytes = bytes;
            }
        }
        if (rc == CBM_OK && s == rocsparse_status_success)
            s = rocsparse_spmm(h->sp, op, rocsparse_operation_none, &one, h->a, bx, &zero, cy,
                               rocsparse_datatype_f32_r, h->alg, rocsparse_spmm_stage_preprocess, &bytes, h->buffer);
        if (rc == CBM_OK && s == rocsparse_status_success) {
            h->pre_trans = trans;
            h->pre_cols = n_cols;
            h->pre_ldx = ldx;
            h->pre_ldy = ldy;
            h->pre_alg = int32_t(h->alg);
        }
    }
    if (rc == CBM_OK && s == rocsparse_status_success) {
        size_t bytes = h->buffer_bytes;
        s = rocsparse_spmm(h->sp, op, rocsparse_operation_none, &one, h->a, bx, &zero, cy, rocsparse_datatype_f32_r,
                           h->alg, rocsparse_spmm_stage_compute, &bytes, h->buffer);
    }
    if (bx) rocsparse_destroy_dnmat_descr(bx);
    if (cy) rocsparse_destroy_dnmat_descr(cy);
    if (rc != CBM_OK) return rc;
    if (s != rocsparse_status_success) return sparse_error(s, "csr_matmul");
    return CBM_OK;
}
