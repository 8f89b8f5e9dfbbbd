// C ABI of the operator surface (include/cbm_hip.h, "operator surface"):
// the eight s_*_csr_int32 ops of cbm/cbm_extensions.cpp:351-907, which take
// the CBM pieces as loose arrays (MKL 4-array CSR with values, branch and
// tree-edge arrays).  Here they run on device arrays with the HIP kernels.
#include "cbm_kernels.hpp"
#include "../../../include/cbm_hip.h"

#include <cstdarg>
#include <cstdio>
#include <string>

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

}  // namespace

extern "C" const char *cbm_op_last_error(void) { return g_error.c_str(); }

extern "C" int cbm_op_spmm(int32_t n_rows, const int32_t *row_b, const int32_t *row_e, const int32_t *col,
                           const float *val, const float *rhs, int64_t ldr, float *dst, int64_t ldd, int32_t n_cols,
                           float self_coef, int32_t add_self, hipStream_t stream) {
    if (n_rows < 0 || n_cols < 0 || !row_b || !rhs || !dst || ldr < n_cols || ldd < n_cols) {
        set_error("cbm_op_spmm: invalid arguments");
        return CBM_ERR_ARG;
    }
    if (n_rows == 0 || n_cols == 0) return CBM_OK;
    cbmk::RowPass rp;
    rp.p_begin = 0;
    rp.p_end = n_rows;
    rp.ptr = row_b;
    rp.ptr_end = row_e;
    rp.idx = col;
    rp.val = val;
    rp.g = rhs;
    rp.ldg = ldr;
    rp.y = dst;
    rp.ldy = ldd;
    rp.ncols = n_cols;
    if (add_self) {
        rp.extra = cbmk::kExtraSelf;
        rp.e = rhs;
        rp.lde = ldr;
        rp.ecoef = self_coef;
    }
    const hipError_t e = cbmk::launch_row_pass(rp, stream);
    if (e != hipSuccess) {
        set_error("cbm_op_spmm: %s", hipGetErrorString(e));
        return CBM_ERR_DEVICE;
    }
    return CBM_OK;
}

extern "C" int cbm_op_update(int32_t n_branches, const int32_t *branches, const int32_t *src, const int32_t *dst,
                             const float *mult, float *y, int64_t ldy, int32_t n_cols, hipStream_t stream) {
    if (n_branches < 0 || n_cols < 0 || !y || ldy < n_cols || (n_branches > 0 && (!branches || !src || !dst))) {
        set_error("cbm_op_update: invalid arguments");
        return CBM_ERR_ARG;
    }
    const hipError_t e = cbmk::launch_branch_update(n_branches, branches, src, dst, mult, y, ldy, n_cols, stream);
    if (e != hipSuccess) {
        set_error("cbm_op_update: %s", hipGetErrorString(e));
        return CBM_ERR_DEVICE;
    }
    return CBM_OK;
}

extern "C" int cbm_op_axpy(int32_t n_rows, int32_t n_cols, float coef, const float *x, int64_t ldx, float *y,
                           int64_t ldy, hipStream_t stream) {
    if (n_rows < 0 || n_cols < 0 || !x || !y || ldx < n_cols || ldy < n_cols) {
        set_error("cbm_op_axpy: invalid arguments");
        return CBM_ERR_ARG;
    }
    const hipError_t e = cbmk::launch_axpy(n_rows, n_cols, coef, x, ldx, y, ldy, stream);
    if (e != hipSuccess) {
        set_error("cbm_op_axpy: %s", hipGetErrorString(e));
        return CBM_ERR_DEVICE;
    }
    return CBM_OK;
}
