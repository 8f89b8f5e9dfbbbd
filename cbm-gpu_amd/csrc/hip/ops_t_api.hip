// C ABI of the transposed operator, s_spmm_csr_t_int32 of
// cbm/cbm_extensions.cpp:742-782 (dst = A^T rhs with A given by rows in the
// MKL 4-array CSR).  A^T is built on the device -- keys (column, row) of
// every entry sorted by a stable radix sort, so each column lists its rows in
// ascending order and the result does not depend on scheduling -- and then
// summed by the row-pass kernel, one output row per column of A.
#include "cbm_kernels.hpp"
#include "../../../include/cbm_hip.h"

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

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

__global__ __launch_bounds__(256) void k_row_lengths(int32_t n_rows, const int32_t *__restrict__ row_b,
                                                     const int32_t *__restrict__ row_e, int64_t *__restrict__ len) {
    const int32_t r = int32_t(blockIdx.x * blockDim.x + threadIdx.x);
    if (r >= n_rows) return;
    const int32_t e = row_e ? row_e[r] : row_b[r + 1];
    len[r] = e > row_b[r] ? int64_t(e - row_b[r]) : 0;
}

// One wavefront per row: the row's entries become keys (column << 32 | row)
// and values at their place in the row-major list.
__global__ __launch_bounds__(256) void k_keys(int32_t n_rows, int64_t nnz, const int32_t *__restrict__ row_b,
                                              const int32_t *__restrict__ row_e, const int32_t *__restrict__ col,
                                              const float *__restrict__ val, const int64_t *__restrict__ off,
                                              uint64_t *__restrict__ keys, float *__restrict__ vals) {
    const int64_t r = int64_t(blockIdx.x) * (blockDim.x / 64) + threadIdx.x / 64;
    if (r >= n_rows) return;
    const int32_t b = row_b[r], e = row_e ? row_e[r] : row_b[r + 1];
    const int64_t o = off[r];
    // a caller's nnz below the rows' total keeps every write inside the lists
    for (int32_t k = b + int32_t(threadIdx.x & 63u); k < e && o + (k - b) < nnz; k += 64) {
        keys[o + (k - b)] = (uint64_t(uint32_t(col[k])) << 32) | uint64_t(r);
        vals[o + (k - b)] = val ? val[k] : 1.0f;
    }
}

// Column offsets of the sorted list: entry i opens the columns after the
// previous entry's up to its own; the last entry closes the rest.
__global__ __launch_bounds__(256) void k_col_ptr(int64_t nnz, int32_t n_cols, const uint64_t *__restrict__ keys,
                                                 int32_t *__restrict__ ptr, int32_t *__restrict__ rows) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (nnz == 0) {
        if (i <= n_cols) ptr[i] = 0;
        return;
    }
    if (i >= nnz) return;
    // columns outside [0, n_cols) (invalid input) are clamped so that no
    // write leaves ptr; their entries then belong to no column
    const int32_t c = int32_t(min<uint64_t>(keys[i] >> 32, uint64_t(n_cols)));
    rows[i] = int32_t(uint32_t(keys[i]));
    const int32_t prev = i > 0 ? int32_t(min<uint64_t>(keys[i - 1] >> 32, uint64_t(n_cols))) : -1;
    for (int32_t q = prev + 1; q <= c && q <= n_cols; ++q) ptr[q] = int32_t(i);
    if (i == nnz - 1)
        for (int32_t q = c + 1; q <= n_cols; ++q) ptr[q] = int32_t(nnz);
}

int64_t align256(int64_t b) { return (b + 255) / 256 * 256; }

struct Layout {
    int64_t len, off, keys, keys2, vals, vals2, ptr, rows, temp, total;
    size_t temp_bytes;
};

// Byte offsets of the workspace parts; temp_bytes from rocPRIM's queries.
int layout(int32_t n_rows, int32_t n_cols, int64_t nnz, Layout &L) {
    size_t scan_bytes = 0, sort_bytes = 0;
    hipError_t e = rocprim::exclusive_scan(nullptr, scan_bytes, static_cast<const int64_t *>(nullptr),
                                           static_cast<int64_t *>(nullptr), int64_t(0), size_t(n_rows),
                                           rocprim::plus<int64_t>());
    if (e == hipSuccess)
        e = rocprim::radix_sort_pairs(nullptr, sort_bytes, static_cast<const uint64_t *>(nullptr),
                                      static_cast<uint64_t *>(nullptr), static_cast<const float *>(nullptr),
                                      static_cast<float *>(nullptr), size_t(nnz));
    if (e != hipSuccess) return -1;
    L.temp_bytes = std::max(scan_bytes, sort_bytes);
    int64_t at = 0;
    L.len = at, at += align256(8 * int64_t(n_rows));
    L.off = at, at += align256(8 * (int64_t(n_rows) + 1));
    L.keys = at, at += align256(8 * nnz);
    L.keys2 = at, at += align256(8 * nnz);
    L.vals = at, at += align256(4 * nnz);
    L.vals2 = at, at += align256(4 * nnz);
    L.ptr = at, at += align256(4 * (int64_t(n_cols) + 1));
    L.rows = at, at += align256(4 * nnz);
    L.temp = at, at += align256(int64_t(L.temp_bytes));
    L.total = at;
    return 0;
}

}  // namespace

extern "C" int64_t cbm_op_spmm_t_workspace_bytes(int32_t n_rows, int32_t n_cols, int64_t nnz) {
    if (n_rows < 0 || n_cols < 0 || nnz < 0 || nnz > INT32_MAX) return -1;
    Layout L;
    if (layout(n_rows, n_cols, nnz, L) != 0) return -1;
    return L.total;
}

extern "C" int cbm_op_spmm_t(int32_t n_rows, int32_t n_cols, const int32_t *row_b, const int32_t *row_e,
                             const int32_t *col, const float *val, int64_t nnz, const float *rhs, int64_t ldr,
                             float *dst, int64_t ldd, int32_t n_feat, void *workspace, int64_t workspace_bytes,
                             hipStream_t stream) {
    if (n_rows < 0 || n_cols < 0 || nnz < 0 || nnz > INT32_MAX || n_feat < 0 || !rhs || !dst || ldr < n_feat ||
        ldd < n_feat || (n_rows > 0 && !row_b) || (nnz > 0 && !col)) {
        set_error("cbm_op_spmm_t: invalid arguments");
        return CBM_ERR_ARG;
    }
    if (n_cols == 0 || n_feat == 0) return CBM_OK;
    Layout L;
    if (layout(n_rows, n_cols, nnz, L) != 0) {
        set_error("cbm_op_spmm_t: rocPRIM workspace query failed");
        return CBM_ERR_DEVICE;
    }
    if (!workspace || workspace_bytes < L.total || (reinterpret_cast<uintptr_t>(workspace) & 255u) != 0) {
        set_error("cbm_op_spmm_t: workspace of %lld bytes (256-byte aligned) needed, got %lld", (long long)L.total,
                  (long long)workspace_bytes);
        return CBM_ERR_ARG;
    }
    char *ws = static_cast<char *>(workspace);
    auto *len = reinterpret_cast<int64_t *>(ws + L.len);
    auto *off = reinterpret_cast<int64_t *>(ws + L.off);
    auto *keys = reinterpret_cast<uint64_t *>(ws + L.keys);
    auto *keys2 = reinterpret_cast<uint64_t *>(ws + L.keys2);
    auto *vals = reinterpret_cast<float *>(ws + L.vals);
    auto *vals2 = reinterpret_cast<float *>(ws + L.vals2);
    auto *ptr = reinterpret_cast<int32_t *>(ws + L.ptr);
    auto *rows = reinterpret_cast<int32_t *>(ws + L.rows);
    void *temp = ws + L.temp;
    size_t temp_bytes = L.temp_bytes;
    hipError_t e = hipSuccess;
    if (n_rows > 0 && nnz > 0) {
        hipLaunchKernelGGL(k_row_lengths, dim3(unsigned((n_rows + 255) / 256)), dim3(256), 0, stream, n_rows, row_b,
                           row_e, len);
        e = hipGetLastError();
        if (e == hipSuccess)
            e = rocprim::exclusive_scan(temp, temp_bytes, len, off, int64_t(0), size_t(n_rows),
                                        rocprim::plus<int64_t>(), stream);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_keys, dim3(unsigned((int64_t(n_rows) + 3) / 4)), dim3(256), 0, stream, n_rows, nnz,
                               row_b, row_e, col, val, off, keys, vals);
            e = hipGetLastError();
        }
        unsigned col_bits = 1;
        while (col_bits < 32 && (int64_t(1) << col_bits) < int64_t(n_cols)) ++col_bits;
        temp_bytes = L.temp_bytes;
        if (e == hipSuccess)
            e = rocprim::radix_sort_pairs(temp, temp_bytes, keys, keys2, vals, vals2, size_t(nnz), 0u, 32u + col_bits,
                                          stream);
    }
    if (e == hipSuccess) {
        const int64_t threads = std::max<int64_t>(nnz, int64_t(n_cols) + 1);
        hipLaunchKernelGGL(k_col_ptr, dim3(unsigned((threads + 255) / 256)), dim3(256), 0, stream,
                           n_rows > 0 ? nnz : int64_t(0), n_cols, keys2, ptr, rows);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        cbmk::RowPass rp;
        rp.p_begin = 0;
        rp.p_end = n_cols;
        rp.ptr = ptr;
        rp.idx = rows;
        rp.val = vals2;
        rp.g = rhs;
        rp.ldg = ldr;
        rp.y = dst;
        rp.ldy = ldd;
        rp.ncols = n_feat;
        e = cbmk::launch_row_pass(rp, stream);
    }
    if (e != hipSuccess) {
        set_error("cbm_op_spmm_t: %s", hipGetErrorString(e));
        return CBM_ERR_DEVICE;
    }
    return CBM_OK;
}
