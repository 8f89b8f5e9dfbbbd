// Row clustering passes (include/cbm_hip.h, "row clustering"): the two data
// passes of the fingerprint clustering of benchmark/clustering_gpu.py:86-128
// over a binary CSR, for a selected set of rows.
//   column counts   counts[c] = selected rows that hold column c  (:99-100,
//                   the in-degrees within the cluster)
//   fingerprints    fp[i] = sum of 2^j over the hash columns hash[j] that
//                   row rows[i] holds (:106-110; a row holds a column at most
//                   once, so the sum is a bitwise or)
// One wavefront per selected row, lanes over the row's entries; the column
// counts are global atomics, the fingerprint an or-reduction across lanes.
#include "../../../include/cbm_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <string>

namespace {

constexpr int kThreads = 256;
constexpr int kMaxHashes = 30;

thread_local std::string g_error;

void set_error(const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_error = buf;
}

__global__ void k_zero(uint32_t *p, int64_t n) {
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0u;
}

__global__ __launch_bounds__(kThreads) void k_column_counts(const int64_t *ptr, const int32_t *col,
                                                            const int32_t *rows, int32_t n_sel, int32_t n_rows,
                                                            int32_t n_cols, uint32_t *counts) {
    const int lane = int(threadIdx.x & 63u);
    const int64_t waves = int64_t(gridDim.x) * (kThreads / 64);
    for (int64_t i = int64_t(blockIdx.x) * (kThreads / 64) + (threadIdx.x >> 6); i < n_sel; i += waves) {
        const int32_t r = rows[i];
        if (uint32_t(r) >= uint32_t(n_rows)) continue;
        const int64_t e = ptr[r + 1];
        for (int64_t q = ptr[r] + lane; q < e; q += 64) {
            const int32_t c = col[q];
            if (uint32_t(c) < uint32_t(n_cols)) atomicAdd(&counts[c], 1u);
        }
    }
}

__global__ __launch_bounds__(kThreads) void k_fingerprints(const int64_t *ptr, const int32_t *col,
                                                           const int32_t *rows, int32_t n_sel, int32_t n_rows,
                                                           const int32_t *hash_cols, int32_t n_hash, int32_t *fp) {
    __shared__ int32_t hc[kMaxHashes];
    if (threadIdx.x < uint32_t(n_hash)) hc[threadIdx.x] = hash_cols[threadIdx.x];
    __syncthreads();
    const int lane = int(threadIdx.x & 63u);
    const int64_t waves = int64_t(gridDim.x) * (kThreads / 64);
    for (int64_t i = int64_t(blockIdx.x) * (kThreads / 64) + (threadIdx.x >> 6); i < n_sel; i += waves) {
        const int32_t r = rows[i];
        int32_t bits = 0;
        if (uint32_t(r) < uint32_t(n_rows)) {
            const int64_t e = ptr[r + 1];
            for (int64_t q = ptr[r] + lane; q < e; q += 64) {
                const int32_t c = col[q];
                for (int32_t j = 0; j < n_hash; ++j) bits |= (c == hc[j]) ? int32_t(1) << j : 0;
            }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) bits |= __shfl_xor(bits, off, 64);
        if (lane == 0) fp[i] = bits;
    }
}

unsigned grid_for(int64_t n_sel) {
    return unsigned(std::max<int64_t>(1, std::min<int64_t>((n_sel + 3) / 4, 8192)));
}

int launched(const char *what) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("%s: %s", what, hipGetErrorString(e));
        return CBM_ERR_DEVICE;
    }
    return CBM_OK;
}

}  // namespace

extern "C" const char *cbm_cluster_last_error(void) { return g_error.c_str(); }

extern "C" int cbm_cluster_column_counts(int32_t n_rows, int32_t n_cols, const int64_t *ptr, const int32_t *col,
                                         const int32_t *rows, int32_t n_sel, int32_t *counts, hipStream_t stream) {
    if (n_rows < 0 || n_cols < 0 || n_sel < 0 || (n_sel > 0 && (!ptr || !col || !rows)) ||
        (n_cols > 0 && !counts)) {
        set_error("cbm_cluster_column_counts: invalid arguments");
        return CBM_ERR_ARG;
    }
    if (n_cols == 0) return CBM_OK;
    auto *c = reinterpret_cast<uint32_t *>(counts);
    k_zero<<<dim3(unsigned(std::min<int64_t>((int64_t(n_cols) + 255) / 256, 8192))), dim3(256), 0, stream>>>(
        c, n_cols);
    if (n_sel > 0)
        k_column_counts<<<dim3(grid_for(n_sel)), dim3(kThreads), 0, stream>>>(ptr, col, rows, n_sel, n_rows, n_cols,
                                                                                c);
    return launched("cbm_cluster_column_counts");
}

extern "C" int cbm_cluster_fingerprints(int32_t n_rows, const int64_t *ptr, const int32_t *col, const int32_t *rows,
                                        int32_t n_sel, const int32_t *hash_cols, int32_t n_hash, int32_t *fp,
                                        hipStream_t stream) {
    if (n_rows < 0 || n_sel < 0 || n_hash < 1 || n_hash > kMaxHashes || !hash_cols ||
        (n_sel > 0 && (!ptr || !col || !rows || !fp))) {
        set_error("cbm_cluster_fingerprints: invalid arguments (n_hash=%d, at most %d)", n_hash, kMaxHashes);
        return CBM_ERR_ARG;
    }
    if (n_sel == 0) return CBM_OK;
    k_fingerprints<<<dim3(grid_for(n_sel)), dim3(kThreads), 0, stream>>>(ptr, col, rows, n_sel, n_rows, hash_cols,
                                                                          n_hash, fp);
    return launched("cbm_cluster_fingerprints");
}
