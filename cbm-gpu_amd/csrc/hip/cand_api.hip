// Candidate edges of the CBM build on the GPU (include/cbm_hip.h, "candidate
// edges").  Replaces the MKL syrk of A A^T and the two serial candidate scans
// of cbm_init_ (cbm/cbm_extensions.cpp:100-182); the arborescence and Delta
// stay on the host (cbm_build_from_candidates in cbm_host.h).
//
// Work: for every row r, the overlaps |A_r & A_c| with every c > r that
// shares a column with r, i.e. sum over k in A_r of the entries of A_k above
// r -- sum_k deg(k)^2 / 2 increments in all, integer work bound by the
// latency of scattered counter updates, not by a dense contraction.
//
// Mapping: a persistent grid of P workgroups; a workgroup takes the next row
// from a global work counter and owns n private uint32 counters in the
// workspace (its own slab, so no two workgroups ever touch one counter).
//   1. its four wavefronts walk the rows k of A_r (one k per wavefront at a
//      time), binary-search the first entry of A_k above r and add 1 to the
//      counter of every such column c (global atomics, performed in L2); a
//      counter going 0 -> 1 appends c to the touched list in LDS;
//   2. if at most kSortCap columns were touched, the list is sorted in LDS
//      (bitonic) and walked in ascending c; otherwise every counter of
//      (r, n) is walked in ascending c.  Each column's counter is read and
//      reset to 0 in the same pass, so the slab is clean for the next row;
//   3. per block of 256 columns, the edges that pass the rule are placed by
//      a workgroup prefix sum: the output is the reference's creation order
//      (root edge, then for c ascending: c -> r, then r -> c).
// The count pass runs 1-3 writing only the number of edges per row; the
// emit pass runs them again and writes the edges at the caller's offsets.
#include "../../../include/cbm_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdarg>
#include <cstdio>
#include <string>

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int32_t kSortCap = 4096;          // touched columns handled by the sorted path
constexpr int64_t kHeaderBytes = 256;       // work counter, padded to keep the slabs aligned
constexpr int32_t kMaxWorkgroups = 2048;    // 8 per CU
constexpr int64_t kDefaultSlabBytes = int64_t(4) << 30;  // counter slabs of the default workspace

thread_local std::string g_error;

void set_error(const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_error = buf;
}

struct CandArgs {
    int32_t n = 0;
    int32_t alpha = 0;
    const int64_t *ptr = nullptr;
    const int32_t *col = nullptr;
    int64_t *row_count = nullptr;      // count pass: edges per row
    const int64_t *row_off = nullptr;  // emit pass: first edge of each row
    int32_t *src = nullptr, *dst = nullptr, *w = nullptr;
    int64_t cap = 0;                   // emit pass: entries of src / dst / w
    uint32_t *next_row = nullptr;
    uint32_t *cnt = nullptr;           // n counters per workgroup
};

// First position of the sorted range [b, e) of col holding a value above r.
__device__ inline int64_t first_above(const int32_t *col, int64_t b, int64_t e, int32_t r) {
    while (b < e) {
        const int64_t m = b + (e - b) / 2;
        if (col[m] <= r) {
            b = m + 1;
        } else {
            e = m;
        }
    }
    return b;
}

// Exclusive prefix sum of v over the workgroup; *total receives the sum.
// Every thread of the workgroup must call it.
__device__ inline int32_t block_exclusive_scan(int32_t v, int32_t *wave_tot, int32_t *total) {
    const int lane = int(threadIdx.x & 63u), wave = int(threadIdx.x >> 6);
    int32_t incl = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int32_t t = __shfl_up(incl, off, 64);
        if (lane >= off) incl += t;
    }
    if (lane == 63) wave_tot[wave] = incl;
    __syncthreads();
    int32_t base = 0, sum = 0;
#pragma unroll
    for (int q = 0; q < kWaves; ++q) {
        const int32_t t = wave_tot[q];
        base += q < wave ? t : 0;
        sum += t;
    }
    *total = sum;
    __syncthreads();  // wave_tot is rewritten by the next call
    return base + incl - v;
}

// The candidate rule (cbm/cbm_extensions.cpp:164-180), in the same int32
// arithmetic as the host scan: h = |A_r xor A_c|.
struct Rule {
    bool c_to_r = false, r_to_c = false;
    int32_t h = 0;
};

__device__ inline Rule decide(int32_t dr, int32_t dc, uint32_t overlap, int32_t alpha) {
    Rule q;
    q.h = dr + dc - 2 * int32_t(overlap);
    q.c_to_r = (q.h < dr - alpha) && (q.h != dr);
    q.r_to_c = (q.h < dc - alpha) && (q.h != dc);
    return q;
}

// Places the edges of one block of columns: thread t holds column c with
// rule q (q all false when it holds none); `run` is the row's edge count so
// far and is advanced by the block's edges.
template <bool EMIT>
__device__ inline void place(const CandArgs &a, int32_t r, int32_t c, const Rule &q, int64_t &run,
                             int32_t *wave_tot) {
    int32_t total;
    const int32_t off = block_exclusive_scan(int32_t(q.c_to_r) + int32_t(q.r_to_c), wave_tot, &total);
    if constexpr (EMIT) {
        // Writes stay inside the caller's arrays whatever the offsets say.
        int64_t o = a.row_off[r] + run + off;
        if (q.c_to_r) {
            if (uint64_t(o) < uint64_t(a.cap)) {
                a.src[o] = c;
                a.dst[o] = r;
                a.w[o] = q.h;
            }
            ++o;
        }
        if (q.r_to_c && uint64_t(o) < uint64_t(a.cap)) {
            a.src[o] = r;
            a.dst[o] = c;
            a.w[o] = q.h;
        }
    }
    run += total;
}

template <bool EMIT>
__global__ __launch_bounds__(kThreads) void k_candidates(const CandArgs a) {
    __shared__ int32_t touched[kSortCap];
    __shared__ int32_t wave_tot[kWaves];
    __shared__ int32_t s_row, s_touched;
    uint32_t *cnt = a.cnt + size_t(blockIdx.x) * size_t(a.n);
    const int tid = int(threadIdx.x), lane = tid & 63, wave = tid >> 6;
    for (;;) {
        if (tid == 0) {
            s_row = int32_t(min(atomicAdd(a.next_row, 1u), uint32_t(INT32_MAX)));
            s_touched = 0;
        }
        __syncthreads();
        const int32_t r = s_row;
        if (r >= a.n) return;  // every thread read the same row: the workgroup leaves together
        const int64_t rb = a.ptr[r], re = a.ptr[r + 1];
        const int32_t dr = int32_t(re - rb);

        // 1. Overlap counts of the columns above r.
        for (int64_t i = rb + wave; i < re; i += kWaves) {
            const int32_t k = a.col[i];
            if (uint32_t(k) >= uint32_t(a.n)) continue;
            const int64_t ke = a.ptr[k + 1];
            for (int64_t q = first_above(a.col, a.ptr[k], ke, r) + lane; q < ke; q += 64) {
                const int32_t c = a.col[q];
                if (uint32_t(c) >= uint32_t(a.n)) continue;
                if (atomicAdd(&cnt[c], 1u) == 0u) {
                    const int32_t pos = atomicAdd(&s_touched, 1);
                    if (pos < kSortCap) touched[pos] = c;
                }
            }
        }
        __syncthreads();
        const int32_t nt = s_touched;

        // The root edge N -> r of weight deg(r) opens the row (:164).
        if (EMIT && tid == 0) {
            const int64_t o = a.row_off[r];
            if (uint64_t(o) < uint64_t(a.cap)) {
                a.src[o] = a.n;
                a.dst[o] = r;
                a.w[o] = dr;
            }
        }
        int64_t run = 1;

        if (nt <= kSortCap) {
            // 2a. Few touched columns: sort them, then walk them in order.
            int32_t m = 1;
            while (m < nt) m <<= 1;
            for (int32_t i = nt + tid; i < m; i += kThreads) touched[i] = INT32_MAX;
            __syncthreads();
            for (int32_t k = 2; k <= m; k <<= 1) {
                for (int32_t j = k >> 1; j > 0; j >>= 1) {
                    for (int32_t i = tid; i < m; i += kThreads) {
                        const int32_t ixj = i ^ j;
                        if (ixj > i) {
                            const int32_t x = touched[i], y = touched[ixj];
                            if ((x > y) == ((i & k) == 0)) {
                                touched[i] = y;
                                touched[ixj] = x;
                            }
                        }
                    }
                    __syncthreads();
                }
            }
            for (int32_t base = 0; base < nt; base += kThreads) {
                const int32_t i = base + tid;
                int32_t c = -1;
                Rule q;
                if (i < nt) {
                    c = touched[i];
                    const uint32_t ov = atomicExch(&cnt[c], 0u);
                    q = decide(dr, int32_t(a.ptr[c + 1] - a.ptr[c]), ov, a.alpha);
                }
                place<EMIT>(a, r, c, q, run, wave_tot);
            }
        } else {
            // 2b. Many touched columns: walk every counter above r in order.
            for (int64_t base = int64_t(r) + 1; base < int64_t(a.n); base += kThreads) {
                const int64_t cc = base + tid;
                int32_t c = -1;
                Rule q;
                if (cc < int64_t(a.n)) {
                    c = int32_t(cc);
                    const uint32_t ov = __hip_atomic_load(&cnt[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (ov != 0u) {
                        __hip_atomic_store(&cnt[c], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        q = decide(dr, int32_t(a.ptr[c + 1] - a.ptr[c]), ov, a.alpha);
                    }
                }
                place<EMIT>(a, r, c, q, run, wave_tot);
            }
        }
        if (!EMIT && tid == 0) a.row_count[r] = run;
        __syncthreads();  // s_row, s_touched and touched are rewritten for the next row
    }
}

__global__ void k_zero_words(uint32_t *p, int64_t n) {
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0u;
}

int32_t workgroups_for(int32_t n_rows, int64_t workspace_bytes) {
    if (n_rows < 1 || workspace_bytes <= kHeaderBytes) return 0;
    const int64_t slab = int64_t(n_rows) * 4;
    return int32_t(std::min<int64_t>({(workspace_bytes - kHeaderBytes) / slab, kMaxWorkgroups, n_rows}));
}

int run_pass(bool emit, int32_t n_rows, const int64_t *ptr, const int32_t *col, int32_t alpha, int64_t *row_count,
             const int64_t *row_off, int32_t *src, int32_t *dst, int32_t *w, int64_t cap, void *workspace,
             int64_t workspace_bytes, hipStream_t stream, const char *what) {
    if (n_rows < 1 || !ptr || !col || !workspace || (!emit && !row_count) ||
        (emit && (!row_off || cap < 0 || (cap > 0 && (!src || !dst || !w))))) {
        set_error("%s: invalid arguments (n_rows=%d)", what, n_rows);
        return CBM_ERR_ARG;
    }
    if ((reinterpret_cast<uintptr_t>(workspace) & 255u) != 0) {
        set_error("%s: the workspace must be 256-byte aligned", what);
        return CBM_ERR_ARG;
    }
    const int32_t wgs = workgroups_for(n_rows, workspace_bytes);
    if (wgs < 1) {
        set_error("%s: workspace of %lld bytes holds no counter slab of %lld bytes", what, (long long)workspace_bytes,
                  (long long)(int64_t(n_rows) * 4 + kHeaderBytes));
        return CBM_ERR_ARG;
    }
    CandArgs a;
    a.n = n_rows;
    a.alpha = alpha;
    a.ptr = ptr;
    a.col = col;
    a.row_count = row_count;
    a.row_off = row_off;
    a.src = src;
    a.dst = dst;
    a.w = w;
    a.cap = cap;
    a.next_row = static_cast<uint32_t *>(workspace);
    a.cnt = reinterpret_cast<uint32_t *>(static_cast<char *>(workspace) + kHeaderBytes);
    // Work counter and every slab start at zero (a fill kernel, so the pass
    // can sit in a captured graph).
    const int64_t words = (kHeaderBytes + int64_t(wgs) * n_rows * 4) / 4;
    k_zero_words<<<dim3(unsigned(std::min<int64_t>((words + 255) / 256, 8192))), dim3(256), 0, stream>>>(
        static_cast<uint32_t *>(workspace), words);
    if (emit) {
        k_candidates<true><<<dim3(unsigned(wgs)), dim3(kThreads), 0, stream>>>(a);
    } else {
        k_candidates<false><<<dim3(unsigned(wgs)), dim3(kThreads), 0, stream>>>(a);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("%s: %s", what, hipGetErrorString(e));
        return CBM_ERR_DEVICE;
    }
    return CBM_OK;
}

}  // namespace

extern "C" const char *cbm_candidates_last_error(void) { return g_error.c_str(); }

extern "C" int64_t cbm_candidates_workspace_bytes(int32_t n_rows, int32_t n_workgroups) {
    if (n_rows < 1) return -1;
    const int64_t slab = int64_t(n_rows) * 4;
    int64_t wgs = n_workgroups > 0 ? n_workgroups : std::max<int64_t>(1, kDefaultSlabBytes / slab);
    wgs = std::min<int64_t>({wgs, kMaxWorkgroups, n_rows});
    return kHeaderBytes + wgs * slab;
}

extern "C" int cbm_candidates_count(int32_t n_rows, const int64_t *ptr, const int32_t *col, int32_t alpha,
                                    int64_t *row_count, void *workspace, int64_t workspace_bytes,
                                    hipStream_t stream) {
    return run_pass(false, n_rows, ptr, col, alpha, row_count, nullptr, nullptr, nullptr, nullptr, 0, workspace,
                    workspace_bytes, stream, "cbm_candidates_count");
}

extern "C" int cbm_candidates_emit(int32_t n_rows, const int64_t *ptr, const int32_t *col, int32_t alpha,
                                   const int64_t *row_offset, int32_t *src, int32_t *dst, int32_t *w, int64_t n_cand,
                                   void *workspace, int64_t workspace_bytes, hipStream_t stream) {
    return run_pass(true, n_rows, ptr, col, alpha, nullptr, row_offset, src, dst, w, n_cand, workspace,
                    workspace_bytes, stream, "cbm_candidates_emit");
}
