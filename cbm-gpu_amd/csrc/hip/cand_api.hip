// Candidate edges of the CBM build on the GPU (include/cbm_hip.h, "candidate
// edges").  Replaces the MKL syrk of A A^T and the two serial candidate scans
// of cbm_init_ (cbm/cbm_extensions.cpp:100-182); the arborescence and Delta
// stay on the host (cbm_build_from_candidates in cbm_host.h).
//
// Work: for every row r, the overlaps |A_r & A_c| with every c > r that
// shares a column with r, i.e. for each column k of A_r the rows above r
// that hold k (row k of A^T, which is A itself for a symmetric A) -- the
// upper triangle of A A^T, sum_k colcount(k)^2 / 2 counter increments.  Integer work bound by
// the rate of scattered counter updates and by list-search latency, not a
// dense contraction.  A persistent grid takes rows from a global work
// counter, one row per workgroup at a time; the wavefronts of a workgroup
// take the columns k of A_r in turn, find the first row of column k in range
// with a 64-way wavefront search, and count the rows after it.  Edges that
// pass the rule are placed by workgroup prefix sums, so each row's output is
// the reference's creation order (root edge, then for c ascending: c -> r,
// then r -> c).  The count pass writes the number of edges per row, the emit
// pass repeats the work and writes the edges at the caller's offsets.
//
// Two ways to hold the counters:
//   tiles  (n <= kTileMaxRows): the rows above r are cut into tiles of
//          kTileCols; a tile's counters live in LDS (128 KB, one 1024-thread
//          workgroup per CU), are counted with LDS atomics and then walked in
//          column order.  Cost per row ~ sum deg(k) + (n - r): for small n.
//   lists  (any n): each 1024-thread workgroup owns n global counters and a
//          touched list of n entries in the workspace.  Counting appends a column to the
//          list when its counter leaves 0; the list is filtered by the rule
//          into a short LDS list that alone is sorted (the passing edges are
//          few), then the touched counters are reset.  Cost per row ~ the
//          touched columns, independent of n.
#include "../../../include/cbm_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdarg>
#include <cstdio>
#include <string>

namespace {

constexpr int64_t kHeaderBytes = 256;                  // work counter, padded to keep the slabs aligned
constexpr int kTileThreads = 1024;
constexpr int32_t kTileCols = 32768;                   // uint32 LDS counters per tile (128 KB)
constexpr int32_t kTileMaxRows = 262144;               // automatic choice: tiles up to this size
constexpr int kListThreads = 1024;                     // 16 wavefronts: the walk is latency-bound
constexpr int32_t kPassCap = 4096;                     // passing columns sorted in LDS
constexpr int32_t kMaxListWorkgroups = 2048;
constexpr int64_t kDefaultListBytes = int64_t(8) << 30;  // counter + list slabs of the default workspace

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
    const int64_t *ptr = nullptr;      // A: row offsets (degrees) and columns
    const int32_t *col = nullptr;
    const int64_t *t_ptr = nullptr;    // A^T: for each column, the rows holding it
    const int32_t *t_col = nullptr;
    int64_t *row_count = nullptr;      // count pass: edges per row
    const int64_t *row_off = nullptr;  // emit pass: first edge of each row
    int32_t *src = nullptr, *dst = nullptr, *w = nullptr;
    int64_t cap = 0;                   // emit pass: entries of src / dst / w
    uint32_t *next_row = nullptr;
    uint32_t *slabs = nullptr;         // lists: 2 n words per workgroup (counters, touched list)
};

// First position p of the sorted range [b, e) of col with col[p] >= t (e if
// none), by the whole wavefront: 64 samples narrow the range 64-fold per
// round, so a list of up to 4096 entries takes two dependent loads.  Every
// lane must call it with the same arguments.
__device__ inline int64_t wave_lower_bound(const int32_t *col, int64_t b, int64_t e, int32_t t, int lane) {
    while (e - b > 64) {
        const int64_t step = (e - b + 63) / 64;
        const int64_t p = b + int64_t(lane) * step;
        const int nb = __popcll(__ballot(p < e && col[p] < t));  // samples below t form a prefix
        if (nb == 0) return b;                                    // col[b] >= t
        const int64_t hi = b + int64_t(nb) * step;                // first sample at or above t, if any
        b = b + int64_t(nb - 1) * step + 1;
        e = std::min(e, hi);
    }
    const int64_t p = b + lane;
    return b + __popcll(__ballot(p < e && col[p] < t));
}

// Exclusive prefix sum of v over a workgroup of WAVES wavefronts; *total
// receives the sum.  Every thread of the workgroup must call it.
template <int WAVES>
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
    for (int q = 0; q < WAVES; ++q) {
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

__device__ inline Rule decide(const CandArgs &a, int32_t dr, int32_t c, uint32_t overlap) {
    const int32_t dc = int32_t(a.ptr[c + 1] - a.ptr[c]);
    Rule q;
    q.h = dr + dc - 2 * int32_t(overlap);
    q.c_to_r = (q.h < dr - a.alpha) && (q.h != dr);
    q.r_to_c = (q.h < dc - a.alpha) && (q.h != dc);
    return q;
}

// Places the edges of one block of columns: thread t holds column c with
// rule q (q all false when it holds none); `run` is the row's edge count so
// far and is advanced by the block's edges.  Every thread must call it.
template <bool EMIT, int WAVES>
__device__ inline void place(const CandArgs &a, int32_t r, int32_t c, const Rule &q, int64_t &run,
                             int32_t *wave_tot) {
    int32_t total;
    const int32_t off = block_exclusive_scan<WAVES>(int32_t(q.c_to_r) + int32_t(q.r_to_c), wave_tot, &total);
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

// The root edge N -> r of weight deg(r) opens the row (:164).
template <bool EMIT>
__device__ inline void root_edge(const CandArgs &a, int32_t r, int32_t dr) {
    if (EMIT && threadIdx.x == 0) {
        const int64_t o = a.row_off[r];
        if (uint64_t(o) < uint64_t(a.cap)) {
            a.src[o] = a.n;
            a.dst[o] = r;
            a.w[o] = dr;
        }
    }
}

// Next row of the persistent grid (the same value for the whole workgroup).
__device__ inline int32_t next_row(const CandArgs &a, int32_t *s_row) {
    if (threadIdx.x == 0) *s_row = int32_t(min(atomicAdd(a.next_row, 1u), uint32_t(INT32_MAX)));
    __syncthreads();
    return *s_row;
}

template <bool EMIT>
__global__ __launch_bounds__(kTileThreads) void k_cand_tiles(const CandArgs a) {
    constexpr int kWaves = kTileThreads / 64;
    __shared__ uint32_t cnt[kTileCols];
    __shared__ int32_t wave_tot[kWaves];
    __shared__ int32_t s_row;
    const int tid = int(threadIdx.x), lane = tid & 63, wave = tid >> 6;
    for (int32_t i = tid; i < kTileCols; i += kTileThreads) cnt[i] = 0u;
    for (;;) {
        const int32_t r = next_row(a, &s_row);
        if (r >= a.n) return;  // the same row for every thread: the workgroup leaves together
        const int64_t rb = a.ptr[r], re = a.ptr[r + 1];
        const int32_t dr = int32_t(re - rb);
        root_edge<EMIT>(a, r, dr);
        int64_t run = 1;
        for (int32_t t0 = r + 1; t0 < a.n; t0 += std::min(kTileCols, a.n - t0)) {
            const int32_t t1 = t0 + std::min(kTileCols, a.n - t0);
            // 1. Overlap counts of the tile's columns.
            for (int64_t i = rb + wave; i < re; i += kWaves) {
                const int32_t k = a.col[i];
                if (uint32_t(k) >= uint32_t(a.n)) continue;
                const int64_t ke = a.t_ptr[k + 1];
                for (int64_t q = wave_lower_bound(a.t_col, a.t_ptr[k], ke, t0, lane) + lane; q < ke; q += 64) {
                    const int64_t off = int64_t(a.t_col[q]) - t0;  // in [0, t1 - t0) for sorted rows
                    if (uint64_t(off) >= uint64_t(t1 - t0)) break;
                    atomicAdd(&cnt[off], 1u);
                }
            }
            __syncthreads();
            // 2. The tile's columns in order; counters back to 0 as they are read.
            for (int32_t base = t0; base < t1; base += kTileThreads) {
                const int32_t c = base + tid;
                Rule q;
                if (c < t1) {
                    const uint32_t ov = cnt[c - t0];
                    if (ov != 0u) {
                        cnt[c - t0] = 0u;
                        q = decide(a, dr, c, ov);
                    }
                }
                place<EMIT, kWaves>(a, r, c, q, run, wave_tot);
            }
        }
        if (!EMIT && tid == 0) a.row_count[r] = run;
        __syncthreads();  // s_row is rewritten for the next row
    }
}

template <bool EMIT>
__global__ __launch_bounds__(kListThreads) void k_cand_lists(const CandArgs a) {
    constexpr int kWaves = kListThreads / 64;
    __shared__ int32_t pass[kPassCap];
    __shared__ int32_t wave_tot[kWaves];
    __shared__ int32_t s_row, s_touched, s_pass;
    uint32_t *cnt = a.slabs + size_t(blockIdx.x) * 2 * size_t(a.n);
    int32_t *touched = reinterpret_cast<int32_t *>(cnt + a.n);
    const int tid = int(threadIdx.x), lane = tid & 63, wave = tid >> 6;
    for (;;) {
        if (tid == 0) {
            s_touched = 0;
            s_pass = 0;
        }
        const int32_t r = next_row(a, &s_row);
        if (r >= a.n) return;  // the same row for every thread: the workgroup leaves together
        const int64_t rb = a.ptr[r], re = a.ptr[r + 1];
        const int32_t dr = int32_t(re - rb);

        // 1. Overlap counts of the columns above r (global atomics, in L2); a
        //    counter leaving 0 puts its column on the touched list.
        for (int64_t i = rb + wave; i < re; i += kWaves) {
            const int32_t k = a.col[i];
            if (uint32_t(k) >= uint32_t(a.n)) continue;
            const int64_t ke = a.t_ptr[k + 1];
            for (int64_t q = wave_lower_bound(a.t_col, a.t_ptr[k], ke, r + 1, lane) + lane; q < ke; q += 64) {
                const int32_t c = a.t_col[q];
                if (uint32_t(c) >= uint32_t(a.n)) continue;
                if (atomicAdd(&cnt[c], 1u) == 0u) touched[atomicAdd(&s_touched, 1)] = c;
            }
        }
        __syncthreads();
        const int32_t nt = s_touched;

        // 2. The touched columns whose edges pass the rule (few of them).
        for (int32_t i = tid; i < nt; i += kListThreads) {
            const int32_t c = touched[i];
            const Rule q = decide(a, dr, c, __hip_atomic_load(&cnt[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
            if (q.c_to_r || q.r_to_c) {
                const int32_t pos = atomicAdd(&s_pass, 1);
                if (pos < kPassCap) pass[pos] = c;
            }
        }
        __syncthreads();
        const int32_t np = s_pass;

        root_edge<EMIT>(a, r, dr);
        int64_t run = 1;
        if (np <= kPassCap) {
            // 3a. Sort the passing columns (bitonic, in LDS) and place them.
            int32_t m = 1;
            while (m < np) m <<= 1;
            for (int32_t i = np + tid; i < m; i += kListThreads) pass[i] = INT32_MAX;
            __syncthreads();
            for (int32_t k = 2; k <= m; k <<= 1) {
                for (int32_t j = k >> 1; j > 0; j >>= 1) {
                    for (int32_t i = tid; i < m; i += kListThreads) {
                        const int32_t ixj = i ^ j;
                        if (ixj > i) {
                            const int32_t x = pass[i], y = pass[ixj];
                            if ((x > y) == ((i & k) == 0)) {
                                pass[i] = y;
                                pass[ixj] = x;
                            }
                        }
                    }
                    __syncthreads();
                }
            }
            for (int32_t base = 0; base < np; base += kListThreads) {
                const int32_t i = base + tid;
                int32_t c = -1;
                Rule q;
                if (i < np) {
                    c = pass[i];
                    q = decide(a, dr, c, __hip_atomic_load(&cnt[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
                }
                place<EMIT, kWaves>(a, r, c, q, run, wave_tot);
            }
        } else {
            // 3b. More passing columns than LDS holds: walk every counter above r.
            for (int64_t base = int64_t(r) + 1; base < int64_t(a.n); base += kListThreads) {
                const int64_t cc = base + tid;
                int32_t c = -1;
                Rule q;
                if (cc < int64_t(a.n)) {
                    c = int32_t(cc);
                    const uint32_t ov = __hip_atomic_load(&cnt[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (ov != 0u) q = decide(a, dr, c, ov);
                }
                place<EMIT, kWaves>(a, r, c, q, run, wave_tot);
            }
        }
        // 4. Touched counters back to 0 for the next row.
        for (int32_t i = tid; i < nt; i += kListThreads)
            __hip_atomic_store(&cnt[touched[i]], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (!EMIT && tid == 0) a.row_count[r] = run;
        __syncthreads();  // s_row, s_touched, s_pass and pass are rewritten for the next row
    }
}

// Partner scan of the gather order (cbm_partners): for every row r of an
// ordered row list, the k rows sharing the most columns with it.  The same
// counting as k_cand_lists, over every other row instead of the rows above
// r, and only through columns held by at most `cap` rows (a column held by
// many rows adds one to each of them and tells little about which share
// most, at a quadratic cost).  The counts are walked once for the best
// kMaxPartners (most overlap, then lowest row), combined over the workgroup
// by a tree in LDS.
constexpr int kMaxPartners = 4;

struct PartnerArgs {
    int32_t n = 0;
    int32_t n_cols = 0;
    int32_t cap = 0;
    int32_t k = 0;                   // partners written per row (1 .. kMaxPartners)
    const int32_t *ptr = nullptr;    // rows: entry offsets and columns (bit 31 a sign)
    const int32_t *col = nullptr;
    const int32_t *t_ptr = nullptr;  // per column: the rows holding it
    const int32_t *t_col = nullptr;
    int32_t *partner = nullptr;      // [k n]
    int32_t *overlap = nullptr;      // [k n]
    uint32_t *next_row = nullptr;
    uint32_t *slabs = nullptr;       // 2 n words per workgroup (counters, touched list)
};

// Ranking key of a (count, row) pair: more overlap first, then the lower row.
__device__ inline uint64_t partner_key(uint32_t ov, int32_t row) {
    return ov ? (uint64_t(ov) << 32) | uint64_t(uint32_t(INT32_MAX - row)) : 0ull;
}

// Inserts x into the descending list t (keys of distinct rows; 0 is empty).
__device__ inline void top_insert(uint64_t (&t)[kMaxPartners], uint64_t x) {
#pragma unroll
    for (int i = 0; i < kMaxPartners; ++i) {
        const uint64_t hi = t[i] > x ? t[i] : x, lo = t[i] > x ? x : t[i];
        t[i] = hi;
        x = lo;
    }
}

__global__ __launch_bounds__(kListThreads) void k_partner_lists(const PartnerArgs a) {
    constexpr int kWaves = kListThreads / 64;
    __shared__ uint64_t best[kMaxPartners * kListThreads];
    __shared__ int32_t s_row, s_touched;
    uint32_t *cnt = a.slabs + size_t(blockIdx.x) * 2 * size_t(a.n);
    int32_t *touched = reinterpret_cast<int32_t *>(cnt + a.n);
    const int tid = int(threadIdx.x), lane = tid & 63, wave = tid >> 6;
    for (;;) {
        if (tid == 0) s_touched = 0;
        if (tid == 0) s_row = int32_t(min(atomicAdd(a.next_row, 1u), uint32_t(INT32_MAX)));
        __syncthreads();
        const int32_t r = s_row;
        if (r >= a.n) return;  // the same row for every thread: the workgroup leaves together
        const int32_t rb = a.ptr[r], re = a.ptr[r + 1];
        for (int32_t i = rb + wave; i < re; i += kWaves) {
            const int32_t k = a.col[i] & 0x7fffffff;
            if (k >= a.n_cols) continue;
            const int32_t kb = a.t_ptr[k], ke = a.t_ptr[k + 1];
            if (ke - kb > a.cap) continue;
            for (int32_t q = kb + lane; q < ke; q += 64) {
                const int32_t c = a.t_col[q];
                if (c == r || uint32_t(c) >= uint32_t(a.n)) continue;
                if (atomicAdd(&cnt[c], 1u) == 0u) touched[atomicAdd(&s_touched, 1)] = c;
            }
        }
        __syncthreads();
        const int32_t nt = s_touched;
        uint64_t t[kMaxPartners] = {};
        for (int32_t i = tid; i < nt; i += kListThreads) {
            const int32_t c = touched[i];
            const uint32_t ov = __hip_atomic_load(&cnt[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&cnt[c], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            top_insert(t, partner_key(ov, c));
        }
        // Tree over the workgroup through LDS: at each step the lower half
        // takes in the upper half's lists.
#pragma unroll
        for (int j = 0; j < kMaxPartners; ++j) best[kMaxPartners * tid + j] = t[j];
        for (int stride = kListThreads / 2; stride > 0; stride >>= 1) {
            __syncthreads();
            if (tid < stride) {
#pragma unroll
                for (int j = 0; j < kMaxPartners; ++j) top_insert(t, best[kMaxPartners * (tid + stride) + j]);
#pragma unroll
                for (int j = 0; j < kMaxPartners; ++j) best[kMaxPartners * tid + j] = t[j];
            }
        }
        if (tid == 0) {
#pragma unroll
            for (int j = 0; j < kMaxPartners; ++j) {
                if (j < a.k) {
                    a.partner[int64_t(a.k) * r + j] = t[j] ? INT32_MAX - int32_t(uint32_t(t[j])) : -1;
                    a.overlap[int64_t(a.k) * r + j] = int32_t(t[j] >> 32);
                }
            }
        }
        __syncthreads();  // s_row, s_touched and best are rewritten for the next row
    }
}

__global__ void k_zero_words(uint32_t *p, int64_t n) {
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0u;
}

// The path a call takes, and how many workgroups run it.
struct Shape {
    int32_t path = 0;
    int32_t wgs = 0;
    int64_t bytes = 0;  // workspace it needs
};

bool resolve(int32_t n_rows, int32_t path, int32_t n_workgroups, Shape &s) {
    if (n_rows < 1 || path < CBM_CAND_AUTO || path > CBM_CAND_LISTS || n_workgroups < 0) return false;
    s.path = path == CBM_CAND_AUTO ? (n_rows <= kTileMaxRows ? CBM_CAND_TILES : CBM_CAND_LISTS) : path;
    int64_t wgs = n_workgroups;
    if (s.path == CBM_CAND_TILES) {
        if (wgs == 0) {
            int dev = 0, cus = 0;
            if (hipGetDevice(&dev) != hipSuccess ||
                hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
                cus = 256;
            wgs = cus;
        }
        s.bytes = kHeaderBytes;
    } else {
        const int64_t slab = int64_t(n_rows) * 8;
        if (wgs == 0) wgs = std::max<int64_t>(1, kDefaultListBytes / slab);
        wgs = std::min<int64_t>(wgs, kMaxListWorkgroups);
        s.bytes = kHeaderBytes + std::min<int64_t>(wgs, n_rows) * slab;
    }
    s.wgs = int32_t(std::min<int64_t>(wgs, n_rows));
    return true;
}

int run_pass(bool emit, int32_t n_rows, const int64_t *ptr, const int32_t *col, const int64_t *t_ptr,
             const int32_t *t_col, int32_t alpha, int32_t path,
             int32_t n_workgroups, int64_t *row_count, const int64_t *row_off, int32_t *src, int32_t *dst, int32_t *w,
             int64_t cap, void *workspace, int64_t workspace_bytes, hipStream_t stream, const char *what) {
    Shape s;
    if (!resolve(n_rows, path, n_workgroups, s) || !ptr || !col || !t_ptr || !t_col || !workspace ||
        (!emit && !row_count) ||
        (emit && (!row_off || cap < 0 || (cap > 0 && (!src || !dst || !w))))) {
        set_error("%s: invalid arguments (n_rows=%d path=%d n_workgroups=%d)", what, n_rows, path, n_workgroups);
        return CBM_ERR_ARG;
    }
    if ((reinterpret_cast<uintptr_t>(workspace) & 255u) != 0) {
        set_error("%s: the workspace must be 256-byte aligned", what);
        return CBM_ERR_ARG;
    }
    if (workspace_bytes < s.bytes) {
        set_error("%s: workspace of %lld bytes, %lld needed", what, (long long)workspace_bytes, (long long)s.bytes);
        return CBM_ERR_ARG;
    }
    CandArgs a;
    a.n = n_rows;
    a.alpha = alpha;
    a.ptr = ptr;
    a.col = col;
    a.t_ptr = t_ptr;
    a.t_col = t_col;
    a.row_count = row_count;
    a.row_off = row_off;
    a.src = src;
    a.dst = dst;
    a.w = w;
    a.cap = cap;
    a.next_row = static_cast<uint32_t *>(workspace);
    a.slabs = reinterpret_cast<uint32_t *>(static_cast<char *>(workspace) + kHeaderBytes);
    // Work counter and counter slabs start at zero (a fill kernel, so the
    // pass can sit in a captured graph).
    const int64_t words = s.bytes / 4;
    k_zero_words<<<dim3(unsigned(std::min<int64_t>((words + 255) / 256, 8192))), dim3(256), 0, stream>>>(
        static_cast<uint32_t *>(workspace), words);
    const dim3 grid(unsigned(s.wgs));
    if (s.path == CBM_CAND_TILES) {
        if (emit) {
            k_cand_tiles<true><<<grid, dim3(kTileThreads), 0, stream>>>(a);
        } else {
            k_cand_tiles<false><<<grid, dim3(kTileThreads), 0, stream>>>(a);
        }
    } else {
        if (emit) {
            k_cand_lists<true><<<grid, dim3(kListThreads), 0, stream>>>(a);
        } else {
            k_cand_lists<false><<<grid, dim3(kListThreads), 0, stream>>>(a);
        }
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

extern "C" int64_t cbm_partners_workspace_bytes(int32_t n_rows, int32_t n_workgroups) {
    Shape s;
    return resolve(n_rows, CBM_CAND_LISTS, n_workgroups, s) ? s.bytes : -1;
}

extern "C" int cbm_partners(int32_t n_rows, const int32_t *ptr, const int32_t *col, int32_t n_cols,
                            const int32_t *t_ptr, const int32_t *t_col, int32_t cap, int32_t k,
                            int32_t n_workgroups, int32_t *partner, int32_t *overlap, void *workspace,
                            int64_t workspace_bytes, hipStream_t stream) {
    Shape s;
    if (!resolve(n_rows, CBM_CAND_LISTS, n_workgroups, s) || n_cols < 0 || cap < 0 || k < 1 ||
        k > kMaxPartners || !ptr || !t_ptr || !partner || !overlap || !workspace) {
        set_error("cbm_partners: invalid arguments (n_rows=%d n_cols=%d cap=%d k=%d)", n_rows, n_cols, cap, k);
        return CBM_ERR_ARG;
    }
    if ((reinterpret_cast<uintptr_t>(workspace) & 255u) != 0) {
        set_error("cbm_partners: the workspace must be 256-byte aligned");
        return CBM_ERR_ARG;
    }
    if (workspace_bytes < s.bytes) {
        set_error("cbm_partners: workspace of %lld bytes, %lld needed", (long long)workspace_bytes,
                  (long long)s.bytes);
        return CBM_ERR_ARG;
    }
    PartnerArgs a;
    a.n = n_rows;
    a.n_cols = n_cols;
    a.cap = cap;
    a.k = k;
    a.ptr = ptr;
    a.col = col;
    a.t_ptr = t_ptr;
    a.t_col = t_col;
    a.partner = partner;
    a.overlap = overlap;
    a.next_row = static_cast<uint32_t *>(workspace);
    a.slabs = reinterpret_cast<uint32_t *>(static_cast<char *>(workspace) + kHeaderBytes);
    const int64_t words = s.bytes / 4;
    k_zero_words<<<dim3(unsigned(std::min<int64_t>((words + 255) / 256, 8192))), dim3(256), 0, stream>>>(
        static_cast<uint32_t *>(workspace), words);
    k_partner_lists<<<dim3(unsigned(s.wgs)), dim3(kListThreads), 0, stream>>>(a);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("cbm_partners: %s", hipGetErrorString(e));
        return CBM_ERR_DEVICE;
    }
    return CBM_OK;
}

extern "C" int64_t cbm_candidates_workspace_bytes(int32_t n_rows, int32_t path, int32_t n_workgroups) {
    Shape s;
    return resolve(n_rows, path, n_workgroups, s) ? s.bytes : -1;
}

extern "C" int cbm_candidates_count(int32_t n_rows, const int64_t *ptr, const int32_t *col, const int64_t *t_ptr,
                                    const int32_t *t_col, int32_t alpha, int32_t path, int32_t n_workgroups,
                                    int64_t *row_count, void *workspace, int64_t workspace_bytes,
                                    hipStream_t stream) {
    return run_pass(false, n_rows, ptr, col, t_ptr, t_col, alpha, path, n_workgroups, row_count, nullptr, nullptr,
                    nullptr, nullptr, 0, workspace, workspace_bytes, stream, "cbm_candidates_count");
}

extern "C" int cbm_candidates_emit(int32_t n_rows, const int64_t *ptr, const int32_t *col, const int64_t *t_ptr,
                                   const int32_t *t_col, int32_t alpha, int32_t path, int32_t n_workgroups,
                                   const int64_t *row_offset, int32_t *src, int32_t *dst, int32_t *w, int64_t n_cand,
                                   void *workspace, int64_t workspace_bytes, hipStream_t stream) {
    return run_pass(true, n_rows, ptr, col, t_ptr, t_col, alpha, path, n_workgroups, nullptr, row_offset, src, dst, w,
                    n_cand, workspace, workspace_bytes, stream, "cbm_candidates_emit");
}
