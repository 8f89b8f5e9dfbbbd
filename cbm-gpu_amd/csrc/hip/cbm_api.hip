// C ABI of the CBM product (include/cbm_hip.h).  Replaces class cusparseCBM
// of cbm/cbm.cu:93-300: no cuSPARSE handle, no per-call descriptors and no
// per-call workspace (cbm/cbm.cu:157-181,203-205), and every launch goes to
// the caller's stream instead of the legacy default stream (:196,256).
#include "cbm_kernels.hpp"
#include "../../../include/cbm_hip.h"
#include "../../../include/cbm_host.h"

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <mutex>
#include <new>
#include <string>
#include <vector>

// Scratch of cbm_matmul_noalloc, for callers that cannot allocate: the
// long-row partial sums and the compensation rows of one forward product.
// Each GPU the library runs on holds its own copy until the process ends;
// host code reaches it only through the address hipGetSymbolAddress gives on
// that device (the array's name is not a device address in host code).
constexpr size_t kNoallocScratchBytes = size_t(64) << 20;
static __device__ __align__(256) float scratch[kNoallocScratchBytes / sizeof(float)];

struct cbm_handle {
    cbm_desc d{};
    // host copies of the per-level tables
    std::vector<int32_t> lvl_ptr, lvl_nlong, fck_ptr, fred_ptr, rev_ptr, rev_nlong, rck_ptr, rred_ptr;
    float *ps_pos = nullptr;  // D A D only: d[node] / d[parent] per level position
    float *scratch = nullptr; // partial sums of long-row chunks
    int64_t scratch_floats = 0;
    int64_t slab_bytes = 0;   // column slab target (gathered bytes), 0 = whole rows
    int32_t kahan_level = 32; // levels from here on propagate with compensation
    int32_t fwd_split = -1;   // forward form: 0 fused per level, 1 split, -1 automatic
    int32_t wave_rows = -1;   // short-row passes: whole wavefront per row (1), never (0), by size (-1)
    int32_t mixed_grid = -1;  // chunks and short rows: one grid (1), two launches (0), by rule (-1)
    int32_t xcd_slabs = 1;    // feature slabs spread over the XCDs (1 = off)
    int32_t xcd_ranges = -1;  // column ranges per XCD: 1 on, 0 off, -1 by size, 8 / 4 that piece set
    int32_t xcd_phases = -1;  // column phases of a range pass: 1, 2, 4, or -1 automatic
    int32_t tree_tails = 1;   // tree tails: 1 deep fragments staged, 2 none staged, 3 all that fit, 0 per level
    int32_t *tf_order = nullptr;  // forward tail fragments: the stageable ones, deepest first, then the rest
    int32_t tf_staged_deep = 0, tf_rows_deep = 0;  // the stageable ones of at least kTailStagedMinSegs levels
    int32_t tf_staged_all = 0, tf_rows_all = 0;    // every stageable one
    // Column-range piece sets: [0] 8 ranges (one per XCD), [1] 4 ranges (each
    // on two XCDs, one column half each).
    struct RangeSet {
        int32_t n_ranges = 0;   // 0: the plan has no such pieces
        int32_t fwd = 0, bwd = 0;
        int32_t f_first[9] = {}, b_first[9] = {};  // host copies of the piece offsets per range
        const int32_t *fpc_b = nullptr, *fpc_e = nullptr, *fpc_slot = nullptr, *frs_slot = nullptr,
                      *frs_cnt = nullptr;
        const int32_t *bpc_b = nullptr, *bpc_e = nullptr, *bpc_mid = nullptr, *bpc_slot = nullptr,
                      *brs_slot = nullptr, *brs_cnt = nullptr, *bpc_row = nullptr;
    } rs[2];
    float *partial = nullptr; // one partial row per range piece
    int64_t partial_floats = 0;
    float *comp = nullptr;    // compensation rows of those levels
    int64_t comp_floats = 0;
    int32_t reserved_cols = 0; // widest product the buffers serve (cbm_reserve)
    int32_t symmetric = 0;     // A = A^T declared: the adjoint runs the forward structure
    cbm_handle *plain = nullptr; // the same product over A itself (no tree), borrowed (cbm_attach_uncompressed)
    int32_t plain_mode = -1;     // products over `plain`: 1 always, 0 never, -1 by the cost model
    int32_t fused = -1;        // one-launch forward: 1 on where possible, 0 off, -1 automatic (off)
    uint32_t *fused_state = nullptr; // ticket, done, epoch, failed waits (cbm_reserve)
    uint32_t *fused_flags = nullptr; // a flag per unit and 64-column slab
    int64_t fused_flag_count = 0;
    const int32_t *order = nullptr; // gather order of the short rows, borrowed (cbm_attach_order)
    int32_t *order_meta = nullptr;  // per ordered row (position, first entry, end, output row), owned
    int32_t n_order = 0;
    int32_t order_mode = 0;    // Delta pass in the gather order: 1 on, 0 off, -1 automatic
    uint32_t *arrive = nullptr; // a counter per long-row scratch slot, zero between launches (cbm_reserve)
    int64_t arrive_count = 0;
    int32_t chunk_finish = -1; // multi-chunk rows finished by their last chunk: 1 on, 0 off, -1 by size
    // Small or sparse matrices: the forward Delta pass's own long-row tables
    // (rows of more than kSmallLongRow entries in chunks of about
    // kSmallChunk), by position, with slots numbered over all levels; one
    // owned allocation, built at the first cbm_reserve.
    bool sc_on = false;
    int32_t *sc_buf = nullptr;
    const int32_t *sc_ck_pos = nullptr, *sc_ck_meta = nullptr, *sc_ck_slot = nullptr;
    const int32_t *sc_red_pos = nullptr, *sc_red_slot = nullptr, *sc_red_cnt = nullptr;
    std::vector<int32_t> sc_ck_lvl, sc_red_lvl;  // first chunk / reduction of each level (L + 1)
    int32_t sc_slots = 0;
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

// First level position that carries a compensation row.  Level 0 has no
// parent term, so compensation starts at level max(kahan_level, 1).
int32_t comp_base(const cbm_handle *h) {
    const int32_t first = std::max<int32_t>(h->kahan_level, 1);
    return h->kahan_level >= 0 && first < h->d.n_levels ? h->lvl_ptr[size_t(first)] : h->d.n_rows;
}

bool kahan_at(const cbm_handle *h, int32_t level) {
    return h->kahan_level >= 0 && level >= h->kahan_level && level > 0;
}

int ensure_comp(cbm_handle *h, int32_t n_cols) {
    const int64_t need = int64_t(h->d.n_rows - comp_base(h)) * n_cols;
    if (need <= h->comp_floats) return CBM_OK;
    if (h->comp) {
        const hipError_t e = hipFree(h->comp);
        h->comp = nullptr;
        h->comp_floats = 0;
        if (e != hipSuccess) return device_error(e, "cbm compensation rows");
    }
    const hipError_t e = hipMalloc(&h->comp, sizeof(float) * size_t(need));
    if (e != hipSuccess) {
        h->comp = nullptr;
        return device_error(e, "cbm compensation rows");
    }
    h->comp_floats = need;
    return CBM_OK;
}

using RangeSet = cbm_handle::RangeSet;

// The piece set the gather passes of an n_cols-wide call run by XCD column
// range with, or null for none; elem is the byte size of the gathered
// matrix's entries (2 for BF16 X or G, 4 for FP32).  Automatic: rows of 64
// entries or more on average (so that a row's pieces are not mostly
// partial-row overhead) and a gathered matrix of 24 to 256 MiB (smaller ones
// hit the L2s well enough without the partial rows: N = 132,534 at F = 64
// measured 1.70 ms per BF16 step without ranges, 1.79 ms with, at 16 MiB of
// BF16 X; its FP32 step, 32 MiB, 2.58 -> 2.06 ms; larger ones leave each
// XCD's range far beyond its L2).  The 4-range set halves the partial rows
// and each gather: it is taken when half a row is 256 bytes or more (F = 256:
// BF16 step 4.40 -> 4.17 ms, FP32 7.51 -> 7.33 ms; at F = 64 FP32, 128-byte
// halves, 2.06 -> 2.18 ms, so the 8-range set stays there).
const RangeSet *pick_ranges(const cbm_handle *h, int32_t n_cols, int elem) {
    auto usable = [](const RangeSet &r) { return r.fwd > 0 && r.bwd > 0; };
    const RangeSet &r8 = h->rs[0], &r4 = h->rs[1];
    switch (h->xcd_ranges) {
        case 0: return nullptr;
        case 8: return usable(r8) ? &r8 : nullptr;
        case 4: return usable(r4) ? &r4 : nullptr;
        case -1: {
            const int64_t bytes = int64_t(h->d.n_rows) * n_cols * elem;
            if (h->d.nnz < 64 * int64_t(h->d.n_rows) || bytes < (int64_t(24) << 20) || bytes > (int64_t(256) << 20))
                return nullptr;
            break;
        }
        default: break;
    }
    if (int64_t(n_cols) * elem >= 512 && usable(r4)) return &r4;
    if (usable(r8)) return &r8;
    return usable(r4) ? &r4 : nullptr;
}

int elem_of(int32_t dtype) { return dtype == CBM_BF16 ? 2 : 4; }

// Column phases of a range pass with piece set rs: each XCD gathers its
// range one column slab after the other, so its L2 holds a smaller working
// set, at the price of narrower gathers and one more read of every index
// per phase.  Automatic: 2 when an XCD's share of the gathered matrix (N x F
// x elem / 8) is 12 MiB or more and a phase's gather stays 256 bytes or more
// (N = 132,534, F = 256 FP32, 17 MiB per XCD: step 7.34 -> 6.21 ms with the
// 4-range set; 4 phases, or 2 at 8.5 MiB per XCD (the BF16 step), measured
// slower than 1).
int32_t range_phases(const cbm_handle *h, const RangeSet &rs, int32_t n_cols, int elem) {
    if (h->xcd_phases > 0) return h->xcd_phases;
    const int64_t per_xcd = int64_t(h->d.n_rows) * n_cols * elem / 8;
    const int64_t gather = int64_t(n_cols) * elem / (8 / rs.n_ranges * 2);
    return per_xcd >= (int64_t(12) << 20) && gather >= 256 ? 2 : 1;
}

// The partial rows are sized for either storage type and either piece set,
// so that a reserved handle serves BF16 and FP32 calls of this width without
// allocating.
int ensure_partial(cbm_handle *h, int32_t n_cols) {
    if (!pick_ranges(h, n_cols, 2) && !pick_ranges(h, n_cols, 4)) return CBM_OK;
    const int64_t need =
        int64_t(std::max({h->rs[0].fwd, h->rs[0].bwd, h->rs[1].fwd, h->rs[1].bwd})) * n_cols;
    if (need <= h->partial_floats) return CBM_OK;
    if (h->partial) {
        const hipError_t e = hipFree(h->partial);
        h->partial = nullptr;
        h->partial_floats = 0;
        if (e != hipSuccess) return device_error(e, "cbm range partials");
    }
    const hipError_t e = hipMalloc(&h->partial, sizeof(float) * size_t(need));
    if (e != hipSuccess) {
        h->partial = nullptr;
        return device_error(e, "cbm range partials");
    }
    h->partial_floats = need;
    return CBM_OK;
}

// The state words and flags of the one-launch forward for n_cols columns
// (all zero when allocated; the kernel keeps them consistent after that).
int ensure_fused(cbm_handle *h, int32_t n_cols) {
    if (h->d.n_units <= 0 || n_cols <= 0) return CBM_OK;
    hipError_t e = hipSuccess;
    if (!h->fused_state) {
        e = hipMalloc(&h->fused_state, 4 * sizeof(uint32_t));
        if (e == hipSuccess) e = hipMemset(h->fused_state, 0, 4 * sizeof(uint32_t));
        if (e != hipSuccess) {
            if (h->fused_state) (void)hipFree(h->fused_state);
            h->fused_state = nullptr;
            return device_error(e, "cbm fused state");
        }
    }
    const int64_t need = int64_t(h->d.n_units) * ((n_cols + 63) / 64);
    if (need <= h->fused_flag_count) return CBM_OK;
    if (h->fused_flags) {
        e = hipFree(h->fused_flags);
        h->fused_flags = nullptr;
        h->fused_flag_count = 0;
        if (e != hipSuccess) return device_error(e, "cbm fused flags");
    }
    e = hipMalloc(&h->fused_flags, sizeof(uint32_t) * size_t(need));
    if (e == hipSuccess) e = hipMemset(h->fused_flags, 0, sizeof(uint32_t) * size_t(need));
    if (e != hipSuccess) {
        if (h->fused_flags) (void)hipFree(h->fused_flags);
        h->fused_flags = nullptr;
        return device_error(e, "cbm fused flags");
    }
    h->fused_flag_count = need;
    return CBM_OK;
}

// The arrival counters of the long-row chunks (one per scratch slot, zero
// when allocated; each launch leaves them zero).
int ensure_arrive(cbm_handle *h, int64_t slots) {
    if (slots <= h->arrive_count) return CBM_OK;
    if (h->arrive) {
        const hipError_t e = hipFree(h->arrive);
        h->arrive = nullptr;
        h->arrive_count = 0;
        if (e != hipSuccess) return device_error(e, "cbm chunk counters");
    }
    hipError_t e = hipMalloc(&h->arrive, sizeof(uint32_t) * size_t(slots));
    if (e == hipSuccess) e = hipMemset(h->arrive, 0, sizeof(uint32_t) * size_t(slots));
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        if (h->arrive) (void)hipFree(h->arrive);
        h->arrive = nullptr;
        return device_error(e, "cbm chunk counters");
    }
    h->arrive_count = slots;
    return CBM_OK;
}

int grow_buffers(cbm_handle *h, int32_t n_cols) {
    int rc = ensure_fused(h, n_cols);
    if (rc == CBM_OK) rc = ensure_comp(h, n_cols);
    if (rc == CBM_OK) rc = ensure_partial(h, n_cols);
    if (rc != CBM_OK) return rc;
    const int64_t slots =
        std::max<int64_t>({h->d.fwd_slots, h->d.bwd_slots, h->d.fwd_gslots, h->d.rev_slots, h->sc_slots});
    rc = ensure_arrive(h, slots);
    if (rc != CBM_OK) return rc;
    const int64_t need = slots * int64_t(n_cols);
    if (need <= h->scratch_floats) return CBM_OK;
    if (h->scratch) {
        const hipError_t e = hipFree(h->scratch);
        h->scratch = nullptr;
        h->scratch_floats = 0;
        if (e != hipSuccess) return device_error(e, "cbm scratch");
    }
    const hipError_t e = hipMalloc(&h->scratch, sizeof(float) * size_t(need));
    if (e != hipSuccess) {
        h->scratch = nullptr;
        return device_error(e, "cbm scratch");
    }
    h->scratch_floats = need;
    return CBM_OK;
}

// Product calls allocate nothing: their buffers are sized by cbm_reserve (and
// kept sized by the settings that change what a width needs), so a call
// wider than the reserved width is refused instead of growing them -- a
// hipMalloc in a call would synchronise the device and fails under HIP-graph
// capture.
int check_reserved(const cbm_handle *h, int32_t n_cols, const char *what) {
    if (n_cols <= h->reserved_cols) return CBM_OK;
    set_error("%s: %d columns exceed the reserved width %d; call cbm_reserve(h, %d) first", what, n_cols,
              h->reserved_cols, n_cols);
    return CBM_ERR_ARG;
}

// The device buffers a forward call works in: long-row partial sums,
// compensation rows, range partial rows.  The handle's own (cbm_reserve) or
// the library's static scratch (cbm_matmul_noalloc).
struct Bufs {
    float *scratch = nullptr;
    float *comp = nullptr;
    float *partial = nullptr;
};

Bufs handle_bufs(const cbm_handle *h) { return Bufs{h->scratch, h->comp, h->partial}; }

// The scalings a forward-structure pass applies: the handle's own product
// (CBM_MODE_*), or kFormDAX = D^-1/2 A, which is (A D^-1/2)^T for a
// symmetric A: Delta rows scaled by D and the propagation factor D_v / D_parent,
// with no column scale.
constexpr int32_t kFormDAX = 3;

bool form_rowscaled(int32_t form) { return form == CBM_MODE_DADX || form == kFormDAX; }
bool form_colscaled(int32_t form) { return form == CBM_MODE_ADX || form == CBM_MODE_DADX; }

// Device address of `scratch` on each device, looked up once per device, by
// the first cbm_matmul_noalloc call with that device current.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float *g_scratch_addr[kMaxDevices] = {};

int noalloc_scratch_address(float **out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return device_error(e, "cbm scratch: hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("cbm scratch: device %d beyond the %d the library serves", dev, kMaxDevices);
        return CBM_ERR_DEVICE;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_addr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return device_error(e, "cbm scratch: hipGetSymbolAddress");
        g_scratch_addr[dev] = static_cast<float *>(p);
    }
    *out = g_scratch_addr[dev];
    return CBM_OK;
}

// Small or sparse matrices run their forward Delta pass with long-row tables
// of their own: rows of more than kSmallLongRow entries (the plan's limit is
// CBM_LONG_ROW) in chunks of about kSmallChunk (CBM_CHUNK).  Such a pass
// does not fill the device, or fills it with short rows, and its time is its
// longest chains of dependent gathers: a row of 256 entries on one wavefront,
// or a 1024-entry chunk, is 16 steps of a lane group, a 256-entry chunk 4.
// N = 19,717, F = 128: step 0.0377 -> 0.0295 ms BF16; N = 169,343 with 14
// entries per row: 0.209 -> 0.176 ms.  Matrices of long rows (N = 132,534
// with 300 entries per row: 1.23 -> 1.52 ms at F = 64) keep the plan's
// tables, and so do the adjoint's reverse and Delta^T passes.
constexpr int32_t kSmallChunkRows = 1 << 16;  // any matrix up to this size
constexpr int32_t kSparseChunkRows = 1 << 18; // up to this size with at most
constexpr int64_t kSparseChunkAvg = 32;       //   this many entries per row
constexpr int32_t kSmallLongRow = 64;
constexpr int32_t kSmallChunk = 256;

bool small_tables_wanted(const cbm_handle *h) {
    const int64_t n = h->d.n_rows;
    return n <= kSmallChunkRows || (n <= kSparseChunkRows && h->d.nnz <= kSparseChunkAvg * n);
}

// Builds the small-matrix tables from the plan's forward row offsets.  Runs
// once, at the first reservation (it reads the offsets back from the device,
// so it waits for the device); until then the plan's tables serve.
int build_small_chunks(cbm_handle *h) {
    const int32_t n = h->d.n_rows, L = h->d.n_levels;
    std::vector<int32_t> ptr(size_t(n) + 1);
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(ptr.data(), h->d.fwd_ptr, sizeof(int32_t) * ptr.size(), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return device_error(e, "cbm_reserve (small-matrix chunks)");
    std::vector<int32_t> ck_pos, ck_meta, ck_slot, red_pos, red_slot, red_cnt;
    h->sc_ck_lvl.assign(size_t(L) + 1, 0);
    h->sc_red_lvl.assign(size_t(L) + 1, 0);
    int32_t slots = 0;
    for (int32_t l = 0; l < L; ++l) {
        h->sc_ck_lvl[size_t(l)] = int32_t(ck_pos.size());
        h->sc_red_lvl[size_t(l)] = int32_t(red_pos.size());
        for (int32_t p = h->lvl_ptr[size_t(l)]; p < h->lvl_ptr[size_t(l) + 1]; ++p) {
            const int64_t len = int64_t(ptr[size_t(p) + 1]) - ptr[size_t(p)];
            if (len <= kSmallLongRow) continue;
            const int64_t cnt = std::min<int64_t>((len + kSmallChunk - 1) / kSmallChunk, 0xFFFF);
            for (int64_t c = 0; c < cnt; ++c) {
                ck_pos.push_back(p);
                ck_meta.push_back(int32_t(uint32_t(c) | (uint32_t(cnt) << 16)));
                ck_slot.push_back(cnt > 1 ? slots + int32_t(c) : -1);
            }
            if (cnt > 1) {
                red_pos.push_back(p);
                red_slot.push_back(slots);
                red_cnt.push_back(int32_t(cnt));
                slots += int32_t(cnt);
            }
        }
    }
    h->sc_ck_lvl[size_t(L)] = int32_t(ck_pos.size());
    h->sc_red_lvl[size_t(L)] = int32_t(red_pos.size());
    const size_t nc = ck_pos.size(), nr = red_pos.size();
    if (nc + nr > 0) {
        std::vector<int32_t> all;
        all.reserve(3 * (nc + nr));
        for (const auto *v : {&ck_pos, &ck_meta, &ck_slot, &red_pos, &red_slot, &red_cnt})
            all.insert(all.end(), v->begin(), v->end());
        e = hipMalloc(&h->sc_buf, sizeof(int32_t) * all.size());
        if (e == hipSuccess) e = hipMemcpy(h->sc_buf, all.data(), sizeof(int32_t) * all.size(), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            if (h->sc_buf) (void)hipFree(h->sc_buf);
            h->sc_buf = nullptr;
            return device_error(e, "cbm_reserve (small-matrix chunks)");
        }
        h->sc_ck_pos = h->sc_buf;
        h->sc_ck_meta = h->sc_buf + nc;
        h->sc_ck_slot = h->sc_buf + 2 * nc;
        h->sc_red_pos = h->sc_buf + 3 * nc;
        h->sc_red_slot = h->sc_buf + 3 * nc + nr;
        h->sc_red_cnt = h->sc_buf + 3 * nc + 2 * nr;
    }
    h->sc_slots = slots;
    h->sc_on = true;
    return CBM_OK;
}

// By default a pass finishes its multi-chunk rows in the chunk kernel when
// the matrix has at most kChunkFinishMaxRows rows and the pass at most
// kChunkFinishMaxLong such rows: there the reduction's launch is a large part
// of a short pass (N = 19,717, F = 128: 22 -> 19 us per product).  Each
// arrival releases its wave's stores at device scope and each last chunk
// acquires, which invalidates its XCD's L2; with many multi-chunk rows that
// costs the L2 hits of the gathers (N = 132,534 with 80M entries, F = 64:
// 0.61 -> 1.11 ms per product), so larger passes keep the reduction launch.
constexpr int32_t kChunkFinishMaxRows = 1 << 18;
constexpr int32_t kChunkFinishMaxLong = 64;

// Whether a range's n_multi multi-chunk rows are finished by their last
// chunk (ChunkArgs::arrive) instead of a reduction launch.
bool chunk_finish_on(const cbm_handle *h, const cbmk::RowPass &rp, int64_t n_multi) {
    if (!h->arrive || rp.xcd_slabs != 1 || rp.ranges || h->chunk_finish == 0) return false;
    return h->chunk_finish == 1 || (h->d.n_rows <= kChunkFinishMaxRows && n_multi <= kChunkFinishMaxLong);
}

// One ordered row range whose first n_long rows are long: the chunk kernel
// and the short-row pass (one grid), then the reduction of multi-chunk rows
// (or none: their last chunks finish them).
int run_range(cbm_handle *h, cbmk::RowPass rp, int32_t p_begin, int32_t p_end, int32_t n_long,
              cbmk::ChunkArgs ck, const cbmk::ReduceArgs &rd, hipStream_t stream, const char *what) {
    hipError_t e = hipSuccess;
    rp.p_begin = p_begin + n_long;
    rp.p_end = p_end;
    const bool finish = rd.count > 0 && ck.count > 0 && chunk_finish_on(h, rp, rd.count);
    ck.arrive = finish ? h->arrive : nullptr;
    if (ck.count > 0) {
        int launched = 0;
        e = cbmk::launch_row_mixed(rp, ck, &launched, stream);
        if (e != hipSuccess) return device_error(e, what);
        h->last_launches += launched;
    } else if (rp.p_end > rp.p_begin) {
        e = cbmk::launch_row_pass(rp, stream);
        if (e != hipSuccess) return device_error(e, what);
        ++h->last_launches;
    }
    if (rd.count > 0 && !finish) {
        e = cbmk::launch_row_reduce(rp, rd, stream);
        if (e != hipSuccess) return device_error(e, what);
        ++h->last_launches;
    }
    return CBM_OK;
}

// Column slabs: the columns are processed in slabs so that the gathered
// matrix's slab (N x width x elem bytes) stays within slab_bytes and keeps
// hitting the device's last-level cache.  Slab starts stay 16-byte aligned.
template <class F>
int for_each_slab(const cbm_handle *h, int32_t n_cols, int elem, F &&fn) {
    int32_t width = n_cols;
    if (h->slab_bytes > 0) {
        const int32_t align = 16 / elem;
        const int64_t fit = h->slab_bytes / (int64_t(h->d.n_rows) * elem);
        if (fit < n_cols) width = std::max<int32_t>(align, int32_t(fit / align) * align);
    }
    for (int32_t c0 = 0; c0 < n_cols; c0 += width) {
        const int rc = fn(c0, std::min(width, n_cols - c0));
        if (rc != CBM_OK) return rc;
    }
    return CBM_OK;
}

template <typename T>
const T *shift(const void *p, int32_t c0, int elem) {
    return reinterpret_cast<const T *>(static_cast<const char *>(p) + int64_t(c0) * elem);
}

// The forward Delta pass over every level; with_parent adds the parent term.
int forward_levels(cbm_handle *h, const Bufs &b, cbmk::RowPass rp, bool with_parent, hipStream_t stream,
                   const char *what) {
    cbmk::ChunkArgs ck;
    ck.pos = h->d.fck_pos;
    ck.meta = h->d.fck_meta;
    ck.slot = h->d.fck_slot;
    ck.scratch = b.scratch;
    cbmk::ReduceArgs rd;
    rd.pos = h->d.fred_pos;
    rd.slot = h->d.fred_slot;
    rd.cnt = h->d.fred_cnt;
    rd.scratch = b.scratch;
    rp.parent_pos = h->d.lvl_ppos;
    rp.comp = b.comp;
    rp.comp_base = comp_base(h);
    rp.ldc = rp.ncols;
    if (h->sc_on) {
        // The small-matrix tables: a level's long rows are not at its front,
        // so its short-row pass covers the whole level and skips them.
        ck.pos = h->sc_ck_pos;
        ck.meta = h->sc_ck_meta;
        ck.slot = h->sc_ck_slot;
        rd.pos = h->sc_red_pos;
        rd.slot = h->sc_red_slot;
        rd.cnt = h->sc_red_cnt;
        rp.skip_above = kSmallLongRow;
    }
    for (int32_t l = 0; l < h->d.n_levels; ++l) {
        rp.extra = (with_parent && l > 0) ? (kahan_at(h, l) ? cbmk::kExtraParentKahan : cbmk::kExtraParent)
                                          : cbmk::kExtraNone;
        const std::vector<int32_t> &cl = h->sc_on ? h->sc_ck_lvl : h->fck_ptr;
        const std::vector<int32_t> &rl = h->sc_on ? h->sc_red_lvl : h->fred_ptr;
        ck.first = cl[size_t(l)];
        ck.count = cl[size_t(l) + 1] - ck.first;
        rd.first = rl[size_t(l)];
        rd.count = rl[size_t(l) + 1] - rd.first;
        const int rc = run_range(h, rp, h->lvl_ptr[size_t(l)], h->lvl_ptr[size_t(l) + 1],
                                 h->sc_on ? 0 : h->lvl_nlong[size_t(l)], ck, rd, stream, what);
        if (rc != CBM_OK) return rc;
    }
    return CBM_OK;
}

// Whether the Delta pass of every level at once runs in the attached gather
// order (cbm_attach_order): an order attached and not switched off, and no
// column slabs (their workgroup mapping is another one).
bool order_on(const cbm_handle *h) {
    return h->order_mode != 0 && h->n_order > 0 && h->slab_bytes == 0 && h->xcd_slabs <= 1;
}

// The Delta pass of every level at once (no parent term): all long-row
// chunks with slots numbered over the levels, one short-row pass over all
// positions that skips the long rows, then all reductions.
int delta_flat(cbm_handle *h, const Bufs &b, cbmk::RowPass rp, hipStream_t stream, const char *what) {
    const int32_t L = h->d.n_levels;
    cbmk::ChunkArgs ck;
    ck.pos = h->d.fck_pos;
    ck.meta = h->d.fck_meta;
    ck.slot = h->d.fck_gslot;
    ck.first = 0;
    ck.count = h->fck_ptr[size_t(L)];
    ck.scratch = b.scratch;
    cbmk::ReduceArgs rd;
    rd.pos = h->d.fred_pos;
    rd.slot = h->d.fred_gslot;
    rd.cnt = h->d.fred_cnt;
    rd.first = 0;
    rd.count = h->fred_ptr[size_t(L)];
    rd.scratch = b.scratch;
    rp.extra = cbmk::kExtraNone;
    rp.skip_above = CBM_LONG_ROW;
    if (h->sc_on) {
        ck.pos = h->sc_ck_pos;
        ck.meta = h->sc_ck_meta;
        ck.slot = h->sc_ck_slot;
        ck.count = h->sc_ck_lvl[size_t(L)];
        rd.pos = h->sc_red_pos;
        rd.slot = h->sc_red_slot;
        rd.cnt = h->sc_red_cnt;
        rd.count = h->sc_red_lvl[size_t(L)];
        rp.skip_above = kSmallLongRow;
    }
    if (order_on(h)) {
        // The short rows in the gather order, XCD by XCD; the long rows by
        // their chunks as ever.
        rp.order = h->order;
        rp.order_meta = h->order_meta;
        return run_range(h, rp, 0, h->n_order, 0, ck, rd, stream, what);
    }
    return run_range(h, rp, 0, h->d.n_rows, 0, ck, rd, stream, what);
}

bool tails_on(const cbm_handle *h) { return h->tree_tails != 0; }

// Forward tail fragments of at least this many levels are staged in LDS by
// default (their chains of dependent levels are the tail's critical path;
// the short ones cost one or two global round trips either way), provided
// the deepest of them has kTailStagedMinChain levels: the staged launch is
// one launch more, which a chain of 41 levels repays (N = 132,534: 66 ->
// 52 us per tail) and one of 19 does not (N = 12,008: 3.5 us lost).
constexpr int32_t kTailStagedMinSegs = 8;
constexpr int32_t kTailStagedMinChain = 24;

// The forward tail of the handle's tail mode: which fragments are staged.
cbmk::TailArgs forward_tail(const cbm_handle *h) {
    cbmk::TailArgs t;
    t.frag = h->d.tf_frag;
    t.seg = h->d.tf_seg;
    t.pos = h->d.tf_pos;
    t.n_frag = h->d.tf_frags;
    if (h->tf_order && (h->tree_tails == 1 || h->tree_tails == 3)) {
        t.order = h->tf_order;
        t.n_staged = h->tree_tails == 1 ? h->tf_staged_deep : h->tf_staged_all;
        t.staged_rows = h->tree_tails == 1 ? h->tf_rows_deep : h->tf_rows_all;
    }
    return t;
}

int32_t tail_launches(const cbmk::TailArgs &t) {
    const int32_t staged = t.order ? t.n_staged : 0;
    return (staged > 0 ? 1 : 0) + (staged < t.n_frag ? 1 : 0);
}

// In-place propagation of levels 1 .. L-1, y_v += s_v y_parent(v), with the
// compensated form from the handle's Kahan level on: one launch per level
// down to the tree tail, then one launch for the tail.
int propagate_levels(cbm_handle *h, const Bufs &b, int32_t form, float *y, int64_t ldy, int32_t n_cols,
                     hipStream_t stream, const char *what) {
    const float *ps = form_rowscaled(form) ? h->ps_pos : nullptr;
    cbmk::CompArgs cm;
    cm.parent_pos = h->d.lvl_ppos;
    cm.comp = b.comp;
    cm.base = comp_base(h);
    cm.ldc = n_cols;
    const int32_t L = h->d.n_levels;
    const bool tail = tails_on(h) && h->d.tf_level < L;
    for (int32_t l = 1; l < (tail ? h->d.tf_level : L); ++l) {
        const int32_t b = h->lvl_ptr[size_t(l)], e = h->lvl_ptr[size_t(l) + 1];
        const hipError_t err =
            kahan_at(h, l) ? cbmk::launch_propagate_comp(b, e, h->d.lvl_node, h->d.lvl_parent, ps, cm, y, ldy, n_cols,
                                                         stream)
                           : cbmk::launch_propagate(b, e, h->d.lvl_node, h->d.lvl_parent, ps, y, ldy, n_cols, stream);
        if (err != hipSuccess) return device_error(err, what);
        ++h->last_launches;
    }
    if (tail) {
        const cbmk::TailArgs t = forward_tail(h);
        const bool comp = h->kahan_level >= 0 && std::max<int32_t>(h->kahan_level, 1) < L;
        const hipError_t err = cbmk::launch_propagate_tail(t, h->d.lvl_node, h->d.lvl_parent, ps, cm,
                                                           comp ? cm.base : INT32_MAX, y, ldy, n_cols, stream);
        if (err != hipSuccess) return device_error(err, what);
        h->last_launches += tail_launches(t);
    }
    return CBM_OK;
}

// Copies the n_ranges + 1 piece offsets of a range table into a pass.
void set_ranges(cbmk::RowPass &rp, const int32_t *first, int32_t n_ranges, int32_t phases) {
    rp.ranges = true;
    rp.range_count = n_ranges;
    rp.range_phases = phases;
    for (int k = 0; k <= n_ranges; ++k) rp.range_first[k] = first[k];
    rp.xcd_slabs = 8 / n_ranges * phases;
    rp.wave_rows = 0;
    rp.skip_above = INT32_MAX;
}

// The forward by XCD column range: one pass over all forward pieces of set
// rs (each gathers one range of X into its partial row), then a
// reduction that adds each row's pieces in order and applies the row scale --
// per tree level with the parent term (compensated from the Kahan level on)
// for the product, or over all rows at once without it for the Delta half
// alone (the level propagation then adds the same parent terms, so the two
// halves give the product's bits).
int forward_ranges(cbm_handle *h, const Bufs &b, const RangeSet &rs, const cbmk::RowPass &base, bool with_parent,
                   hipStream_t stream, const char *what) {
    cbmk::RowPass pc = base;
    pc.node_of = rs.fpc_slot;
    pc.parent_of = nullptr;
    pc.ptr = rs.fpc_b;
    pc.ptr_end = rs.fpc_e;
    pc.rowscale = nullptr;
    pc.escale = nullptr;
    pc.extra = cbmk::kExtraNone;
    pc.y = b.partial;
    pc.ldy = base.ncols;
    set_ranges(pc, rs.f_first, rs.n_ranges, range_phases(h, rs, base.ncols, base.g_bf16 ? 2 : 4));
    hipError_t e = cbmk::launch_row_pass(pc, stream);
    if (e != hipSuccess) return device_error(e, what);
    ++h->last_launches;
    cbmk::RowPass rr = base;
    rr.g = b.partial;
    rr.ldg = base.ncols;
    rr.g_bf16 = false;
    rr.parent_pos = h->d.lvl_ppos;
    rr.comp = b.comp;
    rr.comp_base = comp_base(h);
    rr.ldc = base.ncols;
    cbmk::ReduceArgs rd;
    rd.pos = nullptr;
    rd.slot = rs.frs_slot;
    rd.cnt = rs.frs_cnt;
    rd.scratch = b.partial;
    const int32_t levels = with_parent ? h->d.n_levels : 1;
    for (int32_t l = 0; l < levels; ++l) {
        rr.extra = (with_parent && l > 0) ? (kahan_at(h, l) ? cbmk::kExtraParentKahan : cbmk::kExtraParent)
                                          : cbmk::kExtraNone;
        rd.first = with_parent ? h->lvl_ptr[size_t(l)] : 0;
        rd.count = with_parent ? h->lvl_ptr[size_t(l) + 1] - rd.first : h->d.n_rows;
        e = cbmk::launch_row_reduce(rr, rd, stream);
        if (e != hipSuccess) return device_error(e, what);
        ++h->last_launches;
    }
    return CBM_OK;
}

// Which forward form a call uses.  The split form (one Delta pass over all
// levels, then one light propagation launch per level) keeps the whole chip
// busy on the gathers however thin the levels are, at the price of one more
// FP32 read and write of each non-root row; both forms give the same bits.
bool use_split(const cbm_handle *h) {
    if (h->fwd_split >= 0) return h->fwd_split == 1;
    return h->d.n_levels > 1;
}

cbmk::RowPass forward_pass(const cbm_handle *h, int32_t form, int32_t dtype, const void *x, int64_t ldx, float *y,
                           int64_t ldy,
                           int32_t n_cols) {
    cbmk::RowPass rp;
    rp.node_of = h->d.lvl_node;
    rp.parent_of = h->d.lvl_parent;
    rp.ptr = h->d.fwd_ptr;
    rp.idx = h->d.fwd_col;
    rp.colscale = form_colscaled(form) ? h->d.d : nullptr;
    rp.rowscale = form_rowscaled(form) ? h->d.d : nullptr;
    rp.escale = form_rowscaled(form) ? h->ps_pos : nullptr;
    rp.g = x;
    rp.ldg = ldx;
    rp.g_bf16 = dtype == CBM_BF16;
    rp.e = y;
    rp.lde = ldy;
    rp.e_bf16 = false;
    rp.y = y;
    rp.ldy = ldy;
    rp.ncols = n_cols;
    rp.rows_hint = h->d.n_rows;
    rp.wave_rows = h->wave_rows;
    rp.mixed_grid = h->mixed_grid;
    rp.xcd_slabs = h->xcd_slabs;
    return rp;
}

// Orders the forward tail's fragments for staging: those whose rows fit the
// LDS of a staged workgroup (with compensation rows, the larger case) first,
// most levels first, then the rest in plan order; counts the ones to stage
// by default (at least kTailStagedMinSegs levels) and in all, with the rows
// the largest of each set needs (a multiple of 4, for 16-byte LDS rows).
// Runs once, at the first reservation (it reads the plan's tail tables back
// from the device, so it waits for the device).
int stage_forward_tail(cbm_handle *h) {
    const int32_t nf = h->d.tf_frags;
    std::vector<int32_t> frag(size_t(nf) + 1);
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(frag.data(), h->d.tf_frag, sizeof(int32_t) * frag.size(), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return device_error(e, "cbm_reserve (tree tail)");
    bool ok = frag[0] == 0;
    for (int32_t f = 0; ok && f < nf; ++f) ok = frag[size_t(f) + 1] >= frag[size_t(f)];
    if (!ok) {
        set_error("cbm_reserve: forward tree tail fragments are not ascending");
        return CBM_ERR_ARG;
    }
    std::vector<int32_t> seg(size_t(frag[size_t(nf)]) + 1);
    e = hipMemcpy(seg.data(), h->d.tf_seg, sizeof(int32_t) * seg.size(), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return device_error(e, "cbm_reserve (tree tail)");
    for (size_t k = 0; ok && k + 1 < seg.size(); ++k) ok = seg[k + 1] >= seg[k];
    if (!ok || seg[0] != 0) {
        set_error("cbm_reserve: forward tree tail segments are not ascending");
        return CBM_ERR_ARG;
    }
    auto round4 = [](int32_t r) { return (r + 3) / 4 * 4; };
    std::vector<int32_t> fit, rest;
    for (int32_t f = 0; f < nf; ++f) {
        const int32_t rows = seg[size_t(frag[size_t(f) + 1])] - seg[size_t(frag[size_t(f)])];
        const bool fits = rows > 0 && cbmk::tail_staged_bytes(round4(rows), true) <= cbmk::kTailStagedMaxBytes;
        (fits ? fit : rest).push_back(f);
    }
    auto levels = [&](int32_t f) { return frag[size_t(f) + 1] - frag[size_t(f)]; };
    auto rows_of = [&](int32_t f) { return seg[size_t(frag[size_t(f) + 1])] - seg[size_t(frag[size_t(f)])]; };
    std::stable_sort(fit.begin(), fit.end(), [&](int32_t a, int32_t b) { return levels(a) > levels(b); });
    h->tf_staged_all = int32_t(fit.size());
    h->tf_staged_deep = 0;
    h->tf_rows_all = h->tf_rows_deep = 0;
    const bool chain = !fit.empty() && levels(fit.front()) >= kTailStagedMinChain;
    for (const int32_t f : fit) {
        h->tf_rows_all = std::max(h->tf_rows_all, round4(rows_of(f)));
        if (chain && levels(f) >= kTailStagedMinSegs) {
            ++h->tf_staged_deep;
            h->tf_rows_deep = std::max(h->tf_rows_deep, round4(rows_of(f)));
        }
    }
    fit.insert(fit.end(), rest.begin(), rest.end());
    int32_t *order = nullptr;
    e = hipMalloc(&order, sizeof(int32_t) * size_t(std::max(nf, 1)));
    if (e == hipSuccess) e = hipMemcpy(order, fit.data(), sizeof(int32_t) * fit.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (order) (void)hipFree(order);
        return device_error(e, "cbm_reserve (tree tail)");
    }
    h->tf_order = order;
    return CBM_OK;
}

}  // namespace

extern "C" const char *cbm_last_error(void) { return g_error.c_str(); }

extern "C" int32_t cbm_desc_size(void) { return int32_t(sizeof(cbm_desc)); }

extern "C" int cbm_create(const cbm_desc *desc, hipStream_t stream, cbm_handle **out) {
    if (!desc || !out) {
        set_error("cbm_create: null argument");
        return CBM_ERR_ARG;
    }
    *out = nullptr;
    const cbm_desc &d = *desc;
    const bool tables = d.lvl_ptr_host && d.lvl_nlong_host && d.fck_ptr_host && d.fred_ptr_host && d.rev_ptr_host &&
                        d.rev_nlong_host && d.rck_ptr_host && d.rred_ptr_host;
    if (d.n_rows < 1 || d.n_levels < 1 || d.nnz < 0 || d.n_edges < 0 || d.n_edges >= d.n_rows || !tables ||
        !d.lvl_node || !d.lvl_parent || !d.lvl_ppos || !d.fwd_ptr || !d.bwd_ptr || !d.bwd_node || !d.bwd_mid || !d.ch_ptr ||
        (d.nnz > 0 && (!d.fwd_col || !d.bwd_row)) || (d.n_edges > 0 && (!d.ch_node || !d.rev_node || !d.ch_mid)) ||
        d.fwd_slots < 0 ||
        d.bwd_slots < 0 || d.n_bck < 0 || d.n_bred < 0 || d.bwd_nlong < 0 || d.bwd_nlong > d.n_rows ||
        (d.n_bck > 0 && (!d.bck_pos || !d.bck_meta || !d.bck_slot)) ||
        (d.n_bred > 0 && (!d.bred_pos || !d.bred_slot || !d.bred_cnt))) {
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
    const int32_t L = d.n_levels;
    if (d.lvl_ptr_host[0] != 0 || d.lvl_ptr_host[L] != d.n_rows || d.fck_ptr_host[0] != 0 || d.fred_ptr_host[0] != 0 ||
        d.rev_ptr_host[0] != 0 || d.rev_ptr_host[L] > d.n_rows || d.rck_ptr_host[0] != 0 || d.rred_ptr_host[0] != 0) {
        set_error("cbm_create: level tables do not start at 0 / end at n_rows");
        return CBM_ERR_ARG;
    }
    for (int32_t l = 0; l < L; ++l) {
        if (d.lvl_ptr_host[l + 1] <= d.lvl_ptr_host[l] || d.lvl_nlong_host[l] < 0 ||
            d.lvl_nlong_host[l] > d.lvl_ptr_host[l + 1] - d.lvl_ptr_host[l] ||
            d.fck_ptr_host[l + 1] < d.fck_ptr_host[l] || d.fred_ptr_host[l + 1] < d.fred_ptr_host[l] ||
            d.rev_ptr_host[l + 1] < d.rev_ptr_host[l] || d.rev_nlong_host[l] < 0 ||
            d.rev_nlong_host[l] > d.rev_ptr_host[l + 1] - d.rev_ptr_host[l] ||
            d.rck_ptr_host[l + 1] < d.rck_ptr_host[l] || d.rred_ptr_host[l + 1] < d.rred_ptr_host[l]) {
            set_error("cbm_create: inconsistent level tables at level %d", l);
            return CBM_ERR_ARG;
        }
    }
    if ((d.fck_ptr_host[L] > 0 && (!d.fck_pos || !d.fck_meta || !d.fck_slot || !d.fck_gslot)) ||
        (d.fred_ptr_host[L] > 0 && (!d.fred_pos || !d.fred_slot || !d.fred_cnt || !d.fred_gslot)) ||
        d.fwd_gslots < 0 || d.rev_slots < 0 ||
        (d.rck_ptr_host[L] > 0 && (!d.rck_pos || !d.rck_meta || !d.rck_slot)) ||
        (d.rred_ptr_host[L] > 0 && (!d.rred_pos || !d.rred_slot || !d.rred_cnt))) {
        set_error("cbm_create: missing forward chunk tables");
        return CBM_ERR_ARG;
    }
    for (int k = 0; k < 2; ++k) {
        const int32_t level = k == 0 ? d.tf_level : d.tr_level, frags = k == 0 ? d.tf_frags : d.tr_frags;
        const int32_t *arrs[3] = {k == 0 ? d.tf_frag : d.tr_frag, k == 0 ? d.tf_seg : d.tr_seg,
                                  k == 0 ? d.tf_pos : d.tr_pos};
        const bool bad_level = level < (k == 0 ? 1 : 0) || level > L;
        const bool bad_tables = level < L && (frags < 1 || !arrs[0] || !arrs[1] || !arrs[2]);
        bool long_lists = false;  // the reverse tail holds no long child lists (they take the chunk kernels)
        if (k == 1 && !bad_level)
            for (int32_t l = level; l < L; ++l) long_lists = long_lists || d.rev_nlong_host[l] > 0;
        if (bad_level || bad_tables || long_lists) {
            set_error("cbm_create: invalid %s tree tail (level %d of %d, %d fragments)", k == 0 ? "forward" : "reverse",
                      level, L, frags);
            return CBM_ERR_ARG;
        }
    }
    if (d.n_units < 0 || (d.n_units > 0 && (!d.un_ptr || !d.un_row || !d.un_lpar || !d.un_pc || !d.un_pcb ||
                                             !d.un_pce || !d.un_seg_ptr || !d.un_seg || !d.un_of ||
                                             !d.un_waited))) {
        set_error("cbm_create: incomplete unit tables (%d units)", d.n_units);
        return CBM_ERR_ARG;
    }
    RangeSet sets[2];
    {
        RangeSet &a = sets[0], &b = sets[1];
        a.n_ranges = 8;
        a.fwd = d.fwd_pieces;
        a.bwd = d.bwd_pieces;
        a.fpc_b = d.fpc_b, a.fpc_e = d.fpc_e, a.fpc_slot = d.fpc_slot, a.frs_slot = d.frs_slot;
        a.frs_cnt = d.frs_cnt;
        a.bpc_b = d.bpc_b, a.bpc_e = d.bpc_e, a.bpc_mid = d.bpc_mid, a.bpc_slot = d.bpc_slot;
        a.brs_slot = d.brs_slot, a.brs_cnt = d.brs_cnt, a.bpc_row = d.bpc_row;
        b.n_ranges = 4;
        b.fwd = d.fwd_pieces4;
        b.bwd = d.bwd_pieces4;
        b.fpc_b = d.fpc4_b, b.fpc_e = d.fpc4_e, b.fpc_slot = d.fpc4_slot, b.frs_slot = d.frs4_slot;
        b.frs_cnt = d.frs4_cnt;
        b.bpc_b = d.bpc4_b, b.bpc_e = d.bpc4_e, b.bpc_mid = d.bpc4_mid, b.bpc_slot = d.bpc4_slot;
        b.brs_slot = d.brs4_slot, b.brs_cnt = d.brs4_cnt, b.bpc_row = d.bpc4_row;
        const int32_t *firsts[2][2] = {{d.fpc_range_host, d.bpc_range_host}, {d.fpc4_range_host, d.bpc4_range_host}};
        for (int k = 0; k < 2; ++k) {
            RangeSet &r = sets[k];
            if (r.fwd < 0 || r.bwd < 0 ||
                (r.fwd > 0 && (!firsts[k][0] || !r.fpc_b || !r.fpc_e || !r.fpc_slot || !r.frs_slot || !r.frs_cnt)) ||
                (r.bwd > 0 && (!firsts[k][1] || !r.bpc_b || !r.bpc_e || !r.bpc_mid || !r.bpc_slot || !r.brs_slot ||
                               !r.brs_cnt || !r.bpc_row))) {
                set_error("cbm_create: incomplete range piece tables (%d ranges)", r.n_ranges);
                return CBM_ERR_ARG;
            }
            for (int pass = 0; pass < 2; ++pass) {
                const int32_t *first = firsts[k][pass];
                const int32_t total = pass == 0 ? r.fwd : r.bwd;
                if (total <= 0) continue;
                bool ok = first[0] == 0 && first[r.n_ranges] == total;
                for (int q = 0; q < r.n_ranges; ++q) ok = ok && first[q + 1] >= first[q];
                if (!ok) {
                    set_error("cbm_create: range piece offsets do not run from 0 to the piece count (%d ranges)",
                              r.n_ranges);
                    return CBM_ERR_ARG;
                }
                std::copy(first, first + r.n_ranges + 1, pass == 0 ? r.f_first : r.b_first);
            }
        }
    }
    cbm_handle *h = new (std::nothrow) cbm_handle;
    if (!h) {
        set_error("cbm_create: out of host memory");
        return CBM_ERR_NOMEM;
    }
    h->d = d;
    h->rs[0] = sets[0];
    h->rs[1] = sets[1];
    // the host range tables are the caller's: the handle keeps its own copies
    h->d.fpc_range_host = h->d.bpc_range_host = h->d.fpc4_range_host = h->d.bpc4_range_host = nullptr;
    h->lvl_ptr.assign(d.lvl_ptr_host, d.lvl_ptr_host + L + 1);
    h->lvl_nlong.assign(d.lvl_nlong_host, d.lvl_nlong_host + L);
    h->fck_ptr.assign(d.fck_ptr_host, d.fck_ptr_host + L + 1);
    h->fred_ptr.assign(d.fred_ptr_host, d.fred_ptr_host + L + 1);
    h->rev_ptr.assign(d.rev_ptr_host, d.rev_ptr_host + L + 1);
    h->d.rev_ptr_host = h->rev_ptr.data();
    h->d.lvl_ptr_host = h->lvl_ptr.data();
    h->d.lvl_nlong_host = h->lvl_nlong.data();
    h->d.fck_ptr_host = h->fck_ptr.data();
    h->d.fred_ptr_host = h->fred_ptr.data();
    h->rev_nlong.assign(d.rev_nlong_host, d.rev_nlong_host + L);
    h->rck_ptr.assign(d.rck_ptr_host, d.rck_ptr_host + L + 1);
    h->rred_ptr.assign(d.rred_ptr_host, d.rred_ptr_host + L + 1);
    h->d.rev_nlong_host = h->rev_nlong.data();
    h->d.rck_ptr_host = h->rck_ptr.data();
    h->d.rred_ptr_host = h->rred_ptr.data();
    if (d.mode != CBM_MODE_AX) {  // D A D, and D A for the adjoint of A D over a symmetric A
        hipError_t e = hipMalloc(&h->ps_pos, sizeof(float) * size_t(d.n_rows));
        if (e == hipSuccess) e = cbmk::launch_dad_scales(d.n_rows, d.lvl_node, d.lvl_parent, d.d, h->ps_pos, stream);
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
    hipError_t e = hipSuccess;
    for (void *p : {static_cast<void *>(h->ps_pos), static_cast<void *>(h->scratch), static_cast<void *>(h->comp),
                    static_cast<void *>(h->partial), static_cast<void *>(h->fused_state),
                    static_cast<void *>(h->fused_flags), static_cast<void *>(h->tf_order),
                    static_cast<void *>(h->order_meta), static_cast<void *>(h->arrive),
                    static_cast<void *>(h->sc_buf)}) {
        if (!p) continue;
        const hipError_t ei = hipFree(p);
        if (e == hipSuccess) e = ei;
    }
    delete h;
    if (e != hipSuccess) return device_error(e, "cbm_destroy");
    return CBM_OK;
}

extern "C" int cbm_reserve(cbm_handle *h, int32_t n_cols) {
    if (!h || n_cols < 0) {
        set_error("cbm_reserve: invalid arguments");
        return CBM_ERR_ARG;
    }
    int rc = CBM_OK;
    if (!h->tf_order && h->d.tf_level < h->d.n_levels) rc = stage_forward_tail(h);
    if (rc == CBM_OK && !h->sc_on && small_tables_wanted(h)) rc = build_small_chunks(h);
    if (rc != CBM_OK) return rc;
    const int32_t cols = std::max(h->reserved_cols, n_cols);
    rc = grow_buffers(h, cols);
    if (rc == CBM_OK) h->reserved_cols = cols;
    if (rc == CBM_OK && h->plain) rc = cbm_reserve(h->plain, cols);
    return rc;
}

namespace {

// Launches of one forward product over this handle's layout (split form):
// the Delta pass (one grid, two where chunks and short rows stay apart), the
// reduction of multi-chunk rows unless their last chunks finish them, and the
// propagation (per-level launches down
// to the tree tail, then the tail).
int64_t forward_launches(const cbm_handle *h) {
    const int32_t L = h->d.n_levels;
    const int64_t n_multi = h->sc_on ? h->sc_red_lvl[size_t(L)] : h->fred_ptr[size_t(L)];
    const bool finish = h->arrive && h->xcd_slabs <= 1 && h->chunk_finish != 0 &&
                        (h->chunk_finish == 1 || (h->d.n_rows <= kChunkFinishMaxRows && n_multi <= kChunkFinishMaxLong));
    int64_t n = 1 + (n_multi > 0 && !finish ? 1 : 0);
    if (L > 1) {
        const bool tail = tails_on(h) && h->d.tf_level < L;
        n += (tail ? h->d.tf_level - 1 : L - 1) + (tail ? tail_launches(forward_tail(h)) : 0);
    }
    return n;
}

// The cost model that picks between the compressed product and the same
// product over A (cbm_attach_uncompressed): t = launches x kLaunchUs + the
// bytes the gathers and the output move / the gather rate.  A launch costs
// about 8 us on MI355X whatever its size (a 2,708-row row pass and a 2,708-
// row propagation level alike), and random row gathers run at about 7 TB/s
// (the gather ceiling of the Delta pass).  The tree's launches are worth it
// only when the gathers that Delta saves take longer than they do: with
// these constants N = 2,708 (depth 9, nnz 11k -> 7k) and N = 19,717 (depth
// 12, nnz 92k -> 55k) run over A, while N = 132,534 (nnz 80M -> 41M) and
// N = 2,449,029 (nnz 122M -> 117M) run compressed.
// One rate for all gathers is a simplification: gathers whose rows the L2s
// hold run faster (12-18 TB/s on matrices of 8-68 MB), but so do the small
// levels' launches (4-5 us).  A rate that follows the L2 fit alone measured
// worse on graphs of N = 16,000 (depth 31, F = 256 BF16: it chose A, 0.533
// ms per step, where the compressed product takes 0.459) than this constant
// pair, whose one miss among eight measured cases (N = 12,008, F = 512 BF16:
// compressed 0.208 against 0.193 ms) costs 8%.
constexpr double kLaunchUs = 8.0;
constexpr double kGatherBytesPerUs = 7.0e6;

double forward_cost_us(const cbm_handle *h, int32_t n_cols, int elem) {
    const double bytes = double(h->d.nnz) * n_cols * elem + double(h->d.n_rows) * n_cols * 4.0;
    return double(forward_launches(h)) * kLaunchUs + bytes / kGatherBytesPerUs;
}

// The handle a product call of this width and storage type runs on.
cbm_handle *route(cbm_handle *h, int32_t dtype, int32_t n_cols) {
    if (!h->plain || h->plain_mode == 0) return h;
    if (h->plain_mode == 1) return h->plain;
    const int elem = dtype == CBM_BF16 ? 2 : 4;
    return forward_cost_us(h->plain, n_cols, elem) < forward_cost_us(h, n_cols, elem) ? h->plain : h;
}

}  // namespace

extern "C" int cbm_attach_uncompressed(cbm_handle *h, cbm_handle *plain, int32_t mode) {
    if (!h || mode < -1 || mode > 1 || plain == h) {
        set_error("cbm_attach_uncompressed: invalid arguments");
        return CBM_ERR_ARG;
    }
    if (plain && (plain->d.n_rows != h->d.n_rows || plain->d.n_levels != 1 || plain->d.n_edges != 0 ||
                  plain->d.mode != h->d.mode || plain->plain)) {
        set_error("cbm_attach_uncompressed: the uncompressed handle must have %d rows, no tree edges, mode %d and "
                  "no uncompressed handle of its own",
                  h->d.n_rows, h->d.mode);
        return CBM_ERR_ARG;
    }
    h->plain = plain;
    h->plain_mode = mode;
    if (plain) {
        plain->symmetric = h->symmetric;
        return cbm_reserve(plain, h->reserved_cols);
    }
    return CBM_OK;
}

extern "C" int cbm_uncompressed_active(cbm_handle *h, int32_t dtype, int32_t n_cols) {
    if (!h || (dtype != CBM_F32 && dtype != CBM_BF16) || n_cols < 1) {
        set_error("cbm_uncompressed_active: invalid arguments");
        return CBM_ERR_ARG;
    }
    return route(h, dtype, n_cols) != h ? 1 : 0;
}

extern "C" int cbm_attach_order(cbm_handle *h, int32_t n_order, const int32_t *order, int32_t mode,
                                hipStream_t stream) {
    if (!h || n_order < 0 || n_order > h->d.n_rows || mode < -1 || mode > 1 || (n_order > 0 && !order)) {
        set_error("cbm_attach_order: invalid arguments (n_order=%d mode=%d)", n_order, mode);
        return CBM_ERR_ARG;
    }
    if (h->order_meta) {
        const hipError_t e = hipFree(h->order_meta);
        h->order_meta = nullptr;
        if (e != hipSuccess) return device_error(e, "cbm_attach_order");
    }
    h->order = nullptr;
    h->n_order = 0;
    h->order_mode = 0;
    if (n_order == 0) return CBM_OK;
    hipError_t e = hipMalloc(reinterpret_cast<void **>(&h->order_meta), size_t(n_order) * 4 * sizeof(int32_t));
    if (e != hipSuccess) {
        h->order_meta = nullptr;
        return device_error(e, "cbm_attach_order");
    }
    e = cbmk::launch_order_meta(n_order, order, h->d.n_rows, h->d.fwd_ptr, h->d.lvl_node, h->order_meta, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return device_error(e, "cbm_attach_order");
    h->order = order;
    h->n_order = n_order;
    h->order_mode = mode;
    return CBM_OK;
}

extern "C" int cbm_set_order(cbm_handle *h, int32_t mode) {
    if (!h || mode < -1 || mode > 1) {
        set_error("cbm_set_order: invalid arguments");
        return CBM_ERR_ARG;
    }
    h->order_mode = h->n_order > 0 ? mode : 0;
    return CBM_OK;
}

namespace {

// Whether a forward-structure product of n_cols columns in this storage type
// runs as the one-launch kernel (cbm_set_fused).  It needs the plan's units
// (built only for matrices of up to CBM_UNIT_MAX_ROWS rows whose Delta rows
// fit a unit) and is used only when asked for (fused == 1): measured forward
// times (BF16, one launch vs the multi-launch forms, microseconds): N = 2,708
// (3 unit levels) 20.8 vs 18.4; N = 19,717 at alpha = 0 (5 unit levels) 74
// vs 46, at alpha = 8 (no waits) 28.6 vs 26.8; N = 169,343 182 vs 129.  Its
// units run with fewer waves per CU (64 KB of LDS each) and narrower gathers
// than the row passes, and every wait between units costs a device-scope
// release (an L2 write-back) and a poll, which together cost more than the
// launches they save.
bool fused_possible(const cbm_handle *h, int32_t dtype, int32_t n_cols, bool ranges) {
    if (h->fused == 0 || h->d.n_units <= 0 || !h->fused_state) return false;
    if (int64_t(h->d.n_units) * ((n_cols + 63) / 64) > h->fused_flag_count) return false;
    const bool compensated = h->kahan_level >= 0 && std::max<int32_t>(h->kahan_level, 1) < h->d.n_levels;
    if (compensated || h->slab_bytes > 0 || h->xcd_slabs > 1) return false;
    if (ranges && pick_ranges(h, n_cols, elem_of(dtype))) return false;
    return h->fused == 1;
}

// The forward product in buffers b; ranges: whether the XCD column ranges
// may be used (they need b.partial); fused: whether the one-launch kernel may
// be used (it needs the handle's flags).
int forward_product(cbm_handle *h, const Bufs &b, bool ranges, int32_t form, int32_t dtype, const void *x,
                    int64_t ldx, float *y, int64_t ldy, int32_t n_cols, hipStream_t stream, const char *what,
                    bool fused = true) {
    h->last_launches = 0;
    if (fused && fused_possible(h, dtype, n_cols, ranges)) {
        cbmk::FusedArgs a;
        a.un_ptr = h->d.un_ptr;
        a.un_row = h->d.un_row;
        a.un_lpar = h->d.un_lpar;
        a.un_pc = h->d.un_pc;
        a.un_pcb = h->d.un_pcb;
        a.un_pce = h->d.un_pce;
        a.un_seg_ptr = h->d.un_seg_ptr;
        a.un_seg = h->d.un_seg;
        a.un_of = h->d.un_of;
        a.un_waited = h->d.un_waited;
        a.node_of = h->d.lvl_node;
        a.parent_of = h->d.lvl_parent;
        a.parent_pos = h->d.lvl_ppos;
        a.idx = h->d.fwd_col;
        a.colscale = form_colscaled(form) ? h->d.d : nullptr;
        a.rowscale = form_rowscaled(form) ? h->d.d : nullptr;
        a.escale = form_rowscaled(form) ? h->ps_pos : nullptr;
        a.x = x;
        a.ldx = ldx;
        a.y = y;
        a.ldy = ldy;
        a.ncols = n_cols;
        a.n_units = h->d.n_units;
        a.n_slabs = (n_cols + 63) / 64;
        a.ordered = h->d.unit_levels > 1 ? 1 : 0;
        a.state = h->fused_state;
        a.flags = h->fused_flags;
        const hipError_t e = cbmk::launch_fused_forward(a, dtype == CBM_BF16, stream);
        if (e != hipSuccess) return device_error(e, what);
        h->last_launches = 1;
        return CBM_OK;
    }
    const int elem = dtype == CBM_BF16 ? 2 : 4;
    const bool split = use_split(h);
    return for_each_slab(h, n_cols, elem, [&](int32_t c0, int32_t w) {
        const cbmk::RowPass rp = forward_pass(h, form, dtype, shift<char>(x, c0, elem), ldx, y + c0, ldy, w);
        if (const RangeSet *rs = ranges ? pick_ranges(h, n_cols, elem) : nullptr) {
            // Pieces and one reduction over all rows, then the light level
            // propagation: the bits of the per-level reduction with the
            // parent term (forward_ranges with_parent), in fewer, fuller launches.
            const int r = forward_ranges(h, b, *rs, rp, false, stream, what);
            return r == CBM_OK ? propagate_levels(h, b, form, y + c0, ldy, w, stream, what) : r;
        }
        // A one-level product (no tree) with a gather order runs the split
        // form too: it is the same Delta pass, with nothing to propagate.
        const bool ordered = h->d.n_levels == 1 && order_on(h);
        if (!split && !ordered) return forward_levels(h, b, rp, true, stream, what);
        const int r = delta_flat(h, b, rp, stream, what);
        return r == CBM_OK ? propagate_levels(h, b, form, y + c0, ldy, w, stream, what) : r;
    });
}

}  // namespace

extern "C" int cbm_order_active(cbm_handle *h, int32_t dtype, int32_t n_cols) {
    if (!h || (dtype != CBM_F32 && dtype != CBM_BF16) || n_cols < 1) {
        set_error("cbm_order_active: invalid arguments");
        return CBM_ERR_ARG;
    }
    const cbm_handle *r = route(h, dtype, n_cols);
    if (!order_on(r) || fused_possible(r, dtype, n_cols, true) || pick_ranges(r, n_cols, elem_of(dtype))) return 0;
    return use_split(r) || r->d.n_levels == 1 ? 1 : 0;
}

extern "C" int cbm_matmul(cbm_handle *h, int32_t dtype, const void *x, int64_t ldx, float *y, int64_t ldy,
                          int32_t n_cols, hipStream_t stream) {
    int rc = check_dense(h, dtype, x, ldx, y, ldy, n_cols, "cbm_matmul");
    if (rc == CBM_OK) rc = check_reserved(h, n_cols, "cbm_matmul");
    if (rc != CBM_OK) return rc;
    if (cbm_handle *p = route(h, dtype, n_cols); p != h) {
        rc = cbm_matmul(p, dtype, x, ldx, y, ldy, n_cols, stream);
        h->last_launches = p->last_launches;
        return rc;
    }
    return forward_product(h, handle_bufs(h), true, h->d.mode, dtype, x, ldx, y, ldy, n_cols, stream, "cbm_matmul");
}

namespace {

int check_adjoint(cbm_handle *h, int32_t dtype, const void *g, int64_t ldg, float *dx, int64_t lddx, float *z,
                  int64_t ldz, int32_t n_cols, const char *what) {
    int rc = check_dense(h, dtype, g, ldg, dx, lddx, n_cols, what);
    if (rc != CBM_OK) return rc;
    if (!z || ldz < n_cols) {
        set_error("%s: invalid workspace (ldz=%lld)", what, (long long)ldz);
        return CBM_ERR_ARG;
    }
    const int32_t n = h->d.n_rows;
    if (overlaps(z, span_bytes(n, ldz, n_cols, 4), dx, span_bytes(n, lddx, n_cols, 4)) ||
        overlaps(z, span_bytes(n, ldz, n_cols, 4), g, span_bytes(n, ldg, n_cols, dtype == CBM_BF16 ? 2 : 4))) {
        set_error("%s: workspace overlaps an operand", what);
        return CBM_ERR_ARG;
    }
    return check_reserved(h, n_cols, what);
}

// Reverse pass over the nodes with children, deepest level first:
//   z_u = s_u g_u + sum_{children v} (v has children ? z_v : s_v g_v)
// with s = D^-1/2 for D A D (the adjoint's (I-P)^-T applied to D G) and
// 1 otherwise.  A leaf's z would be s g, so it is never written: both this
// pass and the Delta^T pass read s g straight from G for leaves.
int adjoint_reverse(cbm_handle *h, int32_t dtype, const void *g, int64_t ldg, float *z, int64_t ldz, int32_t n_cols,
                    hipStream_t stream) {
    const float *gs = h->d.mode == CBM_MODE_DADX ? h->d.d : nullptr;
    cbmk::RowPass rv;
    rv.node_of = h->d.rev_node;
    rv.ptr = h->d.ch_ptr;
    rv.idx = h->d.ch_node;
    rv.dual = true;
    rv.mid = h->d.ch_mid;
    rv.g = z;
    rv.ldg = ldz;
    rv.g2 = g;
    rv.ldg2 = ldg;
    rv.g2_bf16 = dtype == CBM_BF16;
    rv.g2scale = gs;
    rv.e = g;
    rv.lde = ldg;
    rv.e_bf16 = dtype == CBM_BF16;
    rv.escale = gs;
    rv.extra = cbmk::kExtraSelf;
    rv.y = z;
    rv.ldy = ldz;
    rv.ncols = n_cols;
    rv.rows_hint = h->d.n_rows;
    rv.wave_rows = h->wave_rows;
    rv.mixed_grid = h->mixed_grid;
    rv.xcd_slabs = h->xcd_slabs;
    cbmk::ChunkArgs ck;
    ck.pos = h->d.rck_pos;
    ck.meta = h->d.rck_meta;
    ck.slot = h->d.rck_slot;
    ck.scratch = h->scratch;
    cbmk::ReduceArgs rd;
    rd.pos = h->d.rred_pos;
    rd.slot = h->d.rred_slot;
    rd.cnt = h->d.rred_cnt;
    rd.scratch = h->scratch;
    int32_t first_level = h->d.n_levels;  // levels from here down run as the tail
    if (tails_on(h) && h->d.tr_level < h->d.n_levels) {
        cbmk::TailArgs t;
        t.frag = h->d.tr_frag;
        t.seg = h->d.tr_seg;
        t.pos = h->d.tr_pos;
        t.n_frag = h->d.tr_frags;
        const hipError_t err = cbmk::launch_row_tail(rv, t, stream);
        if (err != hipSuccess) return device_error(err, "adjoint reverse pass (tail)");
        ++h->last_launches;
        first_level = h->d.tr_level;
    }
    for (int32_t l = first_level - 1; l >= 0; --l) {
        const int32_t b = h->rev_ptr[size_t(l)], e = h->rev_ptr[size_t(l) + 1];
        if (e <= b) continue;
        // A level's launch holds only its own nodes: a thin level gives each
        // node a whole wavefront (its child list in sub-group slices).
        rv.rows_hint = e - b;
        ck.first = h->rck_ptr[size_t(l)];
        ck.count = h->rck_ptr[size_t(l) + 1] - ck.first;
        rd.first = h->rred_ptr[size_t(l)];
        rd.count = h->rred_ptr[size_t(l) + 1] - rd.first;
        const int rc = run_range(h, rv, b, e, h->rev_nlong[size_t(l)], ck, rd, stream, "adjoint reverse pass");
        if (rc != CBM_OK) return rc;
    }
    return CBM_OK;
}

// Delta^T pass: dx_j = r_j * sum_v sign * (z_v or s_v g_v) over column j of
// Delta, r = D^-1/2 for A D and D A D.
int adjoint_delta_t(cbm_handle *h, int32_t dtype, const void *g, int64_t ldg, const float *z, int64_t ldz, float *dx,
                    int64_t lddx, int32_t n_cols, hipStream_t stream) {
    cbmk::RowPass tp;
    tp.node_of = h->d.bwd_node;
    tp.ptr = h->d.bwd_ptr;
    tp.idx = h->d.bwd_row;
    tp.dual = true;
    tp.mid = h->d.bwd_mid;
    tp.g = z;
    tp.ldg = ldz;
    tp.g2 = g;
    tp.ldg2 = ldg;
    tp.g2_bf16 = dtype == CBM_BF16;
    tp.g2scale = h->d.mode == CBM_MODE_DADX ? h->d.d : nullptr;
    tp.y = dx;
    tp.ldy = lddx;
    tp.ncols = n_cols;
    tp.rows_hint = h->d.n_rows;
    tp.wave_rows = h->wave_rows;
    tp.mixed_grid = h->mixed_grid;
    tp.xcd_slabs = h->xcd_slabs;
    tp.rowscale = h->d.mode == CBM_MODE_AX ? nullptr : h->d.d;
    cbmk::ChunkArgs ck;
    ck.pos = h->d.bck_pos;
    ck.meta = h->d.bck_meta;
    ck.slot = h->d.bck_slot;
    ck.first = 0;
    ck.count = h->d.n_bck;
    ck.scratch = h->scratch;
    cbmk::ReduceArgs rd;
    rd.pos = h->d.bred_pos;
    rd.slot = h->d.bred_slot;
    rd.cnt = h->d.bred_cnt;
    rd.first = 0;
    rd.count = h->d.n_bred;
    rd.scratch = h->scratch;
    if (const RangeSet *rs = pick_ranges(h, n_cols, elem_of(dtype))) {
        // By XCD column range: every backward piece (one range of G and Z,
        // its leaf rows before its internal rows) into its partial row, then
        // one reduction over all positions.
        cbmk::RowPass pc = tp;
        pc.node_of = rs->bpc_slot;
        pc.idx = rs->bpc_row;
        pc.ptr = rs->bpc_b;
        pc.ptr_end = rs->bpc_e;
        pc.mid = rs->bpc_mid;
        pc.rowscale = nullptr;
        pc.y = h->partial;
        pc.ldy = n_cols;
        set_ranges(pc, rs->b_first, rs->n_ranges, range_phases(h, *rs, n_cols, elem_of(dtype)));
        hipError_t e = cbmk::launch_row_pass(pc, stream);
        if (e != hipSuccess) return device_error(e, "adjoint Delta^T pass");
        ++h->last_launches;
        cbmk::RowPass rr = tp;
        rr.dual = false;
        rr.g = h->partial;
        rr.ldg = n_cols;
        rr.g_bf16 = false;
        rr.g2 = nullptr;
        cbmk::ReduceArgs pr;
        pr.pos = nullptr;
        pr.slot = rs->brs_slot;
        pr.cnt = rs->brs_cnt;
        pr.first = 0;
        pr.count = h->d.n_rows;
        pr.scratch = h->partial;
        e = cbmk::launch_row_reduce(rr, pr, stream);
        if (e != hipSuccess) return device_error(e, "adjoint Delta^T pass");
        ++h->last_launches;
        return CBM_OK;
    }
    return run_range(h, tp, 0, h->d.n_rows, h->d.bwd_nlong, ck, rd, stream, "adjoint Delta^T pass");
}

}  // namespace

extern "C" int cbm_matmul_t(cbm_handle *h, int32_t dtype, const void *g, int64_t ldg, float *dx, int64_t lddx,
                            float *z, int64_t ldz, int32_t n_cols, hipStream_t stream) {
    if (h && n_cols >= 1 && (dtype == CBM_F32 || dtype == CBM_BF16) && n_cols <= h->reserved_cols) {
        if (cbm_handle *p = route(h, dtype, n_cols); p != h) {
            const int rc = cbm_matmul_t(p, dtype, g, ldg, dx, lddx, z, ldz, n_cols, stream);
            h->last_launches = p->last_launches;
            return rc;
        }
    }
    if (h && h->symmetric) {
        // A = A^T: M^T = M for A and D A D, and (A D)^T = D A, all through
        // the forward structure (no Delta^T pass, no reverse gather, z unused).
        int rc = check_dense(h, dtype, g, ldg, dx, lddx, n_cols, "cbm_matmul_t");
        if (rc == CBM_OK) rc = check_reserved(h, n_cols, "cbm_matmul_t");
        if (rc != CBM_OK) return rc;
        const int32_t form = h->d.mode == CBM_MODE_ADX ? kFormDAX : h->d.mode;
        return forward_product(h, handle_bufs(h), true, form, dtype, g, ldg, dx, lddx, n_cols, stream,
                               "cbm_matmul_t");
    }
    int rc = check_adjoint(h, dtype, g, ldg, dx, lddx, z, ldz, n_cols, "cbm_matmul_t");
    if (rc != CBM_OK) return rc;
    h->last_launches = 0;
    const int elem = dtype == CBM_BF16 ? 2 : 4;
    return for_each_slab(h, n_cols, elem, [&](int32_t c0, int32_t w) {
        const void *gs = shift<char>(g, c0, elem);
        int r = adjoint_reverse(h, dtype, gs, ldg, z + c0, ldz, w, stream);
        if (r == CBM_OK) r = adjoint_delta_t(h, dtype, gs, ldg, z + c0, ldz, dx + c0, lddx, w, stream);
        return r;
    });
}

extern "C" int cbm_adjoint_reverse(cbm_handle *h, int32_t dtype, const void *g, int64_t ldg, float *z, int64_t ldz,
                                   int32_t n_cols, hipStream_t stream) {
    // This half reads g and writes z only.
    int rc = check_dense(h, dtype, g, ldg, z, ldz, n_cols, "cbm_adjoint_reverse");
    if (rc == CBM_OK) rc = check_reserved(h, n_cols, "cbm_adjoint_reverse");
    if (rc != CBM_OK) return rc;
    h->last_launches = 0;
    return adjoint_reverse(h, dtype, g, ldg, z, ldz, n_cols, stream);
}

extern "C" int cbm_adjoint_delta_t(cbm_handle *h, int32_t dtype, const void *g, int64_t ldg, const float *z,
                                   int64_t ldz, float *dx, int64_t lddx, int32_t n_cols, hipStream_t stream) {
    int rc = check_adjoint(h, dtype, g, ldg, dx, lddx, const_cast<float *>(z), ldz, n_cols, "cbm_adjoint_delta_t");
    if (rc != CBM_OK) return rc;
    h->last_launches = 0;
    return adjoint_delta_t(h, dtype, g, ldg, z, ldz, dx, lddx, n_cols, stream);
}

extern "C" int cbm_delta_spmm(cbm_handle *h, int32_t dtype, const void *x, int64_t ldx, float *y, int64_t ldy,
                              int32_t n_cols, hipStream_t stream) {
    int rc = check_dense(h, dtype, x, ldx, y, ldy, n_cols, "cbm_delta_spmm");
    if (rc == CBM_OK) rc = check_reserved(h, n_cols, "cbm_delta_spmm");
    if (rc != CBM_OK) return rc;
    h->last_launches = 0;
    const int elem = dtype == CBM_BF16 ? 2 : 4;
    return for_each_slab(h, n_cols, elem, [&](int32_t c0, int32_t w) {
        const cbmk::RowPass rp = forward_pass(h, h->d.mode, dtype, shift<char>(x, c0, elem), ldx, y + c0, ldy, w);
        if (const RangeSet *rs = pick_ranges(h, n_cols, elem))
            return forward_ranges(h, handle_bufs(h), *rs, rp, false, stream, "cbm_delta_spmm");
        return delta_flat(h, handle_bufs(h), rp, stream, "cbm_delta_spmm");
    });
}

extern "C" int cbm_propagate(cbm_handle *h, float *y, int64_t ldy, int32_t n_cols, hipStream_t stream) {
    if (!h || !y || n_cols < 1 || ldy < n_cols) {
        set_error("cbm_propagate: invalid arguments");
        return CBM_ERR_ARG;
    }
    const int rc = check_reserved(h, n_cols, "cbm_propagate");
    if (rc != CBM_OK) return rc;
    h->last_launches = 0;
    return propagate_levels(h, handle_bufs(h), h->d.mode, y, ldy, n_cols, stream, "cbm_propagate");
}

extern "C" int cbm_matmul_noalloc(cbm_handle *h, int32_t dtype, const void *x, int64_t ldx, float *y, int64_t ldy,
                                  int32_t n_cols, hipStream_t stream) {
    int rc = check_dense(h, dtype, x, ldx, y, ldy, n_cols, "cbm_matmul_noalloc");
    if (rc != CBM_OK) return rc;
    if (cbm_handle *p = route(h, dtype, n_cols); p != h) return cbm_matmul_noalloc(p, dtype, x, ldx, y, ldy, n_cols, stream);
    float *base = nullptr;  // this device's copy of `scratch`
    rc = noalloc_scratch_address(&base);
    if (rc != CBM_OK) return rc;
    // The long-row partial sums of the forward form in use, then the
    // compensation rows, each start 256-byte aligned.
    const int64_t slots = h->sc_on ? h->sc_slots : use_split(h) ? h->d.fwd_gslots : h->d.fwd_slots;
    const int64_t part = (slots * n_cols + 63) / 64 * 64;
    const int64_t comp = int64_t(h->d.n_rows - comp_base(h)) * n_cols;
    const int64_t need = (part + comp) * int64_t(sizeof(float));
    if (need > int64_t(kNoallocScratchBytes)) {
        set_error("cbm_matmul_noalloc: %d columns need %lld bytes of scratch, the static scratch holds %lld",
                  n_cols, (long long)need, (long long)kNoallocScratchBytes);
        return CBM_ERR_ARG;
    }
    Bufs b;
    b.scratch = base;
    b.comp = base + part;
    return forward_product(h, b, false, h->d.mode, dtype, x, ldx, y, ldy, n_cols, stream, "cbm_matmul_noalloc",
                           false);
}

extern "C" int cbm_last_launches(const cbm_handle *h) { return h ? h->last_launches : -1; }

extern "C" int cbm_set_kahan_depth(cbm_handle *h, int32_t level) {
    if (!h) {
        set_error("cbm_set_kahan_depth: null handle");
        return CBM_ERR_ARG;
    }
    h->kahan_level = level < 0 ? -1 : level;
    return grow_buffers(h, h->reserved_cols);  // the compensated levels may need more rows
}

extern "C" int cbm_set_forward_split(cbm_handle *h, int32_t mode) {
    if (!h || mode < -1 || mode > 1) {
        set_error("cbm_set_forward_split: invalid arguments");
        return CBM_ERR_ARG;
    }
    h->fwd_split = mode;
    return CBM_OK;
}

extern "C" int cbm_set_wave_rows(cbm_handle *h, int32_t mode) {
    if (!h || mode < -1 || mode > 2) {
        set_error("cbm_set_wave_rows: invalid arguments");
        return CBM_ERR_ARG;
    }
    h->wave_rows = mode;
    return CBM_OK;
}

extern "C" int cbm_set_mixed_grid(cbm_handle *h, int32_t mode) {
    if (!h || mode < -1 || mode > 1) {
        set_error("cbm_set_mixed_grid: invalid arguments");
        return CBM_ERR_ARG;
    }
    h->mixed_grid = mode;
    return CBM_OK;
}

extern "C" int cbm_set_chunk_finish(cbm_handle *h, int32_t mode) {
    if (!h || mode < -1 || mode > 1) {
        set_error("cbm_set_chunk_finish: invalid arguments");
        return CBM_ERR_ARG;
    }
    h->chunk_finish = mode;
    return CBM_OK;
}

extern "C" int cbm_set_xcd_slabs(cbm_handle *h, int32_t slabs) {
    if (!h || slabs < 1 || slabs > 64) {
        set_error("cbm_set_xcd_slabs: invalid arguments");
        return CBM_ERR_ARG;
    }
    h->xcd_slabs = slabs;
    return CBM_OK;
}

extern "C" int cbm_set_xcd_ranges(cbm_handle *h, int32_t mode) {
    if (!h || !(mode == -1 || mode == 0 || mode == 1 || mode == 4 || mode == 8)) {
        set_error("cbm_set_xcd_ranges: invalid arguments");
        return CBM_ERR_ARG;
    }
    h->xcd_ranges = mode;
    return grow_buffers(h, h->reserved_cols);  // a piece set needs its partial rows
}

extern "C" int cbm_set_fused(cbm_handle *h, int32_t mode) {
    if (!h || mode < -1 || mode > 1) {
        set_error("cbm_set_fused: invalid arguments");
        return CBM_ERR_ARG;
    }
    h->fused = mode;
    return CBM_OK;
}

extern "C" int cbm_fused_active(const cbm_handle *h, int32_t dtype, int32_t n_cols) {
    if (!h || (dtype != CBM_F32 && dtype != CBM_BF16) || n_cols < 1) {
        set_error("cbm_fused_active: invalid arguments");
        return CBM_ERR_ARG;
    }
    return fused_possible(h, dtype, n_cols, true) ? 1 : 0;
}

extern "C" int64_t cbm_fused_errors(cbm_handle *h) {
    if (!h) {
        set_error("cbm_fused_errors: null handle");
        return CBM_ERR_ARG;
    }
    if (!h->fused_state) return 0;
    uint32_t words[4] = {};
    const hipError_t e = hipMemcpy(words, h->fused_state, sizeof words, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return device_error(e, "cbm_fused_errors");
    return int64_t(words[3]);
}

extern "C" int cbm_set_symmetric(cbm_handle *h, int32_t symmetric) {
    if (!h || symmetric < 0 || symmetric > 1) {
        set_error("cbm_set_symmetric: invalid arguments");
        return CBM_ERR_ARG;
    }
    h->symmetric = symmetric;
    if (h->plain) h->plain->symmetric = symmetric;
    return CBM_OK;
}

extern "C" int cbm_set_tree_tails(cbm_handle *h, int32_t mode) {
    if (!h || mode < 0 || mode > 3) {
        set_error("cbm_set_tree_tails: invalid arguments");
        return CBM_ERR_ARG;
    }
    h->tree_tails = mode;
    return CBM_OK;
}

extern "C" int cbm_set_xcd_phases(cbm_handle *h, int32_t phases) {
    if (!h || !(phases == -1 || phases == 1 || phases == 2 || phases == 4)) {
        set_error("cbm_set_xcd_phases: invalid arguments");
        return CBM_ERR_ARG;
    }
    h->xcd_phases = phases;
    return CBM_OK;
}

extern "C" int cbm_xcd_ranges_active(const cbm_handle *h, int32_t dtype, int32_t n_cols) {
    if (!h || (dtype != CBM_F32 && dtype != CBM_BF16) || n_cols < 1) {
        set_error("cbm_xcd_ranges_active: invalid arguments");
        return CBM_ERR_ARG;
    }
    const RangeSet *rs = pick_ranges(h, n_cols, elem_of(dtype));
    return rs ? rs->n_ranges : 0;
}

extern "C" int cbm_set_slab_bytes(cbm_handle *h, int64_t bytes) {
    if (!h || bytes < 0) {
        set_error("cbm_set_slab_bytes: invalid arguments");
        return CBM_ERR_ARG;
    }
    h->slab_bytes = bytes;
    return CBM_OK;
}
