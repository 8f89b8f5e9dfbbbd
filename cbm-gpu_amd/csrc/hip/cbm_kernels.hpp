// Kernel launchers of the CBM product (gfx950).  See cbm_kernels.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../../include/cbm_host.h"

namespace cbmk {

// Where the optional per-row extra term of a row pass comes from.
enum Extra : int32_t {
    kExtraNone = 0,    // y_row = rowscale * sum
    kExtraParent = 1,  // y_row = rowscale * sum + escale[p] * e[parent_of[p]]   (forward level pass)
    kExtraSelf = 2,    // y_row = sum + escale[row] * e[row]                      (adjoint reverse pass)
    kExtraParentKahan = 3  // as kExtraParent, carrying a compensation row per position
                           // (comp) so that deep levels add no rounding drift
};

// One row pass: for every position p in [p_begin, p_end), with row = node_of[p]
// (or p when node_of is null),
//   sum = sum_k  sign_k * colscale[j_k] * val_k * g[j_k, :]   over k in [ptr[p], ptr[p+1]),
// where idx[k] = j_k | sign_k << 31; y[row, :] = rowscale[row] * sum + the
// extra term per `extra`, stored in FP32.
struct RowPass {
    int32_t p_begin = 0, p_end = 0;
    const int32_t *node_of = nullptr;
    const int32_t *parent_of = nullptr;
    const int32_t *ptr = nullptr;
    const int32_t *ptr_end = nullptr; // row ends (4-array CSR); null = ptr[p + 1]
    const int32_t *idx = nullptr;
    const float *val = nullptr;       // per-entry values; null = +-1 from the sign bit
    const float *colscale = nullptr;  // indexed by gathered row; null = 1
    const float *rowscale = nullptr;  // indexed by output row;   null = 1
    const float *escale = nullptr;    // by position (parent) or by row (self); null = 1
    const void *g = nullptr;          // gathered matrix
    int64_t ldg = 0;
    bool g_bf16 = false;
    // Dual pass: the entries [ptr[p], mid[p]) read g2 (scaled by
    // g2scale[row]), the entries [mid[p], end) read g.
    bool dual = false;
    const int32_t *mid = nullptr;
    const void *g2 = nullptr;
    int64_t ldg2 = 0;
    bool g2_bf16 = false;
    const float *g2scale = nullptr;
    const void *e = nullptr;          // extra-term matrix
    int64_t lde = 0;
    bool e_bf16 = false;
    Extra extra = kExtraNone;
    float ecoef = 1.0f;               // extra term multiplier, on top of escale
    // kExtraParentKahan: comp row of position p is comp[(p - comp_base) * ldc];
    // parent_pos[p] is the parent's position (its comp row if >= comp_base).
    const int32_t *parent_pos = nullptr;
    float *comp = nullptr;
    int32_t comp_base = 0;
    int64_t ldc = 0;
    float *y = nullptr;
    int64_t ldy = 0;
    int32_t ncols = 0;
    // The short-row pass skips positions with more entries than this (rows
    // that the chunk kernel covers when one pass spans several levels).
    int32_t skip_above = INT32_MAX;
    // Rows of the whole product (0: unknown).  When rows x LPR lanes cannot
    // fill the device's wave slots, the short-row pass gives each row a whole
    // wavefront: 64 / LPR sub-groups sum slices of its entries, which are
    // then added across sub-groups.  This trades occupancy that would idle
    // anyway for fewer dependent gather rounds per row.  Where more rows than
    // wave slots would take several rounds of wavefronts, rows narrower than
    // 16 lanes get 16 lanes (16 / LPR sub-groups, four rows per wavefront)
    // instead.  wave_rows: -1 by that rule, 0 never, 1 always a wavefront,
    // 2 always 16 lanes.
    int32_t rows_hint = 0;
    int32_t wave_rows = -1;
    // Long-row chunks and short rows of one range (launch_row_mixed): one
    // grid (1), two launches (0), or by the rule there (-1).
    int32_t mixed_grid = -1;
    // XCD feature slabs: the columns are cut into xcd_slabs slabs of slab_w
    // (set by the launcher, at least 16 vectors wide) and workgroup b works
    // on slab b % xcd_slabs.  Workgroups are dispatched round-robin over the
    // XCDs, so each XCD gathers only its own slab of the gathered matrix and
    // its L2 holds more of what it re-reads.  Every row is still summed in
    // one place, in the same order: the same bits as without slabs.
    int32_t xcd_slabs = 1;
    int32_t slab_w = 0;
    // XCD column ranges: when set, the positions are pieces grouped by the
    // column range they gather from, range k of range_count (2, 4 or 8)
    // holding [range_first[k], range_first[k + 1]).  With S = 8 / range_count
    // XCDs per range and range_phases phases, the columns are cut into
    // xcd_slabs = S * range_phases slabs; workgroup b works on range
    // (b % 8) / S, row block (b / 8) % range_blocks, slab
    // ((b % 8) % S) * range_phases + (b / 8) / range_blocks, so each XCD
    // gathers one range and, phase after phase (workgroups are dispatched in
    // order), one slab of it at a time.  p_begin and p_end are ignored;
    // range_blocks is set by the launcher.
    bool ranges = false;
    int32_t range_count = 8;
    int32_t range_phases = 1;
    int32_t range_blocks = 0;
    int32_t range_first[9] = {};
    // Row order of the short-row pass: when set, row slot i of [p_begin,
    // p_end) is position order[i] (-1: none), and the pass's order_blocks
    // workgroups (set by the launcher) take the row blocks XCD by XCD:
    // workgroup b (dispatched to XCD b % 8) takes block b / 8 of XCD b % 8's
    // contiguous eighth, so each XCD walks one stretch of the order.
    const int32_t *order = nullptr;
    int32_t order_blocks = 0;
    // With an order, optionally the rows' (position, first entry, end, output
    // row) in order: int32 [4 x rows], 16-byte aligned, one load per row
    // instead of the chain order -> ptr -> node_of.  A position below 0 is
    // skipped.  Only for passes without an extra term or a dual source.
    const int32_t *order_meta = nullptr;
};

// order_meta[i] = (order[i], ptr[order[i]], ptr[order[i] + 1], node_of[order[i]])
// for i < n (positions outside [0, n_rows) give (-1, 0, 0, 0)).
hipError_t launch_order_meta(int32_t n, const int32_t *order, int32_t n_rows, const int32_t *ptr,
                             const int32_t *node_of, int32_t *meta, hipStream_t stream);

// Wave slots of the device (256 CUs x 4 SIMDs x 8 waves): a pass with fewer
// rows x lanes than 64 x this leaves slots free.
constexpr int64_t kWaveSlots = 8192;

// Launches one row pass over positions [p_begin, p_end); returns the launch status.
hipError_t launch_row_pass(const RowPass &rp, hipStream_t stream);

// Chunks [first, first + count) of the long rows of a row pass: chunk i covers
// part meta & 0xffff of meta >> 16 equal parts of row pos[i]'s entries; its sum
// finishes the row (slot[i] < 0) or goes to scratch row slot[i] (ld = ncols).
// With arrive (one zeroed counter per scratch slot, xcd_slabs 1 only), the
// last chunk of a multi-chunk row to finish adds the row's scratch rows in
// slot order and finishes the row, as launch_row_reduce would (same bits),
// and sets the counter back to zero; no reduction launch follows.
struct ChunkArgs {
    const int32_t *pos = nullptr;
    const int32_t *meta = nullptr;
    const int32_t *slot = nullptr;
    int32_t first = 0, count = 0;
    float *scratch = nullptr;
    uint32_t *arrive = nullptr;
};
hipError_t launch_row_chunks(const RowPass &rp, const ChunkArgs &c, hipStream_t stream);

// The chunks of c and the short-row pass over [p_begin, p_end) of rp: one
// grid, except for a grid made mostly of chunks that would run fewer waves
// per CU than the chunk kernel (then launch_row_chunks, launch_row_pass; the
// same work and bits either way); rp.mixed_grid 1 or 0 forces one form.  *launched gets the
// number of kernels launched.  The range must not be an XCD range pass.
hipError_t launch_row_mixed(const RowPass &rp, const ChunkArgs &c, int *launched, hipStream_t stream);

// Rows [first, first + count) of a reduction list: row pos[i] (i itself when
// pos is null) is the sum of scratch rows slot[i] .. slot[i] + cnt[i] - 1,
// added in order (0 when cnt[i] is 0), then finished.
struct ReduceArgs {
    const int32_t *pos = nullptr;
    const int32_t *slot = nullptr;
    const int32_t *cnt = nullptr;
    int32_t first = 0, count = 0;
    const float *scratch = nullptr;
};
hipError_t launch_row_reduce(const RowPass &rp, const ReduceArgs &r, hipStream_t stream);

// In-place propagation of one level: y[node_of[p]] += parscale[p] * y[parent_of[p]].
hipError_t launch_propagate(int32_t p_begin, int32_t p_end, const int32_t *node_of, const int32_t *parent_of,
                            const float *parscale, float *y, int64_t ldy, int32_t ncols, hipStream_t stream);

// Compensation rows of a compensated propagation: position p >= base owns
// comp[(p - base) * ldc, ...]; parent_pos[p] is the parent's position.
struct CompArgs {
    const int32_t *parent_pos = nullptr;
    float *comp = nullptr;
    int32_t base = 0;
    int64_t ldc = 0;
};

// The same level propagation with the parent's compensation row added and
// this row's rounding error stored (the arithmetic of kExtraParentKahan, so a
// Delta pass followed by this gives the fused pass's bits).
hipError_t launch_propagate_comp(int32_t p_begin, int32_t p_end, const int32_t *node_of, const int32_t *parent_of,
                                 const float *parscale, const CompArgs &c, float *y, int64_t ldy, int32_t ncols,
                                 hipStream_t stream);

// A tree tail (plan arrays CBM_P_TF_* / CBM_P_TR_*): fragment f holds the
// segments [frag[f], frag[f+1]), one per tree level from the fragment's root
// down, segment s the positions pos[seg[s] .. seg[s+1]).  A tail launch gives
// each fragment one 256-thread workgroup, which finishes the segments one
// after another with a workgroup barrier between them.
struct TailArgs {
    const int32_t *frag = nullptr;
    const int32_t *seg = nullptr;
    const int32_t *pos = nullptr;
    int32_t n_frag = 0;
    // Optional order of the fragments: order[0 .. n_staged) are finished with
    // their rows staged in LDS (at most staged_rows rows each, kTailSlab
    // columns per workgroup), order[n_staged .. n_frag) as above.  Without an
    // order, workgroup f takes fragment f.
    const int32_t *order = nullptr;
    int32_t n_staged = 0;
    int32_t staged_rows = 0;
};

// Columns of one workgroup of a staged tail, and the LDS such a workgroup may
// hold: a fragment is staged only if its rows fit (tail_staged_bytes).
constexpr int kTailSlab = 16;
constexpr int64_t kTailStagedMaxBytes = 65536;
inline int64_t tail_staged_bytes(int32_t rows, bool comp) {
    // position, parent slot, parent position and scale, then the row (and
    // its compensation row)
    return int64_t(rows) * (4 * int64_t(sizeof(int32_t)) + int64_t(kTailSlab) * int64_t(sizeof(float)) * (comp ? 2 : 1));
}

// The forward propagation of a tail, root segments first: each position as in
// launch_propagate, or as in launch_propagate_comp from position comp_first
// on (the same arithmetic, so the same bits as the per-level launches).  The
// staged fragments (t.order) run first, in a launch of their own.
hipError_t launch_propagate_tail(const TailArgs &t, const int32_t *node_of, const int32_t *parent_of,
                                 const float *parscale, const CompArgs &c, int32_t comp_first, float *y, int64_t ldy,
                                 int32_t ncols, hipStream_t stream);

// A row pass over a tail, deepest segments first (the adjoint's reverse pass:
// a node's internal children sit one segment deeper in its own fragment).
// Each row is summed by one group of lanes, in entry order.
hipError_t launch_row_tail(const RowPass &rp, const TailArgs &t, hipStream_t stream);

// The forward product in one launch over the plan's units (CBM_P_UN_*):
// Delta pieces, propagation within each unit in LDS, outside parents after
// their unit's flag; see fused_forward.hip.  The scalings are those of a
// RowPass (colscale by gathered row, rowscale by output row, escale by
// position).  state: 4 uint32 in device memory, zero before the first call
// (ticket, done, epoch, failed waits); flags: n_units x n_slabs uint32.
struct FusedArgs {
    const int32_t *un_ptr = nullptr, *un_row = nullptr, *un_lpar = nullptr, *un_pc = nullptr;
    const int32_t *un_pcb = nullptr, *un_pce = nullptr, *un_seg_ptr = nullptr, *un_seg = nullptr;
    const int32_t *un_of = nullptr, *un_waited = nullptr;
    const int32_t *node_of = nullptr, *parent_of = nullptr, *parent_pos = nullptr;
    const int32_t *idx = nullptr;  // FWD_COL (packed signs)
    const float *colscale = nullptr, *rowscale = nullptr, *escale = nullptr;
    const void *x = nullptr;
    int64_t ldx = 0;
    float *y = nullptr;
    int64_t ldy = 0;
    int32_t ncols = 0;
    int32_t n_units = 0;
    int32_t n_slabs = 0;  // ceil(ncols / 64)
    // Whether units wait on one another (more than one component level):
    // only then are units taken by ticket and the state words used; without
    // waits workgroup b takes unit b / n_slabs and no atomic is issued.
    int32_t ordered = 1;
    uint32_t *state = nullptr;
    uint32_t *flags = nullptr;
    uint32_t spin_limit = 1u << 20;
};
hipError_t launch_fused_forward(const FusedArgs &a, bool x_bf16, hipStream_t stream);

// The reference's branch-serial propagation (cbm/cbm.cu:25-91) for the
// operator surface: one workgroup per branch and column block walks the
// branch's edges in order, y[dst] += y[src] (mult null) or
// y[dst] += (y[src] * mult[dst]) / mult[src].
hipError_t launch_branch_update(int32_t n_branches, const int32_t *branches, const int32_t *src, const int32_t *dst,
                                const float *mult, float *y, int64_t ldy, int32_t ncols, hipStream_t stream);

// y[r, :] += coef * x[r, :] for r < n_rows (FP32).
hipError_t launch_axpy(int32_t n_rows, int32_t ncols, float coef, const float *x, int64_t ldx, float *y, int64_t ldy,
                       hipStream_t stream);

// Per-position propagation factor of the D A D product:
// ps[p] = d[node[p]] / d[parent[p]] (0 for roots and zero-degree parents).
hipError_t launch_dad_scales(int32_t n, const int32_t *node_of, const int32_t *parent_of, const float *d, float *ps,
                             hipStream_t stream);

}  // namespace cbmk
