/*
 * cbm_hip.h -- C ABI of the CBM product on AMD Instinct MI355X (gfx950),
 * libcbm_hip.so.
 *
 * A handle binds one CBM matrix, laid out by cbm_plan_create (cbm_host.h)
 * and copied to device memory by the caller, to the HIP kernels:
 *
 *     forward   Y  = M X         M = A, A D, or D A D   (D = D^-1/2, diagonal)
 *     adjoint   dX = M^T G
 *
 * with A = (I - P)^-1 Delta.  X and G are row-major N x F in FP32 or BF16;
 * Y, dX and all accumulation are FP32.  Every call is asynchronous on the
 * given stream.  Product calls allocate nothing: their scratch is sized by
 * cbm_reserve, and a call wider than the reserved width fails with
 * CBM_ERR_ARG, so every product call may be captured in a HIP graph.
 * Calls return CBM_OK (0) or a negative status; cbm_last_error() describes
 * the failure for the calling thread.
 *
 * Replaces class cusparseCBM (cbm/cbm.cu:93-300): its constructor
 * (cbm/cbm.cu:95-142) becomes cbm_create, matmul_AX_ADX (:150-207) and
 * matmul_DAX_DADX (:210-268) become cbm_matmul, and the adjoint is new.
 */
#ifndef CBM_HIP_H
#define CBM_HIP_H

#include <stdint.h>

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes (shared with cbm_host.h). */
#ifndef CBM_OK
#define CBM_OK 0
#define CBM_ERR_ARG -1
#define CBM_ERR_VALUE -2
#define CBM_ERR_NOMEM -3
#define CBM_ERR_INTERNAL -4
#endif
#define CBM_ERR_DEVICE -5 /* a HIP or rocSPARSE call failed */

/* Which product a handle computes. */
enum cbm_mode {
    CBM_MODE_AX = 0,   /* M = A            (cbm4mm,  cbm/cbm4mm.py)   */
    CBM_MODE_ADX = 1,  /* M = A D^-1/2     (cbm4ad,  cbm/cbm4ad.py)   */
    CBM_MODE_DADX = 2  /* M = D^-1/2 A D^-1/2 (cbm4dad, cbm/cbm4dad.py) */
};

/* Storage type of X and G. */
enum cbm_dtype { CBM_F32 = 0, CBM_BF16 = 1 };

/*
 * Device layout of one CBM matrix.  All pointers are borrowed and must stay
 * valid until cbm_destroy.  Arrays are those of cbm_plan_create
 * (include/cbm_host.h); the per-level tables marked "host" are host copies,
 * everything else lives in device memory.
 * d is the device vector D^-1/2 [N] (0 for nodes of degree 0) for ADX/DADX,
 * NULL for AX.
 */
typedef struct cbm_desc {
    int32_t n_rows;
    int32_t n_levels;
    int64_t nnz;
    int64_t n_edges;
    int32_t mode;
    /* host: level offsets [L+1], long rows per level [L], chunk offsets [L+1],
       reduction-row offsets [L+1], internal-node offsets [L+1] */
    const int32_t *lvl_ptr_host;
    const int32_t *lvl_nlong_host;
    const int32_t *fck_ptr_host;
    const int32_t *fred_ptr_host;
    const int32_t *rev_ptr_host;
    /* device: level order, forward Delta; internal nodes and their children */
    const int32_t *lvl_node;
    const int32_t *lvl_parent;
    const int32_t *lvl_ppos;
    const int32_t *fwd_ptr;
    const int32_t *fwd_col;
    const int32_t *rev_node;
    const int32_t *ch_ptr;
    const int32_t *ch_node;
    const int32_t *ch_mid;
    /* device: forward long-row chunks and reductions */
    const int32_t *fck_pos;
    const int32_t *fck_meta;
    const int32_t *fck_slot;
    const int32_t *fred_pos;
    const int32_t *fred_slot;
    const int32_t *fred_cnt;
    int32_t fwd_slots;
    /* device: Delta^T in backward order, its long-row chunks and reductions */
    const int32_t *bwd_node;
    const int32_t *bwd_ptr;
    const int32_t *bwd_row;
    const int32_t *bwd_mid;
    const int32_t *bck_pos;
    const int32_t *bck_meta;
    const int32_t *bck_slot;
    const int32_t *bred_pos;
    const int32_t *bred_slot;
    const int32_t *bred_cnt;
    int32_t bwd_nlong;
    int32_t n_bck;
    int32_t n_bred;
    int32_t bwd_slots;
    const float *d;
    /* device: forward chunk slots numbered over all levels (plan arrays
       FCK_GSLOT, FRED_GSLOT) and their count (plan stat FWD_GSLOTS) */
    const int32_t *fck_gslot;
    const int32_t *fred_gslot;
    int32_t fwd_gslots;
    /* the adjoint's reverse pass: host per-level tables (long child lists per
       level [L], chunk offsets [L+1], reduction-row offsets [L+1]), device
       chunk and reduction tables, and the slots its largest level needs */
    const int32_t *rev_nlong_host;
    const int32_t *rck_ptr_host;
    const int32_t *rred_ptr_host;
    const int32_t *rck_pos;
    const int32_t *rck_meta;
    const int32_t *rck_slot;
    const int32_t *rred_pos;
    const int32_t *rred_slot;
    const int32_t *rred_cnt;
    int32_t rev_slots;
    /* column-range pieces (plan arrays CBM_P_FPC_* .. CBM_P_BRS_*, stats
       FWD_PIECES / BWD_PIECES; counts 0 and NULL pointers when the plan has
       none): piece offsets per range on the host [CBM_XCD_RANGES + 1], the
       piece tables and per-row slot tables on the device */
    const int32_t *fpc_range_host;
    const int32_t *fpc_b;
    const int32_t *fpc_e;
    const int32_t *fpc_slot;
    const int32_t *frs_slot;
    const int32_t *frs_cnt;
    int32_t fwd_pieces;
    const int32_t *bpc_range_host;
    const int32_t *bpc_b;
    const int32_t *bpc_e;
    const int32_t *bpc_mid;
    const int32_t *bpc_slot;
    const int32_t *brs_slot;
    const int32_t *brs_cnt;
    int32_t bwd_pieces;
    const int32_t *bpc_row;
    /* the same for the 4-range set (plan arrays CBM_P_FPC4_* .. CBM_P_BPC4_ROW,
       stats FWD_PIECES4 / BWD_PIECES4; offsets [5]): each of its ranges is
       gathered on two XCDs, one column half each */
    const int32_t *fpc4_range_host;
    const int32_t *fpc4_b;
    const int32_t *fpc4_e;
    const int32_t *fpc4_slot;
    const int32_t *frs4_slot;
    const int32_t *frs4_cnt;
    int32_t fwd_pieces4;
    const int32_t *bpc4_range_host;
    const int32_t *bpc4_b;
    const int32_t *bpc4_e;
    const int32_t *bpc4_mid;
    const int32_t *bpc4_slot;
    const int32_t *brs4_slot;
    const int32_t *brs4_cnt;
    int32_t bwd_pieces4;
    const int32_t *bpc4_row;
    /* tree tails (plan arrays CBM_P_TF_* / CBM_P_TR_*, stats TF_LEVEL /
       TR_LEVEL): the first level of each tail (n_levels: no tail), the number
       of fragments (the fragment offsets hold one more entry), and the device
       tables */
    int32_t tf_level;
    int32_t tf_frags;
    const int32_t *tf_frag;
    const int32_t *tf_seg;
    const int32_t *tf_pos;
    int32_t tr_level;
    int32_t tr_frags;
    const int32_t *tr_frag;
    const int32_t *tr_seg;
    const int32_t *tr_pos;
    /* units of the one-launch forward (plan arrays CBM_P_UN_*, stats UNITS
       and UNIT_LEVELS; n_units 0 and NULL pointers when the plan has none) */
    int32_t n_units;
    int32_t unit_levels;
    const int32_t *un_ptr;
    const int32_t *un_row;
    const int32_t *un_lpar;
    const int32_t *un_pc;
    const int32_t *un_pcb;
    const int32_t *un_pce;
    const int32_t *un_seg_ptr;
    const int32_t *un_seg;
    const int32_t *un_of;
    const int32_t *un_waited;
} cbm_desc;

typedef struct cbm_handle cbm_handle;

/* Message describing the last failure on the calling thread ("" if none). */
const char *cbm_last_error(void);

/* sizeof(cbm_desc) as compiled into the library (a binding's layout check). */
int32_t cbm_desc_size(void);

/*
 * Creates a handle.  Derived per-position scale vectors (N floats each) are
 * allocated here, once, and computed on `stream`; the call synchronises the
 * stream before returning.
 */
int cbm_create(const cbm_desc *desc, hipStream_t stream, cbm_handle **out);
int cbm_destroy(cbm_handle *h);

/*
 * Sizes the handle's scratch for products of up to n_cols columns: the
 * partial sums of long rows, max(fwd_slots, bwd_slots, fwd_gslots, rev_slots) x n_cols
 * FP32, the compensation rows of the compensated levels, and the partial
 * rows of the XCD column ranges.  Reserving never shrinks: the handle serves
 * the widest width reserved so far.  The product calls (cbm_matmul,
 * cbm_matmul_t, cbm_delta_spmm, cbm_propagate, cbm_adjoint_*) never allocate
 * and fail with CBM_ERR_ARG for more columns than reserved.  The settings
 * that change what a width needs (cbm_set_kahan_depth, cbm_set_xcd_ranges)
 * grow the buffers for the reserved width themselves.  The first reservation
 * also orders the forward tail's fragments for staging in LDS (it reads the
 * tail tables back, waiting for the device; see cbm_set_tree_tails), builds
 * the forward Delta pass's long-row tables of a small or sparse matrix (rows
 * of more than 64 entries in chunks of about 256, for up to 65,536 rows, or
 * up to 262,144 with at most 32 entries per row; from the plan's row offsets,
 * read back the same way), and allocates one arrival counter per scratch
 * slot (cbm_set_chunk_finish).  The scratch and the counters serve one call
 * at a time: product calls on one handle that overlap on the device (other
 * streams, other threads) race.  The reference allocates a cuSPARSE
 * workspace on every call instead (cbm/cbm.cu:181,203).
 */
int cbm_reserve(cbm_handle *h, int32_t n_cols);

/*
 * Forward product Y = M X.  x: N x n_cols, leading dimension ldx (elements);
 * y: N x n_cols FP32, leading dimension ldy; x and y must not overlap.
 * Two forms with the same result bits (cbm_set_forward_split): split (the
 * default when the tree has more than one level) sums every Delta row in one
 * pass over all levels, then adds parents level by level in place; fused
 * sums each level's Delta rows and adds the parent's finished row in the same
 * pass.  Long rows are summed by whole workgroups, in chunks.
 * Replaces cusparseCBM::matmul_AX_ADX / matmul_DAX_DADX.
 */
int cbm_matmul(cbm_handle *h, int32_t dtype, const void *x, int64_t ldx, float *y, int64_t ldy, int32_t n_cols,
               hipStream_t stream);

/*
 * Forward product Y = M X that allocates nothing and needs no cbm_reserve.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP-graph capture, serving loops, harnesses that
 * forbid hipMalloc).  Same arguments and result as cbm_matmul, except that
 * the XCD column ranges are not used (their partial rows would not fit), so
 * where cbm_matmul would take them the bits can differ within FP32
 * rounding.  Its long-row partial sums and compensation rows live in a
 * static device array of 64 MiB per GPU, held until the process ends; a call
 * that needs more fails with CBM_ERR_ARG.  Every call on a GPU shares that
 * array: calls that overlap on one GPU (other streams, other threads, other
 * handles) race and must be serialised by the caller.
 */
int cbm_matmul_noalloc(cbm_handle *h, int32_t dtype, const void *x, int64_t ldx, float *y, int64_t ldy,
                       int32_t n_cols, hipStream_t stream);

/*
 * Adjoint dX = M^T G.  z is caller-provided FP32 workspace N x n_cols with
 * leading dimension ldz; it receives (I-P)^-T applied to the scaled G for
 * the nodes that have children (for a leaf that row is just the scaled G,
 * which the kernels read from g directly; z's leaf rows are left untouched).
 * On a handle declared symmetric (cbm_set_symmetric) the adjoint runs
 * through the forward structure instead -- M^T = M for A and D A D, and
 * (A D)^T = D A, i.e. the forward passes with Delta's rows scaled by D and
 * the propagation factor D_v / D_parent -- and z may be NULL (it is unused).
 * No reference counterpart on the GPU; the CPU op s_spmm_csr_t_int32
 * (cbm/cbm_extensions.cpp:742-782) is the closest.
 */
int cbm_matmul_t(cbm_handle *h, int32_t dtype, const void *g, int64_t ldg, float *dx, int64_t lddx, float *z,
                 int64_t ldz, int32_t n_cols, hipStream_t stream);

/*
 * The two halves of the adjoint, for separate timing: cbm_adjoint_reverse
 * writes z (the reverse tree gather), cbm_adjoint_delta_t then writes
 * dx = M^T g from z and g.  Together they equal cbm_matmul_t.
 */
int cbm_adjoint_reverse(cbm_handle *h, int32_t dtype, const void *g, int64_t ldg, float *z, int64_t ldz,
                        int32_t n_cols, hipStream_t stream);
int cbm_adjoint_delta_t(cbm_handle *h, int32_t dtype, const void *g, int64_t ldg, const float *z, int64_t ldz,
                        float *dx, int64_t lddx, int32_t n_cols, hipStream_t stream);

/*
 * The two halves of the forward product, for separate timing and for the
 * operator surface:  cbm_delta_spmm writes Y = Delta' X (Delta' = Delta,
 * Delta D or D Delta D by mode) without propagation; cbm_propagate then adds
 * parents in level order in place, Y_v += s_v Y_parent(v) with s_v = 1 or
 * D_v / D_parent (cbm/cbm.cu:25-91 `update` / `fused_update`).
 */
int cbm_delta_spmm(cbm_handle *h, int32_t dtype, const void *x, int64_t ldx, float *y, int64_t ldy, int32_t n_cols,
                   hipStream_t stream);
int cbm_propagate(cbm_handle *h, float *y, int64_t ldy, int32_t n_cols, hipStream_t stream);
/* cbm_propagate uses the compensated form from the Kahan level on (see
   cbm_set_kahan_depth), so cbm_delta_spmm + cbm_propagate = cbm_matmul. */

/* Per-kernel launch counts of the last call on this handle (for profiling). */
int cbm_last_launches(const cbm_handle *h);

/*
 * Compensated propagation: the forward product's levels from `level` on
 * (default 32; negative switches it off) carry an FP32 compensation term per
 * row (TwoSum / TwoProduct on the parent add), so rounding error no longer
 * grows with tree depth there.  It costs one extra row read and write per
 * node at those levels and a buffer of (nodes at those levels) x n_cols FP32.
 */
int cbm_set_kahan_depth(cbm_handle *h, int32_t level);

/*
 * Column slabs: cbm_matmul, cbm_matmul_t and cbm_delta_spmm process the
 * columns in slabs whose gathered part (N x width x element size) is at most
 * `bytes`, so that the random row gathers of one slab keep hitting the
 * 256 MB last-level cache.  0 (the default) processes whole rows, which
 * measured faster at every slab size on MI355X: partial-row gathers cost
 * more than the cache residency returns.
 */
int cbm_set_slab_bytes(cbm_handle *h, int64_t bytes);

/*
 * Form of cbm_matmul: 1 split (Delta pass over all levels, then per-level
 * propagation), 0 fused per level, -1 automatic (split when the tree has
 * more than one level).  Both give identical results.
 */
int cbm_set_forward_split(cbm_handle *h, int32_t mode);

/*
 * Row mapping of the short-row passes: 1 gives every row a whole wavefront
 * (its entries split into slices summed by sub-groups of lanes, then added
 * across them), 2 gives rows narrower than 16 lanes 16 lanes (four rows per
 * wavefront, in slices the same way), 0 packs 64 / lanes-per-row rows into a
 * wavefront, -1 (the default) chooses the first when N rows cannot fill the
 * device's wave slots packed and fit them in one round as wavefronts, the
 * second when they cannot fill them packed but need several rounds as
 * wavefronts (small matrices, where per-row gather latency, not bandwidth,
 * sets the time).  The mappings add a row's entries in different orders;
 * each is deterministic.
 */
int cbm_set_wave_rows(cbm_handle *h, int32_t mode);

/*
 * Long-row chunks and short rows of a gather pass: 1 runs them as one grid
 * (the chunks' workgroups first), 0 as two launches (chunks, then short
 * rows), -1 (the default) one grid except where a grid made mostly of chunks,
 * in a matrix that fills the device, would run fewer waves per CU than the
 * chunk kernel alone.  Both forms give the same bits.
 */
int cbm_set_mixed_grid(cbm_handle *h, int32_t mode);

/*
 * Long rows of more than one chunk: 1 has the last chunk of each such row to
 * finish (an arrival counter per scratch slot, allocated by cbm_reserve) add
 * the row's chunk sums and finish it inside the gather pass, 0 adds them in
 * a reduction launch after the pass, -1 (the default) the first for passes
 * with at most 64 such rows, over matrices of up to 262,144 rows (the
 * last chunk's device-scope acquire invalidates its XCD's L2, which costs a
 * pass with many such rows its cache hits).  Column slabs and XCD ranges
 * always use the reduction launch.  Both forms add the chunk sums in chunk
 * order: the same bits.  No counterpart in
 * the reference (its Delta SpMM is one cusparseSpMM call, cbm/cbm.cu:184-189).
 */
int cbm_set_chunk_finish(cbm_handle *h, int32_t mode);

/*
 * XCD feature slabs (1..64, default 1 = off): the columns of the gathered
 * matrix are cut into up to `slabs` slabs and workgroup b of a gather pass
 * works on slab b % slabs.  MI355X dispatches workgroups round-robin over
 * its 8 XCDs, so with 2 slabs each XCD gathers half of every row and its
 * 4 MB L2 covers twice as many of the rows it re-reads.  A slab is never
 * narrower than 16 lanes of 16-byte vectors (256 bytes of the narrowest
 * gathered row), so fewer slabs than asked are used on narrow rows (none
 * below 512 bytes).  Every row is still summed in one place, in the same
 * order: results are identical to slabs = 1 under the packed row mapping.
 */
int cbm_set_xcd_slabs(cbm_handle *h, int32_t slabs);

/*
 * XCD column ranges (-1 automatic, the default; 0 off; 1 on when the plan has
 * range pieces, with the piece set chosen by row width; 8 or 4 on with that
 * set).  The gathered matrix's rows are cut into ranges and each XCD
 * (workgroup b runs on XCD b mod 8) gathers from one range only, so its 4 MB
 * L2 serves an eighth of the matrix: with 8 ranges XCD k gathers range k
 * whole; with 4 ranges XCDs 2k and 2k + 1 gather range k, one column half
 * each.  The forward Delta pass and the adjoint Delta^T pass then write one
 * partial row per piece (a row's entries in one range, at most CBM_PIECE of
 * them) and a reduction adds each row's pieces in order; the forward's level
 * propagation follows as usual.  Four ranges halve those partial rows but
 * halve each gather too, so they are chosen when half a gathered row is
 * 256 bytes or more (F x element size >= 512 bytes) and the plan has them.
 * Results are deterministic but rounded in another order than with ranges
 * off (and differently for the two sets).  Automatic: on when the plan has
 * pieces, Delta holds 64 entries per row or more on average, and the
 * gathered matrix (N x F entries of the call's storage type: 2 bytes for
 * BF16, 4 for FP32) is between 24 and 256 MiB, so a BF16 and an FP32 call of
 * one width may choose differently.  Needs F x 4 bytes of scratch per piece;
 * cbm_reserve sizes it for the current choice of either storage type.
 */
int cbm_set_xcd_ranges(cbm_handle *h, int32_t mode);

/*
 * Declares that the matrix A of the handle is symmetric (1) or not (0, the
 * default).  With A = A^T, cbm_matmul_t computes M^T G through the forward
 * structure: the forward's Delta pass and propagation over G (for A D, with
 * the scalings of D A), which gathers only G rows -- no FP32 intermediate
 * rows as the Delta^T pass does -- and needs no workspace.  Rounding differs
 * from the Delta^T form (another summation order); both are deterministic.
 * The cbm4* classes set it from the structure build, which compares A with
 * A^T exactly; a wrong declaration gives a wrong adjoint.
 */
int cbm_set_symmetric(cbm_handle *h, int32_t symmetric);

/*
 * The forward product in one launch (-1 automatic, the default; 0 off; 1 on
 * where possible).  With the plan's units (CBM_P_UN_*), one kernel does the
 * Delta pass and the propagation of every level: a workgroup takes a unit of
 * at most CBM_UNIT_ROWS rows (connected parts of the tree) and a 64-column
 * slab, gathers its rows' Delta entries into LDS, propagates level by level
 * in LDS, and reads a parent outside the unit from the output once that
 * parent's unit has set its completion flag; units are dispatched so that a
 * unit only waits on units already running.  Possible when the plan has
 * units, no level is compensated (cbm_set_kahan_depth), no XCD ranges, slabs
 * or column slabs are in use, and the width is reserved (the flags are sized
 * by cbm_reserve).  Used by cbm_matmul and, on a symmetric handle, by
 * cbm_matmul_t.  Each row sums its entries in pieces of CBM_UNIT_PIECE, so
 * the rounding differs from the multi-launch forms; it is deterministic.
 * Automatic (-1) leaves it off: on MI355X the multi-launch forms measured
 * faster at every size tried (a unit's 64 KB of LDS halves the waves per CU,
 * and each wait between units costs a device-scope release and a poll).
 */
int cbm_set_fused(cbm_handle *h, int32_t mode);

/* Whether a dtype product of n_cols columns would run in one launch now
   (1) or not (0); CBM_ERR_ARG for invalid arguments. */
int cbm_fused_active(const cbm_handle *h, int32_t dtype, int32_t n_cols);

/*
 * The number of dependency waits of the one-launch forward that gave up
 * after their poll limit since the handle was created (0 when every wait
 * ended normally; a nonzero count means some results read a parent row
 * before it was final).  Synchronises the device.
 */
int64_t cbm_fused_errors(cbm_handle *h);

/*
 * Attaches `plain`, a handle of the same product over A itself (a layout
 * with no tree edges, so Delta = A; created by the caller and borrowed until
 * both are destroyed), and chooses when the product calls run over it
 * instead: mode 1 always, 0 never, -1 (automatic) when a cost model --
 * launches per product x 8 us plus gathered and written bytes / 7 TB/s --
 * predicts it faster.  On small graphs the tree's dependent launches cost
 * more than the gathers Delta saves (N = 2,708 and 19,717 run over A), on
 * large ones the reverse.  cbm_matmul, cbm_matmul_t and cbm_matmul_noalloc
 * follow the choice; cbm_delta_spmm, cbm_propagate and the adjoint halves
 * always run the compressed form.  cbm_reserve reserves both handles, and
 * cbm_set_symmetric applies to both.  plain = NULL detaches.  No reference
 * counterpart: the reference always runs the compressed product.
 */
int cbm_attach_uncompressed(cbm_handle *h, cbm_handle *plain, int32_t mode);

/* Whether a dtype product of n_cols columns runs over the attached
   uncompressed handle now (1) or compressed (0); CBM_ERR_ARG for invalid
   arguments. */
int cbm_uncompressed_active(cbm_handle *h, int32_t dtype, int32_t n_cols);

/*
 * Gather order of the Delta pass: order [n_order], device, borrowed (it must
 * outlive the handle's use of it), the short rows of this handle's forward
 * row list (plan array CBM_P_FWD_PTR) in the order of cbm_gather_order
 * (cbm_host.h), every short row exactly once.  The Delta pass over all levels
 * at once (the split forward, cbm_delta_spmm, and the adjoint through the
 * forward structure) then sums the short rows in that order, each XCD
 * taking one contiguous stretch of it, so that rows sharing columns run
 * together on one L2; each row is summed as before, so the results are the
 * same bits.  Not used with XCD column ranges, column slabs or the
 * one-launch forward.  mode: 1 on, 0 off, -1 automatic (the same as on).
 * n_order = 0 detaches.  The handle keeps, per ordered row, its position,
 * entry range and output row in device memory it allocates here (16 bytes a
 * row, one load per row in the pass), computed on `stream`; the call
 * synchronises the stream.  No counterpart in the reference, whose Delta
 * SpMM is one cusparseSpMM call (cbm/cbm.cu:184-189).
 */
int cbm_attach_order(cbm_handle *h, int32_t n_order, const int32_t *order, int32_t mode, hipStream_t stream);
int cbm_set_order(cbm_handle *h, int32_t mode);

/* Whether a forward product of n_cols columns in dtype storage runs its
   Delta pass in the attached gather order (1) or not (0); CBM_ERR_ARG for
   invalid arguments. */
int cbm_order_active(cbm_handle *h, int32_t dtype, int32_t n_cols);

/*
 * Tree tails (1 on, the default; 0 off; 2 and 3 on with other staging).  The
 * deepest tree levels of the forward propagation and of the adjoint's reverse
 * pass, each of which would be a short dependent launch, are finished in one
 * launch: one workgroup per fragment (a node of the tail's first level with
 * its descendants) works through the fragment's levels with a workgroup
 * barrier between them.  The plan chooses each tail's first level by a cost
 * model (cbm_plan_create).  In the forward tail, fragments of at least 8
 * levels whose rows fit 64 KiB of LDS in 16-column slabs run first, in a
 * launch of their own, with their rows staged in LDS (mode 1), when the
 * deepest of them has 24 levels or more; mode 2 stages none, mode 3 every
 * fragment that fits.  The forward gives the same bits in
 * every mode; the reverse pass sums each row in entry order by one group of
 * lanes, which can round differently from the per-level mapping.
 */
int cbm_set_tree_tails(cbm_handle *h, int32_t mode);

/*
 * Column phases of the XCD range passes (-1 automatic, the default; 1, 2 or
 * 4).  With P phases each XCD gathers its range one column slab (1/P of its
 * columns) after the other -- the grid holds every row block of a phase
 * before the next phase's -- so its L2 serves a P times smaller working set,
 * at the price of P times narrower gathers and P reads of every index.
 * Automatic: 2 when an XCD's share of the gathered matrix (N x F x element
 * size / 8) is 12 MiB or more and a phase's gather stays 256 bytes or more,
 * else 1.  Same sums in the same order as 1 phase: the bits do not change.
 */
int cbm_set_xcd_phases(cbm_handle *h, int32_t phases);

/*
 * The number of column ranges (8 or 4) a dtype call of n_cols columns runs
 * its gather passes with under the handle's current setting, 0 when it runs
 * without, CBM_ERR_ARG for invalid arguments.  Reports the choice described
 * above; computes nothing.
 */
int cbm_xcd_ranges_active(const cbm_handle *h, int32_t dtype, int32_t n_cols);

/* ------------------------------------------------------------------------- */
/* Candidate edges of the structure build on the GPU.                         */
/* Replace the MKL syrk of A A^T and the two candidate scans of cbm_init_     */
/* (cbm/cbm_extensions.cpp:100-182).  The result feeds                        */
/* cbm_build_from_candidates (cbm_host.h), which runs the arborescence and    */
/* builds Delta on the host.                                                  */
/* ------------------------------------------------------------------------- */

/*
 * How the overlap counters are held (see cand_api.hip):
 *   CBM_CAND_TILES  in LDS, column tile by column tile; one 1024-thread
 *                   workgroup per CU; cost per row ~ sum of deg(k) + (n - r),
 *                   for matrices up to a few 100k rows
 *   CBM_CAND_LISTS  n global counters and an n-entry touched list per
 *                   workgroup in the workspace; cost per row ~ the columns it
 *                   touches, for any n
 *   CBM_CAND_AUTO   tiles up to 262144 rows, lists above
 */
enum { CBM_CAND_AUTO = 0, CBM_CAND_TILES = 1, CBM_CAND_LISTS = 2 };

/*
 * Bytes of device workspace a call with these arguments needs: 256 bytes,
 * plus for the lists 2 n_rows uint32 words per workgroup (n_workgroups <= 0:
 * the library's choice, about 8 GiB of slabs at most).  -1 for invalid
 * arguments.
 */
int64_t cbm_candidates_workspace_bytes(int32_t n_rows, int32_t path, int32_t n_workgroups);

/*
 * Pass 1.  For the binary CSR (ptr int64 [n_rows + 1], col int32 [nnz],
 * columns sorted and unique per row, all in [0, n_rows)) and the CSR of its
 * transpose (t_ptr, t_col; the same arrays when A is symmetric) -- what
 * cbm_adjacency_create gives -- row_count[r] (int64 [n_rows], device) receives
 * the number of candidate edges created for row r: the root edge N -> r, plus
 * for every c > r sharing a column with r, c -> r when h < deg(r) - alpha and
 * h != deg(r), and r -> c when h < deg(c) - alpha and h != deg(c), where
 * h = |A_r xor A_c|.  path and n_workgroups as for the workspace, which is
 * device memory, 256-byte aligned, of at least cbm_candidates_workspace_bytes
 * bytes.  Asynchronous on `stream`; allocates nothing.
 */
int cbm_candidates_count(int32_t n_rows, const int64_t *ptr, const int32_t *col, const int64_t *t_ptr,
                         const int32_t *t_col, int32_t alpha, int32_t path, int32_t n_workgroups, int64_t *row_count,
                         void *workspace, int64_t workspace_bytes, hipStream_t stream);

/*
 * Pass 2.  Writes the edges of row r to src / dst / w (int32, device) from
 * position row_offset[r] on (the exclusive prefix sum of pass 1's counts), in
 * the reference creation order: the root edge (src = n_rows, w = deg(r)),
 * then for c ascending, c -> r before r -> c (w = h).  n_cand is the length of
 * the three arrays; no write goes past it.  The concatenation over r is the
 * host scan's list (cbm_candidates_create) entry for entry, whatever the path.
 */
int cbm_candidates_emit(int32_t n_rows, const int64_t *ptr, const int32_t *col, const int64_t *t_ptr,
                        const int32_t *t_col, int32_t alpha, int32_t path, int32_t n_workgroups,
                        const int64_t *row_offset, int32_t *src, int32_t *dst, int32_t *w, int64_t n_cand,
                        void *workspace, int64_t workspace_bytes, hipStream_t stream);

/* Message describing the last cbm_candidates_* failure on the calling thread. */
const char *cbm_candidates_last_error(void);

/*
 * Partner scan of the gather order (cbm_gather_order in cbm_host.h; no
 * counterpart in the reference).  For every row r of an ordered row list
 * (ptr int32 [n_rows + 1], col int32 with bit 31 a sign, columns in
 * [0, n_cols)) and, per column, the rows holding it (t_ptr int32
 * [n_cols + 1], t_col int32): partner[k r .. k r + k) receive the k (1 to
 * 4) other rows sharing the most columns with r (most first, ties to the
 * lower row; -1 for none) and overlap[k r ..] the shared counts, counted
 * only through columns held by at most `cap` rows.  Rows whose entries are
 * not wanted (long rows) are simply left out of t_col.  The workspace is
 * device memory, 256-byte aligned, of at least cbm_partners_workspace_bytes
 * bytes (as the list path of the candidate scan).  Asynchronous on
 * `stream`; allocates nothing.
 */
int64_t cbm_partners_workspace_bytes(int32_t n_rows, int32_t n_workgroups);
int cbm_partners(int32_t n_rows, const int32_t *ptr, const int32_t *col, int32_t n_cols, const int32_t *t_ptr,
                 const int32_t *t_col, int32_t cap, int32_t k, int32_t n_workgroups, int32_t *partner,
                 int32_t *overlap, void *workspace, int64_t workspace_bytes, hipStream_t stream);

/* ------------------------------------------------------------------------- */
/* Row clustering: the data passes of the fingerprint clustering of           */
/* benchmark/clustering_gpu.py:86-128 (the recursion, grouping and merging    */
/* run on the host, cbm/clustering.py).                                       */
/* ------------------------------------------------------------------------- */

/*
 * counts[c] (int32 [n_cols], device) = number of the n_sel selected rows
 * rows[] that hold column c in the binary CSR (ptr int64 [n_rows + 1], col
 * int32): the in-degrees within a cluster (clustering_gpu.py:99-100).
 * Rows or columns out of range are skipped.
 */
int cbm_cluster_column_counts(int32_t n_rows, int32_t n_cols, const int64_t *ptr, const int32_t *col,
                              const int32_t *rows, int32_t n_sel, int32_t *counts, hipStream_t stream);

/*
 * fp[i] (int32 [n_sel], device) = sum over j < n_hash of 2^j if row rows[i]
 * holds column hash_cols[j] (clustering_gpu.py:106-110).  1 <= n_hash <= 30.
 */
int cbm_cluster_fingerprints(int32_t n_rows, const int64_t *ptr, const int32_t *col, const int32_t *rows,
                             int32_t n_sel, const int32_t *hash_cols, int32_t n_hash, int32_t *fp,
                             hipStream_t stream);

/* Message describing the last cbm_cluster_* failure on the calling thread. */
const char *cbm_cluster_last_error(void);

/* ------------------------------------------------------------------------- */
/* CSR comparator on the same GPU: rocSPARSE SpMM, FP32.                      */
/* Replaces class cusparseCSR (cbm/csr.cu:24-126).                            */
/* ------------------------------------------------------------------------- */

typedef struct csr_handle csr_handle;

/* Borrowed device CSR (int32 indices, FP32 values) of an n_rows x n_cols matrix. */
int csr_create(int32_t n_rows, int32_t n_cols, int64_t nnz, const int32_t *crow, const int32_t *col,
               const float *val, csr_handle **out);
int csr_destroy(csr_handle *h);

/*
 * Y = A X (trans = 0) or Y = A^T X (trans = 1) with rocsparse_spmm.  The
 * rocSPARSE work buffer is allocated on the first call for a given shape and
 * kept by the handle; later calls of the same or smaller size allocate nothing.
 */
int csr_matmul(csr_handle *h, int32_t trans, const float *x, int64_t ldx, float *y, int64_t ldy, int32_t n_cols,
               hipStream_t stream);

/* Selects the rocsparse_spmm_alg used by csr_matmul (default: rocsparse_spmm_alg_default). */
int csr_set_algorithm(csr_handle *h, int32_t alg);

/* Message describing the last csr_* failure on the calling thread. */
const char *csr_last_error(void);

/* ------------------------------------------------------------------------- */
/* Operator surface: the s_*_csr_int32 ops of cbm/cbm_extensions.cpp:351-907  */
/* on device arrays.  A general CSR is given as MKL-style row begin / end     */
/* arrays (row_e may be NULL: row_b[r + 1] ends row r) with FP32 values.      */
/* ------------------------------------------------------------------------- */

/*
 * dst = A rhs, or dst = A rhs + self_coef * rhs when add_self != 0.
 * Backs s_spmm_csr_int32 (:700-740) and s_spmm_add_csr_int32 (:446-492,
 * self_coef = 1 + epsilon).
 */
int cbm_op_spmm(int32_t n_rows, const int32_t *row_b, const int32_t *row_e, const int32_t *col, const float *val,
                const float *rhs, int64_t ldr, float *dst, int64_t ldd, int32_t n_cols, float self_coef,
                int32_t add_self, hipStream_t stream);

/*
 * Branch-serial propagation in place over y: for each branch b, for each
 * edge e in [branches[b], branches[b+1]) in order, y[dst[e]] += y[src[e]]
 * (mult == NULL) or y[dst[e]] += (y[src[e]] * mult[dst[e]]) / mult[src[e]].
 * Backs s_update_csr_int32 (:784-838) and s_fused_update_csr_int32 (:841-907),
 * and the propagation halves of the s_spmm_*update* ops.
 */
int cbm_op_update(int32_t n_branches, const int32_t *branches, const int32_t *src, const int32_t *dst,
                  const float *mult, float *y, int64_t ldy, int32_t n_cols, hipStream_t stream);

/*
 * dst = A^T rhs for the n_rows x n_cols matrix A given by rows (row_b /
 * row_e as above, nnz = sum of the row lengths, column indices in
 * [0, n_cols)); rhs is n_rows x n_feat, dst n_cols x n_feat.  Backs
 * s_spmm_csr_t_int32 (:742-782).  A^T is formed on the device in the
 * caller's workspace (256-byte aligned, cbm_op_spmm_t_workspace_bytes
 * bytes) by a stable radix sort of the entries by column, so every column
 * lists its rows in ascending order and the result is deterministic; the
 * row-pass kernel then sums each column.  Allocates nothing.
 */
int64_t cbm_op_spmm_t_workspace_bytes(int32_t n_rows, int32_t n_cols, int64_t nnz);
int cbm_op_spmm_t(int32_t n_rows, int32_t n_cols, const int32_t *row_b, const int32_t *row_e, const int32_t *col,
                  const float *val, int64_t nnz, const float *rhs, int64_t ldr, float *dst, int64_t ldd,
                  int32_t n_feat, void *workspace, int64_t workspace_bytes, hipStream_t stream);

/* y += coef * x over n_rows x n_cols (the cblas_saxpy of :491,590). */
int cbm_op_axpy(int32_t n_rows, int32_t n_cols, float coef, const float *x, int64_t ldx, float *y, int64_t ldy,
                hipStream_t stream);

/* Message describing the last cbm_op_* failure on the calling thread. */
const char *cbm_op_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* CBM_HIP_H */
