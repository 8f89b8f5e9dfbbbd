/*
 * cbm_host.h -- C ABI of the host-side CBM structure builder (libcbm_host.so).
 *
 * The CBM (compressed binary matrix) form of a binary adjacency matrix A is
 *     A = (I - P)^-1 * Delta
 * where P[v, parent(v)] = 1 for every tree edge of a minimum spanning
 * arborescence over row Hamming distances, and Delta holds +-1 entries.
 *
 * Everything here runs on the CPU.  Nothing in this header touches a GPU.
 * All functions return CBM_OK (0) on success and a negative status on
 * failure; cbm_host_last_error() then returns a message for this thread.
 */
#ifndef CBM_HOST_H
#define CBM_HOST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes (shared with cbm_hip.h). */
#ifndef CBM_OK
#define CBM_OK 0
#define CBM_ERR_ARG -1      /* invalid argument (null pointer, negative size, index out of range) */
#define CBM_ERR_VALUE -2    /* input violates a precondition (non-binary value, ...)             */
#define CBM_ERR_NOMEM -3    /* host allocation failed                                             */
#define CBM_ERR_INTERNAL -4 /* internal invariant broken                                          */
#endif

/* Message describing the last failure on the calling thread ("" if none). */
const char *cbm_host_last_error(void);

/* ------------------------------------------------------------------------- */
/* Synthetic graphs: the seeded community copying model.                     */
/* ------------------------------------------------------------------------- */

/*
 * Builds an undirected binary graph without self-loops and returns it as a
 * coalesced COO list sorted by (row, col), both directions present.
 *   n_nodes   number of nodes; nodes are numbered contiguously by block
 *   n_blocks  number of equal blocks (the last one takes the remainder)
 *   p_copy    probability of copying each out-link of the prototype
 *   lam       mean of the Poisson number of uniform random out-links per node
 *   seed_deg  out-links drawn inside its block for the first node of a block
 *   seed      64-bit seed; equal arguments give identical graphs everywhere
 * The result is owned by the library: read it with cbm_graph_{size,row,col}
 * and release it with cbm_graph_free.
 * No reference counterpart: it replaces the network dataset loaders of
 * benchmark/utilities.py:62-198.
 */
typedef struct cbm_graph cbm_graph;
int cbm_graph_copying_model(int64_t n_nodes, int32_t n_blocks, double p_copy, double lam,
                            int32_t seed_deg, uint64_t seed, cbm_graph **out);
int64_t cbm_graph_nnz(const cbm_graph *g);
const int32_t *cbm_graph_row(const cbm_graph *g);
const int32_t *cbm_graph_col(const cbm_graph *g);
void cbm_graph_free(cbm_graph *g);

/* ------------------------------------------------------------------------- */
/* Structure build: replaces cbm_init_ (cbm/cbm_extensions.cpp:60-345).      */
/* ------------------------------------------------------------------------- */

/* Arrays held by a cbm_build result, addressed by cbm_build_array(). */
enum {
    CBM_B_DELTA_ROW = 0,   /* int32 [nnz_delta]  COO rows of Delta, construction order  (ref :301-305)    */
    CBM_B_DELTA_COL = 1,   /* int32 [nnz_delta]  COO cols of Delta                      (ref :307-311)    */
    CBM_B_DELTA_VAL = 2,   /* float [nnz_delta]  +1 / -1                                (ref :313-317)    */
    CBM_B_BRANCHES = 3,    /* int32 [nb + 1]     prefix offsets of the non-trivial branches (ref :319-323) */
    CBM_B_TREE_SRC = 4,    /* int32 [E]          parent of each tree edge, BFS order per branch (ref :325) */
    CBM_B_TREE_DST = 5,    /* int32 [E]          child of each tree edge                 (ref :331-335)   */
    CBM_B_PARENT = 6,      /* int32 [N]          parent(v), or -1 when v hangs off the virtual root       */
    CBM_B_CSR_PTR = 7,     /* int32 [N+1]        Delta as CSR: row offsets (rows ascending)               */
    CBM_B_CSR_COL = 8,     /* int32 [nnz_delta]  Delta as CSR: columns, ascending within a row            */
    CBM_B_CSR_VAL = 9,     /* float [nnz_delta]  Delta as CSR: +1 / -1                                    */
    CBM_B_N_ARRAYS = 10
};

/* Scalar statistics of a cbm_build result, addressed by cbm_build_stat(). */
enum {
    CBM_S_N = 0,            /* rows                                                         */
    CBM_S_NNZ_A = 1,        /* nnz(A) after coalescing                                      */
    CBM_S_NNZ_DELTA = 2,    /* nnz(Delta)                                                   */
    CBM_S_CANDIDATES = 3,   /* candidate edges handed to the arborescence (root edges incl.) */
    CBM_S_MSA_WEIGHT = 4,   /* total weight of the minimum arborescence (== nnz(Delta))     */
    CBM_S_TREE_EDGES = 5,   /* E, edges whose parent is a real node                          */
    CBM_S_BRANCHES = 6,     /* nb, non-trivial branches                                     */
    CBM_S_DEPTH = 7,        /* levels of the forest (a forest of singletons has depth 1)    */
    CBM_S_SYMMETRIC = 8,    /* 1 when A = A^T (then M^T runs through the forward structure) */
    CBM_S_N_STATS = 9
};

typedef struct cbm_build cbm_build;

/*
 * Builds the CBM structure of the binary matrix given as COO (row, col, val).
 * Mirrors cbm_mkl_cpp.init(row, col, val, n_rows, alpha):
 *   - candidate edges are created in the reference order (cbm_extensions.cpp:158-182):
 *     for each row r ascending, the root edge N->r of weight deg(r), then for each
 *     c > r sharing a column with r, ascending: c->r if h < deg(r)-alpha and h != deg(r),
 *     and r->c if h < deg(c)-alpha and h != deg(c), where h = |A_r xor A_c|;
 *   - the minimum spanning arborescence rooted at the virtual node N is found by a
 *     Tarjan-style contraction; ties between equal reduced weights go to the edge
 *     created first;
 *   - Delta rows and the branch/BFS arrays follow cbm_extensions.cpp:191-299, with
 *     children visited in ascending node id.
 * val must be all ones (binary matrix); duplicate (row, col) pairs are merged.
 * n_threads <= 0 uses the OpenMP default.
 */
int cbm_build_create(const int32_t *row, const int32_t *col, const float *val, int64_t nnz,
                     int32_t n_rows, int32_t alpha, int32_t n_threads, cbm_build **out);
int64_t cbm_build_size(const cbm_build *b, int32_t which);
const void *cbm_build_array(const cbm_build *b, int32_t which);
int64_t cbm_build_stat(const cbm_build *b, int32_t which);
void cbm_build_free(cbm_build *b);

/*
 * The three stages of cbm_build_create, exported one by one so that the
 * candidate scan can run on the GPU (cbm_candidates_count / _emit in
 * cbm_hip.h) between the host stages.  No reference counterpart as separate
 * calls: together they are cbm_init_ (cbm/cbm_extensions.cpp:60-345).
 *
 * Stage 1, the binary CSR of A (MKL COO->CSR in the reference, :66-99): rows
 * sorted, duplicate (row, col) pairs merged, values must be all ones.
 * ptr is int64 [n_rows + 1], col int32 [nnz].  t_ptr / t_col are the same
 * for A^T (for each column, the rows holding it, ascending); they are the
 * arrays of A itself when A is symmetric (cbm_adjacency_symmetric() == 1).
 * The candidate scans walk A^T to count row overlaps |A_r & A_c|.
 */
typedef struct cbm_adjacency cbm_adjacency;
int cbm_adjacency_create(const int32_t *row, const int32_t *col, const float *val, int64_t nnz,
                         int32_t n_rows, cbm_adjacency **out);
int32_t cbm_adjacency_rows(const cbm_adjacency *adj);
int64_t cbm_adjacency_nnz(const cbm_adjacency *adj);
const int64_t *cbm_adjacency_ptr(const cbm_adjacency *adj);
const int32_t *cbm_adjacency_col(const cbm_adjacency *adj);
int32_t cbm_adjacency_symmetric(const cbm_adjacency *adj);
const int64_t *cbm_adjacency_t_ptr(const cbm_adjacency *adj);
const int32_t *cbm_adjacency_t_col(const cbm_adjacency *adj);
void cbm_adjacency_free(cbm_adjacency *adj);

/*
 * Stage 2, the candidate edges in the reference creation order
 * (cbm/cbm_extensions.cpp:122-182, the rule given at cbm_build_create):
 * arrays src, dst, weight of cbm_candidates_size() entries.
 */
enum { CBM_C_SRC = 0, CBM_C_DST = 1, CBM_C_WEIGHT = 2 };
typedef struct cbm_candidates cbm_candidates;
int cbm_candidates_create(const cbm_adjacency *adj, int32_t alpha, int32_t n_threads, cbm_candidates **out);
int64_t cbm_candidates_size(const cbm_candidates *c);
const int32_t *cbm_candidates_array(const cbm_candidates *c, int32_t which);
void cbm_candidates_free(cbm_candidates *c);

/*
 * Stage 3, the structure from a candidate list (arborescence, parents,
 * branches, Delta; cbm/cbm_extensions.cpp:188-299).  The list must be the
 * candidate list of `adj` in creation order (as stage 2 or the GPU scan gives
 * it) for the result to equal cbm_build_create's.  An edge outside the
 * candidate graph, a list that does not span the rows, or a chosen tree whose
 * weight is not nnz(Delta) (weights that are not the row distances) fails
 * with CBM_ERR_VALUE.
 */
int cbm_build_from_candidates(const cbm_adjacency *adj, const int32_t *src, const int32_t *dst, const int32_t *w,
                              int64_t n_cand, int32_t n_threads, cbm_build **out);

/* ------------------------------------------------------------------------- */
/* Device plan: the layout the HIP kernels read, built on the host.           */
/* ------------------------------------------------------------------------- */

/*
 * Arrays held by a cbm_plan, addressed by cbm_plan_array().  Positions are
 * indices into an ordered row list.  A row with more than CBM_LONG_ROW
 * entries is "long": it is split into chunks of about CBM_CHUNK entries,
 * each summed by one workgroup; a row of several chunks gets one partial
 * slot per chunk, and a second pass adds its slots in order.
 */
enum {
    CBM_P_LVL_PTR = 0,     /* int32 [L+1]  level offsets into the level order                      */
    CBM_P_LVL_NODE = 1,    /* int32 [N]    node at each level position (level 0 = root children);
                                           within a level, long rows first, each part by node id     */
    CBM_P_LVL_PARENT = 2,  /* int32 [N]    parent of the node at each position, -1 for level 0      */
    CBM_P_FWD_PTR = 3,     /* int32 [N+1]  Delta row offsets, rows in level order                   */
    CBM_P_FWD_COL = 4,     /* int32 [nnz]  Delta columns, bit 31 set for a -1 entry                  */
    CBM_P_BWD_PTR = 5,     /* int32 [N+1]  Delta^T row offsets, rows in BWD_NODE order              */
    CBM_P_BWD_ROW = 6,     /* int32 [nnz]  Delta rows per column, bit 31 set for a -1 entry; rows of
                                           leaves (no children) first, then rows of internal nodes    */
    CBM_P_CH_PTR = 7,      /* int32 [I+1]  child offsets per internal-node position                   */
    CBM_P_CH_NODE = 8,     /* int32 [E]    children (node ids): leaves first, then internal nodes     */
    CBM_P_LVL_NLONG = 9,   /* int32 [L]    long rows at the start of each level                       */
    CBM_P_FCK_PTR = 10,    /* int32 [L+1]  forward chunk offsets per level                            */
    CBM_P_FCK_POS = 11,    /* int32 [C]    level position of each forward chunk                       */
    CBM_P_FCK_META = 12,   /* int32 [C]    chunk index (bits 0-15) and chunk count (bits 16-31)      */
    CBM_P_FCK_SLOT = 13,   /* int32 [C]    partial slot of the chunk, -1 for a one-chunk row          */
    CBM_P_FRED_PTR = 14,   /* int32 [L+1]  offsets per level of the rows that need a reduction       */
    CBM_P_FRED_POS = 15,   /* int32 [R]    their level positions                                      */
    CBM_P_FRED_SLOT = 16,  /* int32 [R]    first partial slot of the row                              */
    CBM_P_FRED_CNT = 17,   /* int32 [R]    number of slots                                            */
    CBM_P_BWD_NODE = 18,   /* int32 [N]    column of Delta (row of dX) at each backward position; long first */
    CBM_P_BCK_POS = 19,    /* int32 [Cb]   backward chunks: position                                  */
    CBM_P_BCK_META = 20,   /* int32 [Cb]   chunk index / count                                        */
    CBM_P_BCK_SLOT = 21,   /* int32 [Cb]   partial slot or -1                                         */
    CBM_P_BRED_POS = 22,   /* int32 [Rb]   backward rows that need a reduction                        */
    CBM_P_BRED_SLOT = 23,  /* int32 [Rb]   first slot                                                 */
    CBM_P_BRED_CNT = 24,   /* int32 [Rb]   number of slots                                            */
    CBM_P_REV_PTR = 25,    /* int32 [L+1]  offsets per level into REV_NODE                                 */
    CBM_P_REV_NODE = 26,   /* int32 [I]    internal nodes (those with children) by level; within a level,
                                           those with more than CBM_LONG_ROW children first, then id ascending */
    CBM_P_BWD_MID = 27,    /* int32 [N]    first internal-node entry of each backward row              */
    CBM_P_CH_MID = 28,     /* int32 [I]    first internal child of each internal-node position         */
    CBM_P_LVL_PPOS = 29,   /* int32 [N]    level position of the parent of each position, -1 for level 0 */
    CBM_P_FCK_GSLOT = 30,  /* int32 [C]    partial slot of each forward chunk numbered over all levels
                                           (one pass over every level at once), -1 for a one-chunk row */
    CBM_P_FRED_GSLOT = 31, /* int32 [R]    first such slot of each forward reduction row               */
    CBM_P_REV_NLONG = 32,  /* int32 [L]    internal nodes with long child lists at the start of each level
                                           of REV_NODE                                                  */
    CBM_P_RCK_PTR = 33,    /* int32 [L+1]  reverse-pass chunk offsets per level                         */
    CBM_P_RCK_POS = 34,    /* int32 [Cr]   internal-node position of each reverse chunk                 */
    CBM_P_RCK_META = 35,   /* int32 [Cr]   chunk index / count                                           */
    CBM_P_RCK_SLOT = 36,   /* int32 [Cr]   partial slot or -1                                            */
    CBM_P_RRED_PTR = 37,   /* int32 [L+1]  offsets per level of the reverse rows that need a reduction   */
    CBM_P_RRED_POS = 38,   /* int32 [Rr]   their internal-node positions                                 */
    CBM_P_RRED_SLOT = 39,  /* int32 [Rr]   first slot                                                    */
    CBM_P_RRED_CNT = 40,   /* int32 [Rr]   number of slots                                               */
    /* Column-range pieces (built only for n_rows <= CBM_RANGE_MAX_ROWS, empty otherwise).  The
       gathered index space is cut into CBM_XCD_RANGES equal ranges; the entries of a row that fall
       in one range are cut into pieces of at most CBM_PIECE entries.  Pieces are grouped by range; each writes one partial row (slot),
       and the slots of a row are consecutive, in (range, part, piece) order. */
    CBM_P_FPC_B = 41,      /* int32 [Pf]   forward pieces: first entry (into FWD_COL)                    */
    CBM_P_FPC_E = 42,      /* int32 [Pf]   end entry                                                     */
    CBM_P_FPC_SLOT = 43,   /* int32 [Pf]   partial slot                                                  */
    CBM_P_FPC_RANGE = 44,  /* int32 [R+1]  piece offsets per range                                       */
    CBM_P_FRS_SLOT = 45,   /* int32 [N]    first slot of each level position                             */
    CBM_P_FRS_CNT = 46,    /* int32 [N]    number of slots (pieces) of each level position               */
    CBM_P_BPC_B = 47,      /* int32 [Pb]   backward pieces: first entry (into BPC_ROW)                   */
    CBM_P_BPC_E = 48,      /* int32 [Pb]   end entry                                                     */
    CBM_P_BPC_MID = 49,    /* int32 [Pb]   first entry of an internal row (leaf rows come first)         */
    CBM_P_BPC_SLOT = 50,   /* int32 [Pb]   partial slot                                                  */
    CBM_P_BPC_RANGE = 51,  /* int32 [R+1]  piece offsets per range                                       */
    CBM_P_BRS_SLOT = 52,   /* int32 [N]    first slot of each backward position                          */
    CBM_P_BRS_CNT = 53,    /* int32 [N]    number of slots of each backward position                     */
    CBM_P_BPC_ROW = 54,    /* int32 [nnz]  BWD_ROW's entries rewritten range by range (a row's extent is
                                           unchanged): in each range the leaf rows, then internal rows  */
    /* The same pieces for CBM_XCD_RANGES_WIDE ranges (same layout, same size limit).  Each of these
       ranges is gathered on two XCDs, one column half each, which halves the partial rows per row
       for rows wide enough to split (see cbm_set_xcd_ranges). */
    CBM_P_FPC4_B = 55,
    CBM_P_FPC4_E = 56,
    CBM_P_FPC4_SLOT = 57,
    CBM_P_FPC4_RANGE = 58,
    CBM_P_FRS4_SLOT = 59,
    CBM_P_FRS4_CNT = 60,
    CBM_P_BPC4_B = 61,
    CBM_P_BPC4_E = 62,
    CBM_P_BPC4_MID = 63,
    CBM_P_BPC4_SLOT = 64,
    CBM_P_BPC4_RANGE = 65,
    CBM_P_BRS4_SLOT = 66,
    CBM_P_BRS4_CNT = 67,
    CBM_P_BPC4_ROW = 68,
    /* Tree tails: the deepest levels [TF_LEVEL, L) of the forward propagation (and [TR_LEVEL, L) of
       the adjoint's reverse pass) are finished in one launch, one workgroup per fragment -- a node of
       the tail's first level with its descendants -- whose levels run one after another between
       workgroup barriers.  Fragment f holds segments [FRAG[f], FRAG[f+1]), one per level from the
       fragment's root down; segment s holds positions POS[SEG[s] .. SEG[s+1]) (level positions for
       the forward, internal-node positions of REV_NODE for the reverse), ascending. */
    CBM_P_TF_FRAG = 69,    /* int32 [Ff+1] forward tail: segment offsets per fragment                  */
    CBM_P_TF_SEG = 70,     /* int32 [Sf+1] position offsets per segment                                */
    CBM_P_TF_POS = 71,     /* int32 [Mf]   level positions                                             */
    CBM_P_TR_FRAG = 72,    /* int32 [Fr+1] reverse tail: segment offsets per fragment                  */
    CBM_P_TR_SEG = 73,     /* int32 [Sr+1] position offsets per segment                                */
    CBM_P_TR_POS = 74,     /* int32 [Mr]   internal-node positions                                     */
    /* Units of the one-launch forward (built only for n_rows <= CBM_UNIT_MAX_ROWS when every Delta row
       fits a unit; empty otherwise).  The tree is cut into connected components of at most
       CBM_UNIT_ROWS rows and CBM_UNIT_PIECES pieces (a row's Delta entries cut into pieces of at most
       CBM_UNIT_PIECE entries, at least one per row); components of equal cut depth are packed into
       units, ordered by that depth, so a unit's outside parents all lie in units of lower index.
       Unit rows are numbered unit after unit, each unit's rows in level order. */
    CBM_P_UN_PTR = 75,     /* int32 [U+1]  row offsets per unit                                        */
    CBM_P_UN_ROW = 76,     /* int32 [N]    level position of each unit row                             */
    CBM_P_UN_LPAR = 77,    /* int32 [N]    parent as a row index within the unit, -1 none, -2 in an
                                           earlier unit (read from the output after that unit's flag)  */
    CBM_P_UN_PC = 78,      /* int32 [N+1]  first piece of each unit row (a row's pieces are consecutive) */
    CBM_P_UN_PCB = 79,     /* int32 [P]    piece: first entry (into FWD_COL)                           */
    CBM_P_UN_PCE = 80,     /* int32 [P]    piece: end entry                                            */
    CBM_P_UN_SEG_PTR = 81, /* int32 [U+1]  segment offsets per unit                                    */
    CBM_P_UN_SEG = 82,     /* int32 [S]    first unit row of each segment (rows of one tree level)     */
    CBM_P_UN_OF = 83,      /* int32 [N]    unit of each level position                                 */
    CBM_P_UN_WAITED = 84,  /* int32 [U]    1 if a later unit reads a parent row of this unit, else 0   */
    CBM_P_N_ARRAYS = 85
};

/* Scalars of a cbm_plan, addressed by cbm_plan_stat(). */
enum {
    CBM_PS_DEPTH = 0,      /* L, number of levels                                   */
    CBM_PS_FWD_SLOTS = 1,  /* partial slots needed by the largest forward level      */
    CBM_PS_BWD_SLOTS = 2,  /* partial slots needed by the backward pass              */
    CBM_PS_BWD_NLONG = 3,  /* long rows at the start of the backward order           */
    CBM_PS_FWD_GSLOTS = 4, /* partial slots of all forward levels together (FCK_GSLOT) */
    CBM_PS_REV_SLOTS = 5,  /* partial slots needed by the largest reverse-pass level    */
    CBM_PS_FWD_PIECES = 6, /* forward column-range pieces (0: none built)               */
    CBM_PS_BWD_PIECES = 7, /* backward column-range pieces                              */
    CBM_PS_FWD_PIECES4 = 8, /* forward pieces of the CBM_XCD_RANGES_WIDE set            */
    CBM_PS_BWD_PIECES4 = 9, /* backward pieces of that set                               */
    CBM_PS_TF_LEVEL = 10,  /* first level of the forward tail (L: none)                   */
    CBM_PS_TR_LEVEL = 11,  /* first level of the reverse tail (L: none)                   */
    CBM_PS_UNITS = 12,     /* units of the one-launch forward (0: none built)              */
    CBM_PS_UNIT_LEVELS = 13, /* component levels (1: no unit waits on another)            */
    CBM_PS_N_STATS = 14
};

#define CBM_LONG_ROW 256
#define CBM_CHUNK 1024
#define CBM_XCD_RANGES 8
#define CBM_XCD_RANGES_WIDE 4
#define CBM_PIECE 256
#define CBM_RANGE_MAX_ROWS 524288
#define CBM_UNIT_ROWS 64
#define CBM_UNIT_PIECES 256
#define CBM_UNIT_PIECE 16
#define CBM_UNIT_MAX_ROWS 524288

typedef struct cbm_plan cbm_plan;

/*
 * Lays out a CBM structure for the GPU.  Inputs are the Delta CSR (any value
 * array whose entries are +-1 times a positive scale; only the signs are kept)
 * and the tree edges (src = parent, dst = child) in any order.
 * Fails with CBM_ERR_VALUE if the tree edges do not form a forest over N nodes
 * (a node with two parents, or a cycle).
 */
int cbm_plan_create(int32_t n_rows, const int32_t *crow, const int32_t *col, const float *val,
                    int64_t nnz, const int32_t *tree_src, const int32_t *tree_dst, int64_t n_edges,
                    int32_t n_threads, cbm_plan **out);
int64_t cbm_plan_size(const cbm_plan *p, int32_t which);
const int32_t *cbm_plan_array(const cbm_plan *p, int32_t which);
int64_t cbm_plan_stat(const cbm_plan *p, int32_t which);
int32_t cbm_plan_depth(const cbm_plan *p);
void cbm_plan_free(cbm_plan *p);

/* ------------------------------------------------------------------------- */
/* Gather order of the Delta pass.                                            */
/* ------------------------------------------------------------------------- */

/*
 * No counterpart in the reference (its Delta SpMM is one cusparseSpMM call,
 * cbm/cbm.cu:184-189).  An order of the short rows (at most CBM_LONG_ROW
 * entries) of an ordered row list (ptr [n+1], as CBM_P_FWD_PTR of a plan) in
 * which rows that share columns sit next to one another, for the row pass
 * to take XCD by XCD (cbm_attach_order in cbm_hip.h): a depth-first walk,
 * heaviest child first, of the maximum spanning forest of the partner edges
 * (partner [n * n_partners], -1 for none, with their overlap counts; from
 * cbm_partners in cbm_hip.h).  Writes the n_order short rows to order [n]
 * (caller-provided, at least n entries).  Any partner lists give a valid
 * order; the order never changes a result, only which rows run together.
 */
int cbm_gather_order(int32_t n, const int32_t *ptr, const int32_t *partner, const int32_t *overlap,
                     int32_t n_partners, int32_t *order, int32_t *n_order);

#ifdef __cplusplus
}
#endif

#endif /* CBM_HOST_H */
