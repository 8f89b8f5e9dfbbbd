// Host CBM structure builder.  Reference: cbm_init_ in
// cbm/cbm_extensions.cpp:60-345 (MKL COO->CSR, MKL syrk, candidate scan,
// arbok TarjanPQ, Delta rows by BFS per branch).
//
// Differences in method, not in result:
//   * A*A^T is never materialised: each row's overlap counts are accumulated
//     in a dense per-thread counter (walking A^T, which is A itself when A is
//     symmetric) and consumed at once (the reference exports the full upper
//     triangle from mkl_sparse_syrk, :101-120);
//   * the candidate list is built once, in parallel over row chunks, and
//     concatenated in row order, which is exactly the creation order of the
//     reference's two serial scans (:122-151 and :158-182);
//   * the arborescence is our own (arborescence.cpp), because arbok is not
//     available; its tie-break rule is documented there;
//   * Delta row lengths are counted first, so the rows are written in parallel
//     straight into their final construction-order slots;
//   * the three stages are also exported one by one (adjacency CSR, candidate
//     edges, structure from a candidate list), so that the candidate scan can
//     run elsewhere (the GPU kernel of cand_api.hip) and feed the same
//     arborescence and Delta code.
#include "arborescence.hpp"
#include "common.hpp"
#include "../../../include/cbm_host.h"

#include <algorithm>
#include <memory>
#include <new>
#include <numeric>

#include <omp.h>

struct cbm_build {
    std::vector<int32_t> arr_i32[CBM_B_N_ARRAYS];
    std::vector<float> delta_val, csr_val;
    int64_t stats[CBM_S_N_STATS] = {};
};

using namespace cbmh;

namespace {

// Binary CSR of A with sorted, unique columns per row.
struct Csr {
    int32_t n = 0;
    std::vector<int64_t> ptr;
    std::vector<int32_t> col;
    int32_t deg(int32_t r) const { return int32_t(ptr[size_t(r) + 1] - ptr[size_t(r)]); }
};

int build_csr(const int32_t *row, const int32_t *col, const float *val, int64_t nnz, int32_t n, Csr &a) {
    a.n = n;
    a.ptr.assign(size_t(n) + 1, 0);
    for (int64_t e = 0; e < nnz; ++e) {
        const int32_t r = row[e], c = col[e];
        if (r < 0 || r >= n || c < 0 || c >= n) {
            set_error("init: entry %lld = (%d, %d) is outside the %d x %d matrix", (long long)e, r, c, n, n);
            return CBM_ERR_ARG;
        }
        if (val && val[e] != 1.0f) {
            set_error("init: entry %lld has value %g; the CBM format needs a binary matrix (all values 1)",
                      (long long)e, double(val[e]));
            return CBM_ERR_VALUE;
        }
        ++a.ptr[size_t(r) + 1];
    }
    std::partial_sum(a.ptr.begin(), a.ptr.end(), a.ptr.begin());
    a.col.resize(size_t(nnz));
    {
        std::vector<int64_t> fill(a.ptr.begin(), a.ptr.end() - 1);
        for (int64_t e = 0; e < nnz; ++e) a.col[size_t(fill[size_t(row[e])]++)] = col[e];
    }
    // Sort each row and merge duplicate entries (the matrix is a 0/1 pattern).
    std::vector<int64_t> keep(size_t(n) + 1, 0);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int32_t r = 0; r < n; ++r) {
        auto b = a.col.begin() + a.ptr[size_t(r)], e = a.col.begin() + a.ptr[size_t(r) + 1];
        std::sort(b, e);
        keep[size_t(r) + 1] = std::unique(b, e) - b;
    }
    std::partial_sum(keep.begin(), keep.end(), keep.begin());
    if (keep[size_t(n)] != nnz) {
        std::vector<int32_t> packed(size_t(keep[size_t(n)]));
#pragma omp parallel for schedule(static)
        for (int32_t r = 0; r < n; ++r)
            std::copy(a.col.begin() + a.ptr[size_t(r)], a.col.begin() + a.ptr[size_t(r)] + (keep[size_t(r) + 1] - keep[size_t(r)]),
                      packed.begin() + keep[size_t(r)]);
        a.col.swap(packed);
        a.ptr.swap(keep);
    }
    return CBM_OK;
}

// Candidate edges in the reference creation order (cbm_extensions.cpp:158-182).
struct Candidates {
    std::vector<int32_t> src, dst, w;
};

}  // namespace

struct cbm_adjacency {
    Csr a;
    Csr at;                 // A^T (the rows holding each column); empty when A is symmetric
    bool symmetric = false;
    const Csr &columns() const { return symmetric ? a : at; }
};

struct cbm_candidates {
    Candidates c;
};

namespace {

// A^T by counting sort: rows of A^T (columns of A) ascending, each with its
// row ids ascending.
void transpose(const Csr &a, Csr &at) {
    const int32_t n = a.n;
    at.n = n;
    at.ptr.assign(size_t(n) + 1, 0);
    for (const int32_t c : a.col) ++at.ptr[size_t(c) + 1];
    std::partial_sum(at.ptr.begin(), at.ptr.end(), at.ptr.begin());
    at.col.resize(a.col.size());
    std::vector<int64_t> fill(at.ptr.begin(), at.ptr.end() - 1);
    for (int32_t r = 0; r < n; ++r)
        for (int64_t i = a.ptr[size_t(r)]; i < a.ptr[size_t(r) + 1]; ++i) at.col[size_t(fill[size_t(a.col[size_t(i)])]++)] = r;
}

// Builds A^T and keeps it only if A is not symmetric.
void attach_columns(cbm_adjacency &adj) {
    transpose(adj.a, adj.at);
    adj.symmetric = adj.at.ptr == adj.a.ptr && adj.at.col == adj.a.col;
    if (adj.symmetric) {
        std::vector<int64_t>().swap(adj.at.ptr);
        std::vector<int32_t>().swap(adj.at.col);
    }
}

// Overlaps |A_r & A_c| = sum over columns k of A_r of [c holds k]: for each
// k in A_r, the rows c > r of column k (row k of `at` = A^T), which is the
// upper triangle of A A^T (the reference's mkl_sparse_syrk, :101-120).
void scan_candidates(const Csr &a, const Csr &at, int32_t alpha, Candidates &out) {
    const int32_t n = a.n;
    const int32_t chunk_rows = 512;
    const int64_t n_chunks = (int64_t(n) + chunk_rows - 1) / chunk_rows;
    std::vector<std::vector<int32_t>> c_src(static_cast<size_t>(n_chunks)), c_dst(static_cast<size_t>(n_chunks)), c_w(static_cast<size_t>(n_chunks));
#pragma omp parallel
    {
        std::vector<int32_t> overlap(size_t(n), 0);  // |A_r and A_c| for the current row r
        std::vector<int32_t> touched, passing;
#pragma omp for schedule(dynamic, 1)
        for (int64_t ch = 0; ch < n_chunks; ++ch) {
            auto &os = c_src[size_t(ch)], &od = c_dst[size_t(ch)], &ow = c_w[size_t(ch)];
            const int32_t r0 = int32_t(ch * chunk_rows);
            const int32_t r1 = int32_t(std::min<int64_t>(n, int64_t(r0) + chunk_rows));
            for (int32_t r = r0; r < r1; ++r) {
                const int32_t dr = a.deg(r);
                // Root edge N -> r with weight deg(r) (cbm_extensions.cpp:164).
                os.push_back(n);
                od.push_back(r);
                ow.push_back(dr);
                touched.clear();
                for (int64_t i = a.ptr[size_t(r)]; i < a.ptr[size_t(r) + 1]; ++i) {
                    const int32_t k = a.col[size_t(i)];
                    // Rows holding column k are sorted: skip to those above r.
                    const int32_t *kb = at.col.data() + at.ptr[size_t(k)];
                    const int32_t *ke = at.col.data() + at.ptr[size_t(k) + 1];
                    for (const int32_t *q = std::upper_bound(kb, ke, r); q < ke; ++q)
                        if (overlap[size_t(*q)]++ == 0) touched.push_back(*q);
                }
                passing.clear();
                for (const int32_t c : touched) {
                    const int32_t dc = a.deg(c);
                    const int32_t h = dr + dc - 2 * overlap[size_t(c)];
                    const bool c_to_r = (h < dr - alpha) && (h != dr);
                    const bool r_to_c = (h < dc - alpha) && (h != dc);
                    if (c_to_r || r_to_c) passing.push_back(c);
                }
                std::sort(passing.begin(), passing.end());
                for (const int32_t c : passing) {
                    const int32_t dc = a.deg(c);
                    const int32_t h = dr + dc - 2 * overlap[size_t(c)];
                    if ((h < dr - alpha) && (h != dr)) {  // c -> r   (:173-175)
                        os.push_back(c);
                        od.push_back(r);
                        ow.push_back(h);
                    }
                    if ((h < dc - alpha) && (h != dc)) {  // r -> c   (:178-180)
                        os.push_back(r);
                        od.push_back(c);
                        ow.push_back(h);
                    }
                }
                for (const int32_t c : touched) overlap[size_t(c)] = 0;
            }
        }
    }
    std::vector<size_t> off(size_t(n_chunks) + 1, 0);
    for (int64_t ch = 0; ch < n_chunks; ++ch) off[size_t(ch) + 1] = off[size_t(ch)] + c_src[size_t(ch)].size();
    out.src.resize(off.back());
    out.dst.resize(off.back());
    out.w.resize(off.back());
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t ch = 0; ch < n_chunks; ++ch) {
        std::copy(c_src[size_t(ch)].begin(), c_src[size_t(ch)].end(), out.src.begin() + int64_t(off[size_t(ch)]));
        std::copy(c_dst[size_t(ch)].begin(), c_dst[size_t(ch)].end(), out.dst.begin() + int64_t(off[size_t(ch)]));
        std::copy(c_w[size_t(ch)].begin(), c_w[size_t(ch)].end(), out.w.begin() + int64_t(off[size_t(ch)]));
        std::vector<int32_t>().swap(c_src[size_t(ch)]);
        std::vector<int32_t>().swap(c_dst[size_t(ch)]);
        std::vector<int32_t>().swap(c_w[size_t(ch)]);
    }
}

// |A_u xor A_v| of two sorted rows.
int64_t symdiff_size(const int32_t *ub, const int32_t *ue, const int32_t *vb, const int32_t *ve) {
    int64_t common = 0;
    const int64_t nu = ue - ub, nv = ve - vb;
    while (ub < ue && vb < ve) {
        if (*ub == *vb) {
            ++common;
            ++ub;
            ++vb;
        } else if (*ub < *vb) {
            ++ub;
        } else {
            ++vb;
        }
    }
    return nu + nv - 2 * common;
}

// Writes row v of Delta = A_v - A_u (u = parent, or none) into (col, val),
// walking both sorted rows like cbm_extensions.cpp:254-291: +1 for columns
// only in v, -1 for columns only in u, in the merge order of the reference.
void write_delta_row(const Csr &a, int32_t v, int32_t u, int32_t *col, float *val) {
    const int32_t *vb = a.col.data() + a.ptr[size_t(v)], *ve = a.col.data() + a.ptr[size_t(v) + 1];
    if (u < 0) {
        for (; vb < ve; ++vb) {
            *col++ = *vb;
            *val++ = 1.0f;
        }
        return;
    }
    const int32_t *ub = a.col.data() + a.ptr[size_t(u)], *ue = a.col.data() + a.ptr[size_t(u) + 1];
    while (ub < ue && vb < ve) {
        if (*ub == *vb) {
            ++ub;
            ++vb;
        } else if (*ub > *vb) {
            *col++ = *vb++;
            *val++ = 1.0f;
        } else {
            *col++ = *ub++;
            *val++ = -1.0f;
        }
    }
    for (; ub < ue; ++ub) {
        *col++ = *ub;
        *val++ = -1.0f;
    }
    for (; vb < ve; ++vb) {
        *col++ = *vb;
        *val++ = 1.0f;
    }
}

// Threads for one call (n_threads <= 0: the OpenMP default), restored on exit.
struct ThreadScope {
    int saved;
    explicit ThreadScope(int32_t n_threads) : saved(omp_get_max_threads()) {
        if (n_threads > 0) omp_set_num_threads(n_threads);
    }
    ~ThreadScope() { omp_set_num_threads(saved); }
};

// The structure of A from its candidate edges (cbm_extensions.cpp:188-299):
// minimum arborescence rooted at N, parents, branches in BFS order, Delta.
// `owned`: the library's own candidate list, released once the parents are
// known; null when the caller supplied the list (then inconsistent weights
// are the caller's error).
int build_from(const Csr &a, bool symmetric, const int32_t *csrc, const int32_t *cdst, const int32_t *cw, int64_t m,
               Candidates *owned, cbm_build **out) {
    const int32_t n = a.n;
    try {
        ArborescenceResult msa = min_arborescence(n + 1, n, csrc, cdst, cw, m);
        if (!msa.spanning || !msa.consistent) {
            set_error("init: arborescence construction failed (spanning=%d consistent=%d)", int(msa.spanning),
                      int(msa.consistent));
            return owned || msa.spanning ? CBM_ERR_INTERNAL : CBM_ERR_VALUE;
        }
        auto *b = new cbm_build;
        std::unique_ptr<cbm_build> guard(b);
        b->stats[CBM_S_N] = n;
        b->stats[CBM_S_NNZ_A] = int64_t(a.col.size());
        b->stats[CBM_S_CANDIDATES] = m;
        b->stats[CBM_S_MSA_WEIGHT] = msa.weight;
        b->stats[CBM_S_SYMMETRIC] = symmetric ? 1 : 0;

        std::vector<int32_t> &parent = b->arr_i32[CBM_B_PARENT];
        parent.resize(size_t(n));
        for (int32_t v = 0; v < n; ++v) {
            const int32_t s = csrc[size_t(msa.in_edge[size_t(v)])];
            parent[size_t(v)] = (s == n) ? -1 : s;
        }
        if (owned) {
            std::vector<int32_t>().swap(owned->src);
            std::vector<int32_t>().swap(owned->dst);
            std::vector<int32_t>().swap(owned->w);
        }

        // Children lists by parent, ascending child id; slot n holds the root's children.
        std::vector<int64_t> cptr(size_t(n) + 2, 0);
        for (int32_t v = 0; v < n; ++v) ++cptr[size_t(parent[size_t(v)] < 0 ? n : parent[size_t(v)]) + 1];
        std::partial_sum(cptr.begin(), cptr.end(), cptr.begin());
        std::vector<int32_t> child(static_cast<size_t>(n));
        {
            std::vector<int64_t> fill(cptr.begin(), cptr.end() - 1);
            for (int32_t v = 0; v < n; ++v) child[size_t(fill[size_t(parent[size_t(v)] < 0 ? n : parent[size_t(v)])]++)] = v;
        }
        auto kids_begin = [&](int32_t u) { return child.data() + cptr[size_t(u)]; };
        auto kids_end = [&](int32_t u) { return child.data() + cptr[size_t(u) + 1]; };

        // Tree edges in BFS order per non-trivial branch (cbm_extensions.cpp:206-217,
        // 232-294); branches are the root children that have children, ascending.
        std::vector<int32_t> &bsrc = b->arr_i32[CBM_B_TREE_SRC];
        std::vector<int32_t> &bdst = b->arr_i32[CBM_B_TREE_DST];
        std::vector<int32_t> &branches = b->arr_i32[CBM_B_BRANCHES];
        branches.push_back(0);
        int32_t depth = 1;
        std::vector<int32_t> level(size_t(n), 0), queue;
        for (const int32_t *r = kids_begin(n); r < kids_end(n); ++r) {
            if (kids_begin(*r) == kids_end(*r)) continue;
            queue.assign(1, *r);
            for (size_t head = 0; head < queue.size(); ++head) {
                const int32_t u = queue[head];
                for (const int32_t *v = kids_begin(u); v < kids_end(u); ++v) {
                    queue.push_back(*v);
                    bsrc.push_back(u);
                    bdst.push_back(*v);
                    level[size_t(*v)] = level[size_t(u)] + 1;
                    depth = std::max(depth, level[size_t(*v)] + 1);
                }
            }
            branches.push_back(int32_t(bsrc.size()));
        }
        b->stats[CBM_S_TREE_EDGES] = int64_t(bsrc.size());
        b->stats[CBM_S_BRANCHES] = int64_t(branches.size()) - 1;
        b->stats[CBM_S_DEPTH] = depth;

        // Delta rows: root children in root-child order, then tree edges in BFS order.
        const int64_t n_root_kids = kids_end(n) - kids_begin(n);
        const int64_t n_rows_out = n_root_kids + int64_t(bdst.size());
        if (n_rows_out != n) {
            set_error("init: tree covers %lld of %d rows", (long long)n_rows_out, n);
            return CBM_ERR_INTERNAL;
        }
        std::vector<int32_t> seq_node(static_cast<size_t>(n)), seq_parent(static_cast<size_t>(n));
        for (int64_t i = 0; i < n_root_kids; ++i) {
            seq_node[size_t(i)] = kids_begin(n)[i];
            seq_parent[size_t(i)] = -1;
        }
        for (size_t e = 0; e < bdst.size(); ++e) {
            seq_node[size_t(n_root_kids) + e] = bdst[e];
            seq_parent[size_t(n_root_kids) + e] = bsrc[e];
        }
        std::vector<int64_t> dptr(size_t(n) + 1, 0);
#pragma omp parallel for schedule(dynamic, 2048)
        for (int32_t i = 0; i < n; ++i) {
            const int32_t v = seq_node[size_t(i)], u = seq_parent[size_t(i)];
            dptr[size_t(i) + 1] = (u < 0) ? a.deg(v)
                                          : symdiff_size(a.col.data() + a.ptr[size_t(u)], a.col.data() + a.ptr[size_t(u) + 1],
                                                         a.col.data() + a.ptr[size_t(v)], a.col.data() + a.ptr[size_t(v) + 1]);
        }
        std::partial_sum(dptr.begin(), dptr.end(), dptr.begin());
        const int64_t nd = dptr[size_t(n)];
        if (nd != msa.weight) {
            set_error("init: nnz(Delta)=%lld differs from the arborescence weight %lld%s", (long long)nd,
                      (long long)msa.weight, owned ? "" : " (candidate weights are not the row distances)");
            return owned ? CBM_ERR_INTERNAL : CBM_ERR_VALUE;
        }
        b->stats[CBM_S_NNZ_DELTA] = nd;
        std::vector<int32_t> &drow = b->arr_i32[CBM_B_DELTA_ROW];
        std::vector<int32_t> &dcol = b->arr_i32[CBM_B_DELTA_COL];
        drow.resize(size_t(nd));
        dcol.resize(size_t(nd));
        b->delta_val.resize(size_t(nd));
#pragma omp parallel for schedule(dynamic, 2048)
        for (int32_t i = 0; i < n; ++i) {
            const int64_t o = dptr[size_t(i)];
            std::fill(drow.begin() + o, drow.begin() + dptr[size_t(i) + 1], seq_node[size_t(i)]);
            write_delta_row(a, seq_node[size_t(i)], seq_parent[size_t(i)], dcol.data() + o, b->delta_val.data() + o);
        }
        // The same rows as CSR (the order torch's coalesce().to_sparse_csr() gives,
        // cbm/cbm4mm.py:26-30): rows ascending, each row's columns already ascending.
        std::vector<int32_t> &cptr_out = b->arr_i32[CBM_B_CSR_PTR];
        std::vector<int32_t> &ccol = b->arr_i32[CBM_B_CSR_COL];
        std::vector<int32_t> seq_of(static_cast<size_t>(n));
        for (int32_t i = 0; i < n; ++i) seq_of[size_t(seq_node[size_t(i)])] = i;
        cptr_out.assign(size_t(n) + 1, 0);
        for (int32_t v = 0; v < n; ++v) {
            const int32_t i = seq_of[size_t(v)];
            cptr_out[size_t(v) + 1] = cptr_out[size_t(v)] + int32_t(dptr[size_t(i) + 1] - dptr[size_t(i)]);
        }
        ccol.resize(size_t(nd));
        b->csr_val.resize(size_t(nd));
#pragma omp parallel for schedule(dynamic, 2048)
        for (int32_t v = 0; v < n; ++v) {
            const int32_t i = seq_of[size_t(v)];
            std::copy(dcol.begin() + dptr[size_t(i)], dcol.begin() + dptr[size_t(i) + 1], ccol.begin() + cptr_out[size_t(v)]);
            std::copy(b->delta_val.begin() + dptr[size_t(i)], b->delta_val.begin() + dptr[size_t(i) + 1],
                      b->csr_val.begin() + cptr_out[size_t(v)]);
        }
        *out = guard.release();
        return CBM_OK;
    } catch (const std::bad_alloc &) {
        set_error("init: out of host memory");
        return CBM_ERR_NOMEM;
    }
}

}  // namespace

extern "C" int cbm_build_create(const int32_t *row, const int32_t *col, const float *val, int64_t nnz,
                                int32_t n_rows, int32_t alpha, int32_t n_threads, cbm_build **out) {
    if (!out || (nnz > 0 && (!row || !col)) || nnz < 0 || n_rows < 1 || n_rows >= INT32_MAX) {
        set_error("init: invalid arguments (n_rows=%d nnz=%lld)", n_rows, (long long)nnz);
        return CBM_ERR_ARG;
    }
    *out = nullptr;
    ThreadScope threads(n_threads);
    try {
        cbm_adjacency adj;
        int rc = build_csr(row, col, val, nnz, n_rows, adj.a);
        if (rc != CBM_OK) return rc;
        attach_columns(adj);
        Candidates cand;
        scan_candidates(adj.a, adj.columns(), alpha, cand);
        std::vector<int64_t>().swap(adj.at.ptr);
        std::vector<int32_t>().swap(adj.at.col);
        return build_from(adj.a, adj.symmetric, cand.src.data(), cand.dst.data(), cand.w.data(), int64_t(cand.src.size()), &cand,
                          out);
    } catch (const std::bad_alloc &) {
        set_error("init: out of host memory");
        return CBM_ERR_NOMEM;
    }
}

extern "C" int cbm_adjacency_create(const int32_t *row, const int32_t *col, const float *val, int64_t nnz,
                                    int32_t n_rows, cbm_adjacency **out) {
    if (!out || (nnz > 0 && (!row || !col)) || nnz < 0 || n_rows < 1 || n_rows >= INT32_MAX) {
        set_error("adjacency: invalid arguments (n_rows=%d nnz=%lld)", n_rows, (long long)nnz);
        return CBM_ERR_ARG;
    }
    *out = nullptr;
    try {
        std::unique_ptr<cbm_adjacency> adj(new cbm_adjacency);
        const int rc = build_csr(row, col, val, nnz, n_rows, adj->a);
        if (rc != CBM_OK) return rc;
        attach_columns(*adj);
        *out = adj.release();
        return CBM_OK;
    } catch (const std::bad_alloc &) {
        set_error("adjacency: out of host memory");
        return CBM_ERR_NOMEM;
    }
}

extern "C" int32_t cbm_adjacency_rows(const cbm_adjacency *adj) { return adj ? adj->a.n : -1; }
extern "C" int64_t cbm_adjacency_nnz(const cbm_adjacency *adj) { return adj ? int64_t(adj->a.col.size()) : -1; }
extern "C" const int64_t *cbm_adjacency_ptr(const cbm_adjacency *adj) { return adj ? adj->a.ptr.data() : nullptr; }
extern "C" const int32_t *cbm_adjacency_col(const cbm_adjacency *adj) { return adj ? adj->a.col.data() : nullptr; }
extern "C" int32_t cbm_adjacency_symmetric(const cbm_adjacency *adj) { return adj ? int32_t(adj->symmetric) : -1; }
extern "C" const int64_t *cbm_adjacency_t_ptr(const cbm_adjacency *adj) {
    return adj ? adj->columns().ptr.data() : nullptr;
}
extern "C" const int32_t *cbm_adjacency_t_col(const cbm_adjacency *adj) {
    return adj ? adj->columns().col.data() : nullptr;
}
extern "C" void cbm_adjacency_free(cbm_adjacency *adj) { delete adj; }

extern "C" int cbm_candidates_create(const cbm_adjacency *adj, int32_t alpha, int32_t n_threads,
                                     cbm_candidates **out) {
    if (!adj || !out) {
        set_error("candidates: invalid arguments");
        return CBM_ERR_ARG;
    }
    *out = nullptr;
    ThreadScope threads(n_threads);
    try {
        std::unique_ptr<cbm_candidates> c(new cbm_candidates);
        scan_candidates(adj->a, adj->columns(), alpha, c->c);
        *out = c.release();
        return CBM_OK;
    } catch (const std::bad_alloc &) {
        set_error("candidates: out of host memory");
        return CBM_ERR_NOMEM;
    }
}

extern "C" int64_t cbm_candidates_size(const cbm_candidates *c) { return c ? int64_t(c->c.src.size()) : -1; }

extern "C" const int32_t *cbm_candidates_array(const cbm_candidates *c, int32_t which) {
    if (!c) return nullptr;
    switch (which) {
        case CBM_C_SRC: return c->c.src.data();
        case CBM_C_DST: return c->c.dst.data();
        case CBM_C_WEIGHT: return c->c.w.data();
        default: return nullptr;
    }
}

extern "C" void cbm_candidates_free(cbm_candidates *c) { delete c; }

extern "C" int cbm_build_from_candidates(const cbm_adjacency *adj, const int32_t *src, const int32_t *dst,
                                         const int32_t *w, int64_t n_cand, int32_t n_threads, cbm_build **out) {
    if (!adj || !out || n_cand < 0 || (n_cand > 0 && (!src || !dst || !w))) {
        set_error("init from candidates: invalid arguments (n_cand=%lld)", (long long)n_cand);
        return CBM_ERR_ARG;
    }
    *out = nullptr;
    const int32_t n = adj->a.n;
    for (int64_t e = 0; e < n_cand; ++e) {
        if (src[e] < 0 || src[e] > n || dst[e] < 0 || dst[e] >= n || src[e] == dst[e] || w[e] < 0) {
            set_error("init from candidates: edge %lld = (%d -> %d, w %d) is not an edge of the %d-row candidate graph",
                      (long long)e, src[e], dst[e], w[e], n);
            return CBM_ERR_VALUE;
        }
    }
    ThreadScope threads(n_threads);
    return build_from(adj->a, adj->symmetric, src, dst, w, n_cand, nullptr, out);
}

extern "C" int64_t cbm_build_size(const cbm_build *b, int32_t which) {
    if (!b || which < 0 || which >= CBM_B_N_ARRAYS) return -1;
    if (which == CBM_B_DELTA_VAL) return int64_t(b->delta_val.size());
    if (which == CBM_B_CSR_VAL) return int64_t(b->csr_val.size());
    return int64_t(b->arr_i32[which].size());
}

extern "C" const void *cbm_build_array(const cbm_build *b, int32_t which) {
    if (!b || which < 0 || which >= CBM_B_N_ARRAYS) return nullptr;
    if (which == CBM_B_DELTA_VAL) return b->delta_val.data();
    if (which == CBM_B_CSR_VAL) return b->csr_val.data();
    return b->arr_i32[which].data();
}

extern "C" int64_t cbm_build_stat(const cbm_build *b, int32_t which) {
    if (!b || which < 0 || which >= CBM_S_N_STATS) return -1;
    return b->stats[which];
}

extern "C" void cbm_build_free(cbm_build *b) { delete b; }
