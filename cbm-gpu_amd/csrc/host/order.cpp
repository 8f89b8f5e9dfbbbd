// Gather order of the Delta pass (include/cbm_host.h, "gather order").  No
// counterpart in the reference, whose Delta SpMM is one cusparseSpMM call
// (cbm/cbm.cu:184-189): this is an order in which to sum the same rows.
//
// On a matrix far larger than the L2s every row gather of the Delta pass is
// a fabric request unless a row summed shortly before, on the same XCD,
// gathered the same X row.  Rows that share many columns are therefore
// listed next to one another: each row's best-overlapping partners (from
// the GPU partner scan) give the edges of a maximum spanning forest
// (Kruskal, heaviest overlap first), and a depth-first walk of that forest,
// heaviest child first, is the order.  The row pass then gives each XCD one
// contiguous stretch of it (RowPass::order in cbm_kernels.hpp).  Every row
// is still summed by itself, in its own entry order: the order changes which
// rows run together, never a result bit.
//
// Only short rows (at most CBM_LONG_ROW entries) are listed; long rows keep
// their chunks.
#include "../../../include/cbm_host.h"
#include "common.hpp"

#include <omp.h>

#include <algorithm>
#include <new>
#include <numeric>
#include <vector>

using cbmh::set_error;

namespace {

struct Dsu {
    std::vector<int32_t> p;
    explicit Dsu(int32_t n) : p(size_t(n)) { std::iota(p.begin(), p.end(), 0); }
    int32_t find(int32_t x) {
        while (p[size_t(x)] != x) {
            p[size_t(x)] = p[size_t(p[size_t(x)])];
            x = p[size_t(x)];
        }
        return x;
    }
};

struct Edge {
    int32_t w, a, b;
};

}  // namespace

extern "C" int cbm_gather_order(int32_t n, const int32_t *ptr, const int32_t *partner, const int32_t *overlap,
                                int32_t n_partners, int32_t *order, int32_t *n_order) {
    if (n < 1 || !ptr || !order || !n_order || n_partners < 0 || (n_partners > 0 && (!partner || !overlap))) {
        set_error("gather order: invalid arguments (n=%d partners=%d)", n, n_partners);
        return CBM_ERR_ARG;
    }
    *n_order = 0;
    if (ptr[0] != 0) {
        set_error("gather order: row offsets start at %d, expected 0", ptr[0]);
        return CBM_ERR_ARG;
    }
    for (int32_t p = 0; p < n; ++p)
        if (ptr[p + 1] < ptr[p]) {
            set_error("gather order: row offsets decrease at position %d", p);
            return CBM_ERR_ARG;
        }
    try {
        auto is_short = [&](int32_t p) { return ptr[p + 1] - ptr[p] <= CBM_LONG_ROW; };
        // Partner edges between short rows, heaviest first: a counting sort
        // on the overlap, stable in (row, partner rank) order, so the forest
        // is deterministic.
        auto usable = [&](int32_t p, int32_t j, int32_t &q, int32_t &w) {
            q = partner[int64_t(p) * n_partners + j];
            w = overlap[int64_t(p) * n_partners + j];
            return q >= 0 && q < n && q != p && w >= 1 && is_short(q);
        };
        int32_t w_max = 0;
        int64_t n_cand = 0;
        for (int32_t p = 0; p < n; ++p) {
            if (!is_short(p)) continue;
            for (int32_t j = 0; j < n_partners; ++j) {
                int32_t q, w;
                if (usable(p, j, q, w)) {
                    w_max = std::max(w_max, w);
                    ++n_cand;
                }
            }
        }
        std::vector<int64_t> wptr(size_t(w_max) + 2, 0);
        for (int32_t p = 0; p < n; ++p) {
            if (!is_short(p)) continue;
            for (int32_t j = 0; j < n_partners; ++j) {
                int32_t q, w;
                if (usable(p, j, q, w)) ++wptr[size_t(w_max - w) + 1];
            }
        }
        std::partial_sum(wptr.begin(), wptr.end(), wptr.begin());
        std::vector<Edge> edges(static_cast<size_t>(n_cand));
        for (int32_t p = 0; p < n; ++p) {
            if (!is_short(p)) continue;
            for (int32_t j = 0; j < n_partners; ++j) {
                int32_t q, w;
                if (usable(p, j, q, w)) edges[size_t(wptr[size_t(w_max - w)]++)] = Edge{w, p, q};
            }
        }
        // Maximum spanning forest (Kruskal); a pair listed from both ends is
        // taken once, since its second copy joins an already joined pair.
        Dsu dsu(n);
        std::vector<int32_t> deg(size_t(n) + 1, 0);
        std::vector<Edge> tree;
        for (const Edge &e : edges) {
            const int32_t ra = dsu.find(e.a), rb = dsu.find(e.b);
            if (ra == rb) continue;
            dsu.p[size_t(ra)] = rb;
            tree.push_back(e);
            ++deg[size_t(e.a) + 1];
            ++deg[size_t(e.b) + 1];
        }
        std::partial_sum(deg.begin(), deg.end(), deg.begin());
        // Forest adjacency, each list heaviest first (tree is sorted so).
        std::vector<int32_t> adj(static_cast<size_t>(deg[size_t(n)]));
        {
            std::vector<int32_t> fill(deg.begin(), deg.end() - 1);
            for (const Edge &e : tree) {
                adj[size_t(fill[size_t(e.a)]++)] = e.b;
                adj[size_t(fill[size_t(e.b)]++)] = e.a;
            }
        }
        // Depth-first walk from the roots in position order, heaviest child
        // first: children are pushed lightest first.
        std::vector<uint8_t> seen(static_cast<size_t>(n), 0);
        std::vector<int32_t> stack;
        int32_t k = 0;
        for (int32_t r = 0; r < n; ++r) {
            if (!is_short(r) || seen[size_t(r)]) continue;
            seen[size_t(r)] = 1;
            stack.push_back(r);
            while (!stack.empty()) {
                const int32_t x = stack.back();
                stack.pop_back();
                order[k++] = x;
                for (int32_t i = deg[size_t(x) + 1] - 1; i >= deg[size_t(x)]; --i) {
                    const int32_t y = adj[size_t(i)];
                    if (!seen[size_t(y)]) {
                        seen[size_t(y)] = 1;
                        stack.push_back(y);
                    }
                }
            }
        }
        *n_order = k;
        return CBM_OK;
    } catch (const std::bad_alloc &) {
        set_error("gather order: out of host memory");
        return CBM_ERR_NOMEM;
    }
}
