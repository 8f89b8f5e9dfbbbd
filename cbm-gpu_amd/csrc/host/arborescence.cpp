// Minimum spanning arborescence by Tarjan's contraction method.  The
// algorithm's description and determinism rule are in arborescence.hpp.
#include "arborescence.hpp"

#include <algorithm>
#include <utility>

namespace cbmh {
namespace {

// Leftist heaps over edge ids.  key[x] is exact for a heap root; lazy[x] is
// an addend still owed to both subtrees of x.
struct LeftistForest {
    std::vector<int32_t> key, lazy, left, right;
    std::vector<uint8_t> npl;  // null-path length of the right spine; 0 for "no node"

    LeftistForest(const int32_t *w, size_t m) : key(w, w + m), lazy(m, 0), left(m, -1), right(m, -1), npl(m, 1) {}

    int npl_of(int32_t x) const { return x < 0 ? 0 : npl[size_t(x)]; }

    void shift(int32_t x, int32_t d) {
        if (x >= 0) {
            key[size_t(x)] += d;
            lazy[size_t(x)] += d;
        }
    }

    void settle(int32_t x) {
        const int32_t d = lazy[size_t(x)];
        if (d != 0) {
            shift(left[size_t(x)], d);
            shift(right[size_t(x)], d);
            lazy[size_t(x)] = 0;
        }
    }

    // (reduced weight, edge id) order: lower weight first, then earlier edge.
    bool precedes(int32_t a, int32_t b) const {
        return key[size_t(a)] != key[size_t(b)] ? key[size_t(a)] < key[size_t(b)] : a < b;
    }

    // Melds two heaps; recursion runs down right spines only, so its depth
    // is at most the sum of two null-path lengths, O(log m).
    int32_t meld(int32_t a, int32_t b) {
        if (a < 0) return b;
        if (b < 0) return a;
        if (precedes(b, a)) std::swap(a, b);
        settle(a);
        const int32_t r = meld(right[size_t(a)], b);
        right[size_t(a)] = r;
        if (npl_of(left[size_t(a)]) < npl_of(r)) std::swap(left[size_t(a)], right[size_t(a)]);
        npl[size_t(a)] = uint8_t(npl_of(right[size_t(a)]) + 1);
        return a;
    }

    int32_t pop(int32_t a) {
        settle(a);
        return meld(left[size_t(a)], right[size_t(a)]);
    }
};

// Union-find with union by size and an undo log (no path compression, so a
// union can be reverted exactly).
struct UndoUnionFind {
    std::vector<int32_t> up, size;
    std::vector<std::pair<int32_t, int32_t>> log;  // (attached root, new root)

    explicit UndoUnionFind(int32_t n) : up(size_t(n)), size(size_t(n), 1) {
        for (int32_t i = 0; i < n; ++i) up[size_t(i)] = i;
    }
    int32_t find(int32_t x) const {
        while (up[size_t(x)] != x) x = up[size_t(x)];
        return x;
    }
    void unite(int32_t a, int32_t b) {
        a = find(a);
        b = find(b);
        if (a == b) return;
        if (size[size_t(a)] < size[size_t(b)]) std::swap(a, b);
        up[size_t(b)] = a;
        size[size_t(a)] += size[size_t(b)];
        log.emplace_back(b, a);
    }
    size_t stamp() const { return log.size(); }
    void undo_to(size_t t) {
        while (log.size() > t) {
            const auto [b, a] = log.back();
            log.pop_back();
            up[size_t(b)] = b;
            size[size_t(a)] -= size[size_t(b)];
        }
    }
};

struct Contraction {
    int32_t rep;         // representative of the contracted cycle
    size_t uf_stamp;     // union-find log length before the cycle was merged
    size_t first, last;  // its cycle edges in cycle_edges[first, last)
};

}  // namespace

ArborescenceResult min_arborescence(int32_t n, int32_t root, const int32_t *src, const int32_t *dst,
                                    const int32_t *w, int64_t n_edges) {
    ArborescenceResult res;
    const size_t m = size_t(n_edges);
    LeftistForest heaps(w, m);
    std::vector<int32_t> heap(size_t(n), -1);
    {
        // Initial heaps, built in parallel: each vertex's entering edges as a
        // chain in (weight, id) order, every edge the left child of the one
        // before it -- a leftist heap whose right spines have length one.  A
        // heap's shape never changes a result: pops return the unique least
        // (reduced weight, id), whatever the melds made of the tree.
        std::vector<int64_t> start(size_t(n) + 1, 0);
        for (size_t e = 0; e < m; ++e) ++start[size_t(dst[e]) + 1];
        for (int32_t v = 0; v < n; ++v) start[size_t(v) + 1] += start[size_t(v)];
        std::vector<int32_t> order(m);
        {
            std::vector<int64_t> fill(start.begin(), start.end() - 1);
            for (size_t e = 0; e < m; ++e) order[size_t(fill[size_t(dst[e])]++)] = int32_t(e);
        }
#pragma omp parallel for schedule(dynamic, 256)
        for (int32_t v = 0; v < n; ++v) {
            const int64_t b = start[size_t(v)], e = start[size_t(v) + 1];
            if (b == e) continue;
            std::sort(order.begin() + b, order.begin() + e, [&](int32_t x, int32_t y) {
                return w[size_t(x)] != w[size_t(y)] ? w[size_t(x)] < w[size_t(y)] : x < y;
            });
            for (int64_t k = b; k + 1 < e; ++k) heaps.left[size_t(order[size_t(k)])] = order[size_t(k) + 1];
            heap[size_t(v)] = order[size_t(b)];
        }
    }

    UndoUnionFind uf(n);
    std::vector<int32_t> walk_mark(size_t(n), -1);  // start vertex of the walk holding this super-node
    std::vector<uint8_t> finished(size_t(n), 0);
    std::vector<int32_t> chosen(size_t(n), -1);     // entering edge picked for a super-node
    std::vector<int32_t> path;
    std::vector<int32_t> cycle_edges;
    std::vector<Contraction> contractions;
    finished[size_t(root)] = 1;
    int64_t total = 0;

    for (int32_t s = 0; s < n; ++s) {
        int32_t u = uf.find(s);
        if (finished[size_t(u)]) continue;
        path.clear();
        while (!finished[size_t(u)]) {
            // Cheapest entering edge that leaves the super-node u.
            int32_t e = heap[size_t(u)];
            while (e >= 0 && uf.find(src[size_t(e)]) == u) {
                heap[size_t(u)] = heaps.pop(e);
                e = heap[size_t(u)];
            }
            if (e < 0) return res;  // u unreachable from the root
            const int32_t we = heaps.key[size_t(e)];
            heap[size_t(u)] = heaps.pop(e);
            heaps.shift(heap[size_t(u)], -we);  // reduced costs relative to the pick
            total += we;
            chosen[size_t(u)] = e;
            walk_mark[size_t(u)] = s;
            path.push_back(u);
            const int32_t v = uf.find(src[size_t(e)]);
            if (walk_mark[size_t(v)] == s && !finished[size_t(v)]) {
                // The walk closed a cycle v -> ... -> u -> v: contract it.
                Contraction c;
                c.uf_stamp = uf.stamp();
                c.first = cycle_edges.size();
                int32_t merged = -1, x;
                do {
                    x = path.back();
                    path.pop_back();
                    cycle_edges.push_back(chosen[size_t(x)]);
                    merged = heaps.meld(merged, heap[size_t(x)]);
                    heap[size_t(x)] = -1;
                    uf.unite(x, v);
                } while (x != v);
                c.rep = uf.find(v);
                c.last = cycle_edges.size();
                contractions.push_back(c);
                heap[size_t(c.rep)] = merged;
                walk_mark[size_t(c.rep)] = -1;
                u = c.rep;
            } else {
                u = v;
            }
        }
        for (const int32_t x : path) finished[size_t(x)] = 1;
    }

    // Expand the contractions, outermost (latest) first.
    std::vector<int32_t> in(size_t(n), -1);
    for (int32_t x = 0; x < n; ++x)
        if (x != root && uf.find(x) == x) in[size_t(x)] = chosen[size_t(x)];
    for (auto it = contractions.rbegin(); it != contractions.rend(); ++it) {
        uf.undo_to(it->uf_stamp);
        const int32_t entering = in[size_t(it->rep)];
        for (size_t k = it->first; k < it->last; ++k) {
            const int32_t e = cycle_edges[k];
            in[size_t(uf.find(dst[size_t(e)]))] = e;
        }
        in[size_t(uf.find(dst[size_t(entering)]))] = entering;
    }
    int64_t check = 0;
    for (int32_t x = 0; x < n; ++x)
        if (x != root) check += w[size_t(in[size_t(x)])];
    res.weight = check;
    res.in_edge = std::move(in);
    res.spanning = true;
    res.consistent = (check == total);
    return res;
}

}  // namespace cbmh
