// Seeded community copying model (synthetic stand-in for the dataset loaders
// of benchmark/utilities.py:62-198, which need network downloads).
//
// Model, node by node in id order:
//   * nodes are split into n_blocks contiguous blocks of equal size;
//   * the first node of a block draws `seed_deg` distinct-ish out-links
//     uniformly inside its block;
//   * every other node i picks a prototype uniformly among the earlier nodes
//     of its block and copies each out-link of the prototype with
//     probability p_copy (out-links include the ones the prototype copied,
//     so copying is recursive);
//   * every node adds Poisson(lam) out-links to nodes drawn uniformly from
//     the whole graph.
// The directed out-links are then symmetrised, self-loops dropped and
// duplicates merged.  If the last node ends up isolated it is linked to node
// N-2 so that "N = max index + 1" (cbm/cbm4mm.py:12) recovers n_nodes.
#include "common.hpp"
#include "../../../include/cbm_host.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <new>

#include <omp.h>

struct cbm_graph {
    std::vector<int32_t> row, col;
};

namespace cbmh {

static thread_local std::string g_err;

void set_error(const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
}

const char *get_error() { return g_err.c_str(); }

int64_t Rng::poisson(double lam) {
    int64_t total = 0;
    while (lam > 30.0) {
        total += poisson(30.0);
        lam -= 30.0;
    }
    if (lam <= 0.0) return total;
    const double limit = std::exp(-lam);
    double prod = uniform();
    int64_t k = 0;
    while (prod > limit) {
        ++k;
        prod *= uniform();
    }
    return total + k;
}

}  // namespace cbmh

using namespace cbmh;

extern "C" const char *cbm_host_last_error(void) { return get_error(); }

// Sorts (row, col) pairs packed into 64-bit keys with an LSD radix sort on
// 16-bit digits (4 passes); used for the 10^8-edge graphs of the largest
// workload where std::sort would dominate generation time.
static void radix_sort_u64(std::vector<uint64_t> &keys, int key_bits) {
    std::vector<uint64_t> tmp(keys.size());
    const int passes = (key_bits + 15) / 16;
    const int nt = omp_get_max_threads();
    std::vector<int64_t> hist(size_t(nt) * 65536);
    for (int pass = 0; pass < passes; ++pass) {
        const int shift = 16 * pass;
        const int64_t n = int64_t(keys.size());
        std::fill(hist.begin(), hist.end(), 0);
#pragma omp parallel num_threads(nt)
        {
            const int t = omp_get_thread_num();
            const int64_t lo = n * t / nt, hi = n * (t + 1) / nt;
            int64_t *h = hist.data() + size_t(t) * 65536;
            for (int64_t i = lo; i < hi; ++i) ++h[(keys[i] >> shift) & 0xFFFF];
#pragma omp barrier
#pragma omp single
            {
                int64_t run = 0;
                for (int d = 0; d < 65536; ++d)
                    for (int tt = 0; tt < nt; ++tt) {
                        int64_t &c = hist[size_t(tt) * 65536 + d];
                        const int64_t v = c;
                        c = run;
                        run += v;
                    }
            }
            for (int64_t i = lo; i < hi; ++i) {
                const uint64_t k = keys[i];
                tmp[h[(k >> shift) & 0xFFFF]++] = k;
            }
        }
        keys.swap(tmp);
    }
}

extern "C" int cbm_graph_copying_model(int64_t n_nodes, int32_t n_blocks, double p_copy, double lam,
                                       int32_t seed_deg, uint64_t seed, cbm_graph **out) {
    if (!out) {
        set_error("cbm_graph_copying_model: out is null");
        return CBM_ERR_ARG;
    }
    *out = nullptr;
    if (n_nodes < 2 || n_nodes > INT32_MAX || n_blocks < 1 || n_blocks > n_nodes || p_copy < 0.0 ||
        p_copy > 1.0 || lam < 0.0 || seed_deg < 0) {
        set_error("cbm_graph_copying_model: invalid arguments (n=%lld blocks=%d p=%g lam=%g seed_deg=%d)",
                  (long long)n_nodes, n_blocks, p_copy, lam, seed_deg);
        return CBM_ERR_ARG;
    }
    try {
        const int32_t n = int32_t(n_nodes);
        const int64_t bsize = n_nodes / n_blocks;
        // Out-links are appended node by node, so they form a CSR by construction.
        std::vector<int64_t> optr(size_t(n) + 1, 0);
        std::vector<int32_t> olink;
        olink.reserve(size_t(n_nodes) * size_t(std::max(2.0, lam / std::max(1e-3, 1.0 - p_copy))));
        Rng rng(seed);
        for (int32_t i = 0; i < n; ++i) {
            const int64_t b = std::min<int64_t>(i / bsize, n_blocks - 1);
            const int64_t b0 = b * bsize;
            const int64_t b1 = (b == n_blocks - 1) ? n_nodes : b0 + bsize;
            if (i == b0) {
                for (int32_t k = 0; k < seed_deg; ++k)
                    olink.push_back(int32_t(b0 + int64_t(rng.below(uint64_t(b1 - b0)))));
            } else {
                const int64_t proto = b0 + int64_t(rng.below(uint64_t(i - b0)));
                for (int64_t e = optr[proto]; e < optr[proto + 1]; ++e)
                    if (rng.uniform() < p_copy) olink.push_back(olink[e]);
            }
            const int64_t r = rng.poisson(lam);
            for (int64_t k = 0; k < r; ++k) olink.push_back(int32_t(rng.below(uint64_t(n_nodes))));
            optr[i + 1] = int64_t(olink.size());
        }
        // Symmetrise into packed (row << 32 | col) keys, drop self-loops.
        std::vector<uint64_t> keys;
        keys.reserve(olink.size() * 2 + 2);
        for (int32_t i = 0; i < n; ++i)
            for (int64_t e = optr[i]; e < optr[i + 1]; ++e) {
                const int32_t j = olink[e];
                if (j == i) continue;
                keys.push_back((uint64_t(uint32_t(i)) << 32) | uint32_t(j));
                keys.push_back((uint64_t(uint32_t(j)) << 32) | uint32_t(i));
            }
        std::vector<int32_t>().swap(olink);
        std::vector<int64_t>().swap(optr);
        int bits = 1;
        while ((int64_t(1) << bits) < n_nodes) ++bits;
        // Keys are (row << 32 | col); compress to (row << bits | col) for fewer passes.
        for (auto &k : keys) k = ((k >> 32) << bits) | (k & 0xFFFFFFFFull);
        radix_sort_u64(keys, 2 * bits);
        keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
        const uint64_t cmask = (uint64_t(1) << bits) - 1;
        bool last_has_edge = !keys.empty() && int64_t(keys.back() >> bits) == n_nodes - 1;
        auto *g = new cbm_graph;
        g->row.resize(keys.size());
        g->col.resize(keys.size());
        const int64_t m = int64_t(keys.size());
#pragma omp parallel for schedule(static)
        for (int64_t e = 0; e < m; ++e) {
            g->row[e] = int32_t(keys[e] >> bits);
            g->col[e] = int32_t(keys[e] & cmask);
        }
        if (!last_has_edge) {
            // Link the isolated last node to N-2; keep (row, col) order sorted.
            const int32_t a = n - 2, z = n - 1;
            auto pos = std::lower_bound(keys.begin(), keys.end(), (uint64_t(uint32_t(a)) << bits) | uint32_t(z));
            const int64_t at = pos - keys.begin();
            g->row.insert(g->row.begin() + at, a);
            g->col.insert(g->col.begin() + at, z);
            g->row.push_back(z);
            g->col.push_back(a);
        }
        *out = g;
        return CBM_OK;
    } catch (const std::bad_alloc &) {
        set_error("cbm_graph_copying_model: out of host memory");
        return CBM_ERR_NOMEM;
    }
}

extern "C" int64_t cbm_graph_nnz(const cbm_graph *g) { return g ? int64_t(g->row.size()) : -1; }
extern "C" const int32_t *cbm_graph_row(const cbm_graph *g) { return g ? g->row.data() : nullptr; }
extern "C" const int32_t *cbm_graph_col(const cbm_graph *g) { return g ? g->col.data() : nullptr; }
extern "C" void cbm_graph_free(cbm_graph *g) { delete g; }
