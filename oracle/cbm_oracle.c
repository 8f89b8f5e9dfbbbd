/*
 * cbm_oracle.c -- CPU restatement of the reference CBM construction.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links, loads or calls
 * this file; it is the checker of the parity tests, smoke() and the
 * cpu_baseline leg of bench.py.
 *
 * Follows cbm_init_ in cbm/cbm_extensions.cpp:60-345 with deliberately
 * simple methods that differ from the product's:
 *   - row overlaps by marking one row in a dense array and scanning every
 *     later row (no A*A^T product, no streaming accumulator);
 *   - the minimum arborescence by the textbook Chu-Liu/Edmonds contraction,
 *     O(V E), recursive (the product uses Tarjan's heap-based method);
 *   - Delta rows by a merge of two sorted rows, in the reference's
 *     construction order (root children first, then BFS per branch).
 * Plain C99, single threaded, for graphs up to a few thousand rows.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* Candidate edges, cbm/cbm_extensions.cpp:122-182.                     */
/* ------------------------------------------------------------------ */

/*
 * indptr/indices: CSR of the binary matrix, columns sorted and unique.
 * Writes up to `cap` edges (src, dst, w) in the reference creation order
 * and returns the number of candidates (which may exceed cap: call again).
 */
int64_t oracle_candidates(int32_t n, const int64_t *indptr, const int32_t *indices, int32_t alpha, int64_t cap,
                          int32_t *src, int32_t *dst, int32_t *w) {
    unsigned char *mark = (unsigned char *)calloc((size_t)n, 1);
    int64_t m = 0;
    if (!mark) return -1;
    for (int32_t r = 0; r < n; ++r) {
        const int32_t dr = (int32_t)(indptr[r + 1] - indptr[r]);
        if (m < cap) {
            src[m] = n;
            dst[m] = r;
            w[m] = dr;
        }
        ++m;
        for (int64_t k = indptr[r]; k < indptr[r + 1]; ++k) mark[indices[k]] = 1;
        for (int32_t c = r + 1; c < n; ++c) {
            int32_t common = 0;
            for (int64_t k = indptr[c]; k < indptr[c + 1]; ++k) common += mark[indices[k]];
            if (common == 0) continue; /* the syrk output holds no entry here */
            const int32_t dc = (int32_t)(indptr[c + 1] - indptr[c]);
            const int32_t h = dr + dc - 2 * common;
            if (h < dr - alpha && h != dr) {
                if (m < cap) {
                    src[m] = c;
                    dst[m] = r;
                    w[m] = h;
                }
                ++m;
            }
            if (h < dc - alpha && h != dc) {
                if (m < cap) {
                    src[m] = r;
                    dst[m] = c;
                    w[m] = h;
                }
                ++m;
            }
        }
        for (int64_t k = indptr[r]; k < indptr[r + 1]; ++k) mark[indices[k]] = 0;
    }
    free(mark);
    return m;
}

/* ------------------------------------------------------------------ */
/* Minimum spanning arborescence: Chu-Liu/Edmonds.                     */
/* ------------------------------------------------------------------ */

typedef struct {
    int32_t u, v;
    int64_t w;
    int32_t id; /* index into the caller's edge list at this recursion level */
} OEdge;

/* Returns the weight, or -1 if some vertex has no entering edge.
 * in_edge[v] receives the index (into `e`) of the edge entering v. */
static int64_t edmonds(int32_t n, int32_t root, const OEdge *e, int64_t m, int32_t *in_edge) {
    int32_t *best = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    for (int32_t v = 0; v < n; ++v) best[v] = -1;
    for (int64_t k = 0; k < m; ++k) {
        const OEdge *x = &e[k];
        if (x->u == x->v || x->v == root) continue;
        const int32_t b = best[x->v];
        if (b < 0 || x->w < e[b].w) best[x->v] = (int32_t)k; /* ties keep the earlier edge */
    }
    for (int32_t v = 0; v < n; ++v)
        if (v != root && best[v] < 0) {
            free(best);
            return -1;
        }
    /* Cycles of the best-edge graph: walk parents, colour by start vertex. */
    int32_t *comp = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    int32_t *seen = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    int32_t *on_cycle = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    for (int32_t v = 0; v < n; ++v) {
        comp[v] = -1;
        seen[v] = -1;
        on_cycle[v] = -1;
    }
    int32_t n_new = 0;
    for (int32_t s = 0; s < n; ++s) {
        int32_t v = s;
        while (v != root && seen[v] < 0 && on_cycle[v] < 0) {
            seen[v] = s;
            v = e[best[v]].u;
        }
        if (v != root && seen[v] == s && on_cycle[v] < 0) {
            /* v lies on a new cycle */
            int32_t x = v;
            do {
                on_cycle[x] = n_new;
                comp[x] = n_new;
                x = e[best[x]].u;
            } while (x != v);
            ++n_new;
        }
    }
    if (n_new == 0) {
        int64_t total = 0;
        for (int32_t v = 0; v < n; ++v) {
            in_edge[v] = (v == root) ? -1 : best[v];
            if (v != root) total += e[best[v]].w;
        }
        free(best);
        free(comp);
        free(seen);
        free(on_cycle);
        return total;
    }
    for (int32_t v = 0; v < n; ++v)
        if (comp[v] < 0) comp[v] = n_new++;
    /* Contracted graph: edges entering a cycle vertex v cost w - w(best[v]). */
    OEdge *ce = (OEdge *)malloc(sizeof(OEdge) * (size_t)(m > 0 ? m : 1));
    int64_t cm = 0;
    for (int64_t k = 0; k < m; ++k) {
        const int32_t cu = comp[e[k].u], cv = comp[e[k].v];
        if (cu == cv) continue;
        ce[cm].u = cu;
        ce[cm].v = cv;
        ce[cm].w = e[k].w - (on_cycle[e[k].v] >= 0 ? e[best[e[k].v]].w : 0);
        ce[cm].id = (int32_t)k;
        ++cm;
    }
    int32_t *sub_in = (int32_t *)malloc(sizeof(int32_t) * (size_t)n_new);
    const int64_t sw = edmonds(n_new, comp[root], ce, cm, sub_in);
    int64_t total = -1;
    if (sw >= 0) {
        /* Expand: the edge chosen into a contracted cycle replaces the cycle
         * edge of the vertex it enters; all other cycle vertices keep theirs. */
        for (int32_t v = 0; v < n; ++v) in_edge[v] = (v == root) ? -1 : best[v];
        for (int32_t x = 0; x < n_new; ++x) {
            if (x == comp[root]) continue;
            const int32_t k = ce[sub_in[x]].id;
            in_edge[e[k].v] = k;
        }
        total = 0;
        for (int32_t v = 0; v < n; ++v)
            if (v != root) total += e[in_edge[v]].w;
    }
    free(sub_in);
    free(ce);
    free(best);
    free(comp);
    free(seen);
    free(on_cycle);
    return total;
}

/* Minimum arborescence of (src, dst, w) rooted at `root`; in_edge[v] gets
 * the chosen edge id (-1 at the root).  Returns the weight or -1. */
int64_t oracle_msa(int32_t n, int32_t root, int64_t m, const int32_t *src, const int32_t *dst, const int32_t *w,
                   int32_t *in_edge) {
    OEdge *e = (OEdge *)malloc(sizeof(OEdge) * (size_t)(m > 0 ? m : 1));
    for (int64_t k = 0; k < m; ++k) {
        e[k].u = src[k];
        e[k].v = dst[k];
        e[k].w = w[k];
        e[k].id = (int32_t)k;
    }
    const int64_t r = edmonds(n, root, e, m, in_edge);
    free(e);
    return r;
}

/* ------------------------------------------------------------------ */
/* Delta and the branch arrays for a given parent array,               */
/* cbm/cbm_extensions.cpp:191-299.                                     */
/* ------------------------------------------------------------------ */

/*
 * parent[v] = -1 for children of the virtual root.  Children are visited
 * in ascending id.  Outputs (caller-sized: nnz of Delta is at most
 * 2 * nnz(A), tree edges at most n, branches at most n + 1):
 *   drow/dcol/dval  Delta COO in construction order; returns its length
 *   bsrc/bdst       tree edges in BFS order per branch; *n_edges
 *   branches        prefix offsets, *n_branches + 1 entries
 */
int64_t oracle_delta(int32_t n, const int64_t *indptr, const int32_t *indices, const int32_t *parent, int32_t *drow,
                     int32_t *dcol, float *dval, int32_t *bsrc, int32_t *bdst, int32_t *n_edges, int32_t *branches,
                     int32_t *n_branches) {
    /* children lists, ascending */
    int32_t *cnt = (int32_t *)calloc((size_t)n + 2, sizeof(int32_t));
    for (int32_t v = 0; v < n; ++v) ++cnt[(parent[v] < 0 ? n : parent[v]) + 1];
    for (int32_t u = 0; u <= n; ++u) cnt[u + 1] += cnt[u];
    int32_t *kid = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    int32_t *fill = (int32_t *)malloc(sizeof(int32_t) * ((size_t)n + 1));
    memcpy(fill, cnt, sizeof(int32_t) * ((size_t)n + 1));
    for (int32_t v = 0; v < n; ++v) kid[fill[parent[v] < 0 ? n : parent[v]]++] = v;
    free(fill);

    int64_t nd = 0;
    /* root children: Delta_v = A_v */
    for (int32_t i = cnt[n]; i < cnt[n + 1]; ++i) {
        const int32_t v = kid[i];
        for (int64_t k = indptr[v]; k < indptr[v + 1]; ++k) {
            drow[nd] = v;
            dcol[nd] = indices[k];
            dval[nd] = 1.0f;
            ++nd;
        }
    }
    int32_t *queue = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    int32_t ne = 0, nb = 0;
    branches[0] = 0;
    for (int32_t i = cnt[n]; i < cnt[n + 1]; ++i) {
        const int32_t r = kid[i];
        if (cnt[r] == cnt[r + 1]) continue;
        int32_t head = 0, tail = 0;
        queue[tail++] = r;
        while (head < tail) {
            const int32_t u = queue[head++];
            for (int32_t j = cnt[u]; j < cnt[u + 1]; ++j) {
                const int32_t v = kid[j];
                queue[tail++] = v;
                bsrc[ne] = u;
                bdst[ne] = v;
                ++ne;
                int64_t a = indptr[u], b = indptr[v];
                while (a < indptr[u + 1] && b < indptr[v + 1]) {
                    if (indices[a] == indices[b]) {
                        ++a;
                        ++b;
                    } else if (indices[a] > indices[b]) {
                        drow[nd] = v;
                        dcol[nd] = indices[b++];
                        dval[nd++] = 1.0f;
                    } else {
                        drow[nd] = v;
                        dcol[nd] = indices[a++];
                        dval[nd++] = -1.0f;
                    }
                }
                while (a < indptr[u + 1]) {
                    drow[nd] = v;
                    dcol[nd] = indices[a++];
                    dval[nd++] = -1.0f;
                }
                while (b < indptr[v + 1]) {
                    drow[nd] = v;
                    dcol[nd] = indices[b++];
                    dval[nd++] = 1.0f;
                }
            }
        }
        branches[++nb] = ne;
    }
    *n_edges = ne;
    *n_branches = nb;
    free(queue);
    free(kid);
    free(cnt);
    return nd;
}

/* ------------------------------------------------------------------ */
/* The GPU path's propagation, cbm/cbm.cu:25-91, in FP64.              */
/* ------------------------------------------------------------------ */

/*
 * y (n x f, row-major) holds Delta' X on entry.  For each branch, for each
 * edge in order: y[dst] += y[src] (mult == NULL), or
 * y[dst] += y[src] * mult[dst] / mult[src] (fused_update, :88).
 */
void oracle_update(int32_t f, double *y, int32_t n_branches, const int32_t *branches, const int32_t *src,
                   const int32_t *dst, const double *mult) {
    for (int32_t b = 0; b < n_branches; ++b)
        for (int32_t k = branches[b]; k < branches[b + 1]; ++k) {
            const double *ys = y + (int64_t)src[k] * f;
            double *yd = y + (int64_t)dst[k] * f;
            for (int32_t c = 0; c < f; ++c) yd[c] += mult ? ys[c] * mult[dst[k]] / mult[src[k]] : ys[c];
        }
}
