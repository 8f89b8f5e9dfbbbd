// Device plan: lays a CBM structure out the way the HIP kernels read it.
//
//   level order  nodes sorted by (depth, id); depth 0 = children of the
//                virtual root.  A node's parent always sits in an earlier
//                level, so one pass per level (or a dependency wait on the
//                parent) gives a race-free propagation.
//   forward      Delta rows permuted into level order; each column index
//                carries the entry's sign in bit 31 (no value array).
//   backward     Delta^T (the CSC of Delta) in natural row order, signs packed
//                the same way, for the adjoint's Delta^T product.
//   children     per level position, the node ids of the children, for the
//                adjoint's reverse (leaves-to-roots) gather.
//   pieces       (matrices up to CBM_RANGE_MAX_ROWS rows) each forward and
//                backward row cut by column range of the gathered matrix, so
//                that the workgroups of one XCD gather one range only; two
//                sets, for CBM_XCD_RANGES and CBM_XCD_RANGES_WIDE ranges.
#include "common.hpp"
#include "../../../include/cbm_host.h"

#include <algorithm>
#include <memory>
#include <new>
#include <numeric>

#include <omp.h>

struct cbm_plan {
    std::vector<int32_t> arr[CBM_P_N_ARRAYS];
    int64_t stat[CBM_PS_N_STATS] = {};
    int32_t depth = 0;
};

using namespace cbmh;

static constexpr uint32_t kSignBit = 0x80000000u;

namespace {

// Chunk and reduction tables of the long rows [b, b + n_long) of an ordered
// row list with offsets ptr.  Slots are numbered from 0 for every call, so
// ranges processed one after another (tree levels) share the scratch rows.
// Returns the number of slots used.
struct Split {
    std::vector<int32_t> ck_pos, ck_meta, ck_slot, red_pos, red_slot, red_cnt;
};

int32_t split_long_rows(const std::vector<int32_t> &ptr, int32_t b, int32_t n_long, Split &s) {
    int32_t slots = 0;
    for (int32_t p = b; p < b + n_long; ++p) {
        const int64_t len = int64_t(ptr[size_t(p) + 1]) - ptr[size_t(p)];
        int64_t cnt = (len + CBM_CHUNK - 1) / CBM_CHUNK;
        if (cnt > 0xFFFF) cnt = 0xFFFF;  // chunks grow instead
        if (cnt < 1) cnt = 1;
        for (int64_t c = 0; c < cnt; ++c) {
            s.ck_pos.push_back(p);
            s.ck_meta.push_back(int32_t(uint32_t(c) | (uint32_t(cnt) << 16)));
            s.ck_slot.push_back(cnt > 1 ? slots + int32_t(c) : -1);
        }
        if (cnt > 1) {
            s.red_pos.push_back(p);
            s.red_slot.push_back(slots);
            s.red_cnt.push_back(int32_t(cnt));
            slots += int32_t(cnt);
        }
    }
    return slots;
}

// Column-range pieces of an ordered row list (see CBM_P_FPC_B in cbm_host.h).
// Row p consists of one or two parts, [ptr[p], cut[p]) and [cut[p], ptr[p+1])
// (cut null: one part), each sorted by gathered index (bit 31 is a sign).
// With two parts the row's entries are rewritten into `order` range by
// range, part 0's entries of a range before part 1's, so that a range is one
// contiguous run [b, e) whose part-0 entries end at mid: a dual pass reads
// [b, mid) from its second source and [mid, e) from its first.
struct Pieces {
    std::vector<int32_t> b, e, mid, slot, range_first, rs_slot, rs_cnt, order;
};

// R ranges of equal width over the gathered index space.  Returns false (and
// builds nothing) if the pieces would not fit int32 numbering.
bool range_pieces(int32_t n, const std::vector<int32_t> &ptr, const std::vector<int32_t> &idx,
                  const std::vector<int32_t> *cut, int32_t R, Pieces &pc) {
    const int64_t width = (int64_t(n) + R - 1) / R;
    const int parts = cut ? 2 : 1;
    // bound[p][q][k]: first entry of part q of row p in range k (k = 0..R).
    const size_t stride = size_t(parts) * (R + 1);
    std::vector<int32_t> bound(size_t(n) * stride);
    std::vector<int32_t> count(size_t(n) * R, 0);  // pieces of row p in range k
#pragma omp parallel for schedule(dynamic, 4096)
    for (int32_t p = 0; p < n; ++p) {
        for (int q = 0; q < parts; ++q) {
            const int32_t lo = q == 0 ? ptr[size_t(p)] : (*cut)[size_t(p)];
            const int32_t hi = (q == 0 && cut) ? (*cut)[size_t(p)] : ptr[size_t(p) + 1];
            int32_t *bd = bound.data() + size_t(p) * stride + size_t(q) * (R + 1);
            int32_t i = lo;
            for (int32_t k = 0; k <= R; ++k) {
                const int64_t start = std::min<int64_t>(int64_t(k) * width, int64_t(n));
                while (i < hi && int64_t(uint32_t(idx[size_t(i)]) & ~kSignBit) < start) ++i;
                bd[k] = k == R ? hi : i;
            }
        }
        for (int32_t k = 0; k < R; ++k) {
            int32_t len = 0;
            for (int q = 0; q < parts; ++q) {
                const int32_t *bd = bound.data() + size_t(p) * stride + size_t(q) * (R + 1);
                len += bd[k + 1] - bd[k];
            }
            count[size_t(p) * R + size_t(k)] = (len + CBM_PIECE - 1) / CBM_PIECE;
        }
    }
    int64_t all = 0;
    for (const int32_t c : count) all += c;
    if (all > INT32_MAX) return false;
    // Slots: a row's pieces are consecutive, in (range, part, piece) order.
    pc.rs_slot.assign(size_t(n), 0);
    pc.rs_cnt.assign(size_t(n), 0);
    int64_t total = 0;
    for (int32_t p = 0; p < n; ++p) {
        int32_t c = 0;
        for (int32_t k = 0; k < R; ++k) c += count[size_t(p) * R + size_t(k)];
        pc.rs_slot[size_t(p)] = int32_t(total);
        pc.rs_cnt[size_t(p)] = c;
        total += c;
    }
    // Piece order: grouped by range, rows ascending inside a range.
    pc.range_first.assign(size_t(R) + 1, 0);
    std::vector<int64_t> at(size_t(n) * R);
    int64_t next = 0;
    for (int32_t k = 0; k < R; ++k) {
        pc.range_first[size_t(k)] = int32_t(next);
        for (int32_t p = 0; p < n; ++p) {
            at[size_t(p) * R + size_t(k)] = next;
            next += count[size_t(p) * R + size_t(k)];
        }
    }
    pc.range_first[size_t(R)] = int32_t(next);
    pc.b.resize(size_t(total));
    pc.e.resize(size_t(total));
    pc.slot.resize(size_t(total));
    if (cut) {
        pc.mid.resize(size_t(total));
        pc.order.resize(idx.size());
    }
#pragma omp parallel for schedule(dynamic, 4096)
    for (int32_t p = 0; p < n; ++p) {
        int32_t s = pc.rs_slot[size_t(p)];
        int32_t w = ptr[size_t(p)];  // next entry of the rewritten row (two parts)
        for (int32_t k = 0; k < R; ++k) {
            int64_t o = at[size_t(p) * R + size_t(k)];
            // The range's run [rb, re) and the end of its part-0 entries.
            int32_t rb, re, mid;
            if (cut) {
                const int32_t *b0 = bound.data() + size_t(p) * stride;
                const int32_t *b1 = b0 + (R + 1);
                rb = w;
                for (int32_t i = b0[k]; i < b0[k + 1]; ++i) pc.order[size_t(w++)] = idx[size_t(i)];
                mid = w;
                for (int32_t i = b1[k]; i < b1[k + 1]; ++i) pc.order[size_t(w++)] = idx[size_t(i)];
                re = w;
            } else {
                const int32_t *bd = bound.data() + size_t(p) * stride;
                rb = bd[k];
                re = bd[k + 1];
                mid = re;
            }
            for (int32_t c0 = rb; c0 < re; c0 += CBM_PIECE, ++o, ++s) {
                const int32_t c1 = std::min(re, c0 + CBM_PIECE);
                pc.b[size_t(o)] = c0;
                pc.e[size_t(o)] = c1;
                pc.slot[size_t(o)] = s;
                if (cut) pc.mid[size_t(o)] = std::min(std::max(mid, c0), c1);
            }
        }
    }
    return true;
}

// Tree tail of one pass (see CBM_P_TF_FRAG in cbm_host.h): the deepest levels
// [level, L) in one launch, one workgroup per fragment.  member(v): whether
// the pass computes node v (every node for the forward propagation, internal
// nodes for the reverse pass); pos_of[v]: v's position in the pass's order.
// The first tail level is chosen by a cost model in microseconds: a level of
// its own costs t_level (a dependent launch); the tail costs t_launch plus
// t_step per step, a fragment taking a step per level it spans and one per
// four of its nodes (a 256-thread workgroup finishes four of the widest rows
// at a time), and the tail as many steps as its slowest fragment or as all
// fragments' steps over the 2048 workgroups the device runs at once, if
// more (many small fragments: N = 2,449,029 with 396k tree edges measured
// 0.13 ms per-level against 0.53 ms as a tail from level 1).  No tail starts
// above min_level; none is built (level = L) unless it saves time.
// Units of the one-launch forward (kernel k_fused_forward): the tree is cut
// into connected pieces ("components", a node with some of its descendants)
// of at most CBM_UNIT_ROWS rows whose Delta rows, cut into pieces of at most
// CBM_UNIT_PIECE entries, make at most CBM_UNIT_PIECES pieces.  Bottom-up:
// a node keeps its children's components while they fit and cuts the
// heaviest ones off otherwise.  A component's level is the number of cuts
// above it; components of one level are packed into units (one workgroup
// each) in position order, so a unit only ever waits on units of lower
// index: the one a component root's parent row lives in.
struct Units {
    std::vector<int32_t> ptr, row, lpar, pc, pcb, pce, seg_ptr, seg, of, waited;
    int32_t levels = 0;
};

bool build_units(int32_t n, int32_t levels, const std::vector<int32_t> &parent, const std::vector<int32_t> &kptr,
                 const std::vector<int32_t> &kid, const std::vector<int32_t> &lvl_node,
                 const std::vector<int32_t> &lvl_ptr, const std::vector<int32_t> &pos_of_node,
                 const std::vector<int32_t> &fptr, Units &u) {
    auto pieces_of = [&](int32_t v) {
        const int32_t len = fptr[size_t(pos_of_node[size_t(v)]) + 1] - fptr[size_t(pos_of_node[size_t(v)])];
        return std::max<int32_t>(1, (len + CBM_UNIT_PIECE - 1) / CBM_UNIT_PIECE);
    };
    // rows / pieces still attached to each node's component, deepest level first
    std::vector<int32_t> rows(size_t(n), 1), pcs(size_t(n), 0);
    std::vector<char> cut(size_t(n), 0);
    std::vector<int32_t> ch;
    for (int32_t p = n - 1; p >= 0; --p) {
        const int32_t v = lvl_node[size_t(p)];
        pcs[size_t(v)] = pieces_of(v);
        if (pcs[size_t(v)] > CBM_UNIT_PIECES) return false;  // a row longer than a unit holds
        int32_t r = 1, q = pcs[size_t(v)];
        ch.assign(kid.begin() + kptr[size_t(v)], kid.begin() + kptr[size_t(v) + 1]);
        for (int32_t c : ch) r += rows[size_t(c)], q += pcs[size_t(c)];
        if (r > CBM_UNIT_ROWS || q > CBM_UNIT_PIECES) {
            // cut the heaviest children first (pieces, then rows, then id)
            std::sort(ch.begin(), ch.end(), [&](int32_t a, int32_t b) {
                if (pcs[size_t(a)] != pcs[size_t(b)]) return pcs[size_t(a)] > pcs[size_t(b)];
                if (rows[size_t(a)] != rows[size_t(b)]) return rows[size_t(a)] > rows[size_t(b)];
                return a < b;
            });
            for (int32_t c : ch) {
                if (r <= CBM_UNIT_ROWS && q <= CBM_UNIT_PIECES) break;
                cut[size_t(c)] = 1;
                r -= rows[size_t(c)];
                q -= pcs[size_t(c)];
            }
        }
        rows[size_t(v)] = r;
        pcs[size_t(v)] = q;
    }
    // component of each node (its nearest cut or root ancestor) and component levels, top down
    std::vector<int32_t> comp_of(static_cast<size_t>(n)), clevel(static_cast<size_t>(n), 0);
    int32_t max_clevel = 0;
    for (int32_t p = 0; p < n; ++p) {
        const int32_t v = lvl_node[size_t(p)], par = parent[size_t(v)];
        if (par < 0) {
            comp_of[size_t(v)] = v;
            clevel[size_t(v)] = 0;
        } else if (cut[size_t(v)]) {
            comp_of[size_t(v)] = v;
            clevel[size_t(v)] = clevel[size_t(comp_of[size_t(par)])] + 1;
        } else {
            comp_of[size_t(v)] = comp_of[size_t(par)];
        }
        max_clevel = std::max(max_clevel, clevel[size_t(comp_of[size_t(v)])]);
    }
    u.levels = max_clevel + 1;
    // members per component root, in position order (the level order)
    std::vector<int32_t> mptr(size_t(n) + 1, 0), memb(static_cast<size_t>(n));
    for (int32_t v = 0; v < n; ++v) ++mptr[size_t(comp_of[size_t(v)]) + 1];
    std::partial_sum(mptr.begin(), mptr.end(), mptr.begin());
    {
        std::vector<int32_t> fill(mptr.begin(), mptr.end() - 1);
        for (int32_t p = 0; p < n; ++p) {
            const int32_t v = lvl_node[size_t(p)];
            memb[size_t(fill[size_t(comp_of[size_t(v)])]++)] = p;
        }
    }
    // component roots by (component level, root position)
    std::vector<int32_t> roots;
    for (int32_t p = 0; p < n; ++p) {
        const int32_t v = lvl_node[size_t(p)];
        if (comp_of[size_t(v)] == v) roots.push_back(v);
    }
    std::stable_sort(roots.begin(), roots.end(),
                     [&](int32_t a, int32_t b) { return clevel[size_t(a)] < clevel[size_t(b)]; });
    // pack components of one level into units
    u.ptr.assign(1, 0);
    u.of.assign(size_t(n), -1);
    u.row.clear();
    int32_t unit_rows = 0, unit_pcs = 0, unit_level = -1;
    std::vector<int32_t> cur;
    auto close_unit = [&]() {
        if (cur.empty()) return;
        std::sort(cur.begin(), cur.end());  // positions ascending = level order
        const int32_t id = int32_t(u.ptr.size()) - 1;
        for (int32_t p : cur) {
            u.of[size_t(p)] = id;
            u.row.push_back(p);
        }
        u.ptr.push_back(int32_t(u.row.size()));
        cur.clear();
        unit_rows = unit_pcs = 0;
    };
    for (int32_t r : roots) {
        const int32_t cr = rows[size_t(r)], cq = pcs[size_t(r)];
        if (clevel[size_t(r)] != unit_level || unit_rows + cr > CBM_UNIT_ROWS || unit_pcs + cq > CBM_UNIT_PIECES)
            close_unit();
        unit_level = clevel[size_t(r)];
        for (int32_t k = mptr[size_t(r)]; k < mptr[size_t(r) + 1]; ++k) cur.push_back(memb[size_t(k)]);
        unit_rows += cr;
        unit_pcs += cq;
    }
    close_unit();
    const int32_t n_units = int32_t(u.ptr.size()) - 1;
    // local parents, pieces and level segments
    std::vector<int32_t> local(size_t(n), -1);  // index of a position within its unit
    u.lpar.resize(size_t(n));
    u.pc.assign(size_t(n) + 1, 0);
    u.pcb.clear();
    u.pce.clear();
    u.seg_ptr.assign(1, 0);
    u.seg.clear();
    std::vector<int32_t> level_of_pos(static_cast<size_t>(n));
    for (int32_t l = 0; l < levels; ++l)
        for (int32_t p = lvl_ptr[size_t(l)]; p < lvl_ptr[size_t(l) + 1]; ++p) level_of_pos[size_t(p)] = l;
    for (int32_t k = 0; k < n_units; ++k) {
        const int32_t b = u.ptr[size_t(k)], e = u.ptr[size_t(k) + 1];
        for (int32_t i = b; i < e; ++i) local[size_t(u.row[size_t(i)])] = i - b;
        int32_t last_level = -1;
        for (int32_t i = b; i < e; ++i) {
            const int32_t p = u.row[size_t(i)];
            const int32_t v = lvl_node[size_t(p)], par = parent[size_t(v)];
            if (par < 0) {
                u.lpar[size_t(i)] = -1;
            } else {
                const int32_t pp = pos_of_node[size_t(par)];
                u.lpar[size_t(i)] = u.of[size_t(pp)] == k ? local[size_t(pp)] : -2;
            }
            if (level_of_pos[size_t(p)] != last_level) {
                u.seg.push_back(i);
                last_level = level_of_pos[size_t(p)];
            }
            const int32_t kb = fptr[size_t(p)], len = fptr[size_t(p) + 1] - kb;
            const int32_t np = std::max<int32_t>(1, (len + CBM_UNIT_PIECE - 1) / CBM_UNIT_PIECE);
            u.pc[size_t(i)] = int32_t(u.pcb.size());
            for (int32_t q = 0; q < np; ++q) {
                u.pcb.push_back(kb + int32_t(int64_t(len) * q / np));
                u.pce.push_back(kb + int32_t(int64_t(len) * (q + 1) / np));
            }
        }
        u.seg_ptr.push_back(int32_t(u.seg.size()));
    }
    u.pc[size_t(n)] = int32_t(u.pcb.size());
    // units holding a parent that a later unit reads: only these publish a flag
    u.waited.assign(size_t(n_units), 0);
    for (int32_t i = 0; i < n; ++i)
        if (u.lpar[size_t(i)] == -2)
            u.waited[size_t(u.of[size_t(pos_of_node[size_t(parent[size_t(lvl_node[size_t(u.row[size_t(i)])])])])])] = 1;
    return true;
}

struct Tail {
    int32_t level = 0;
    std::vector<int32_t> frag, seg, pos;
};

template <class Member>
Tail tree_tail(int32_t n, int32_t levels, const std::vector<int32_t> &depth, const std::vector<int32_t> &parent,
               const std::vector<int32_t> &kptr, const std::vector<int32_t> &kid, const std::vector<int32_t> &by_level,
               const std::vector<int32_t> &level_ptr, Member member, const std::vector<int32_t> &pos_of,
               int32_t min_level, double t_level, double t_launch, double t_step) {
    Tail t;
    t.level = levels;
    // Subtree aggregates, deepest nodes first: members below and at v, and
    // the deepest level holding one (-1: none).
    std::vector<int32_t> size(size_t(n), 0), deepest(size_t(n), -1);
    for (int64_t i = int64_t(n) - 1; i >= 0; --i) {
        const int32_t v = by_level[size_t(i)];
        if (member(v)) {
            size[size_t(v)] += 1;
            deepest[size_t(v)] = std::max(deepest[size_t(v)], depth[size_t(v)]);
        }
        const int32_t u = parent[size_t(v)];
        if (u >= 0) {
            size[size_t(u)] += size[size_t(v)];
            deepest[size_t(u)] = std::max(deepest[size_t(u)], deepest[size_t(v)]);
        }
    }
    std::vector<int32_t> busy(size_t(levels) + 1, 0);  // levels from l down holding members
    for (int32_t l = levels - 1; l >= 0; --l) {
        bool any = false;
        for (int32_t i = level_ptr[size_t(l)]; i < level_ptr[size_t(l) + 1] && !any; ++i) any = member(by_level[size_t(i)]);
        busy[size_t(l)] = busy[size_t(l) + 1] + (any ? 1 : 0);
    }
    constexpr double kResidentGroups = 2048.0;
    double best = 0.0;
    for (int32_t a = std::max(min_level, 0); a < levels; ++a) {
        double slowest = 0.0, all = 0.0;
        for (int32_t i = level_ptr[size_t(a)]; i < level_ptr[size_t(a) + 1]; ++i) {
            const int32_t r = by_level[size_t(i)];
            if (size[size_t(r)] == 0) continue;
            const double steps = double(deepest[size_t(r)] - a + 1) + double((size[size_t(r)] + 3) / 4);
            slowest = std::max(slowest, steps);
            all += steps;
        }
        const double gain =
            t_level * busy[size_t(a)] - (t_launch + t_step * std::max(slowest, all / kResidentGroups));
        if (gain > best) {
            best = gain;
            t.level = a;
        }
    }
    if (t.level >= levels) return t;
    // Fragments: the level's nodes with members below, in level order; each
    // fragment's members level by level (a breadth-first walk), ascending
    // position inside a level.
    const int32_t a = t.level;
    t.frag.push_back(0);
    t.seg.push_back(0);
    std::vector<int32_t> cur, next;
    for (int32_t i = level_ptr[size_t(a)]; i < level_ptr[size_t(a) + 1]; ++i) {
        const int32_t r = by_level[size_t(i)];
        if (size[size_t(r)] == 0) continue;
        cur.assign(1, r);
        while (!cur.empty()) {
            const size_t first = t.pos.size();
            next.clear();
            for (const int32_t v : cur) {
                if (member(v)) t.pos.push_back(pos_of[size_t(v)]);
                for (int32_t k = kptr[size_t(v)]; k < kptr[size_t(v) + 1]; ++k)
                    if (size[size_t(kid[size_t(k)])] > 0) next.push_back(kid[size_t(k)]);
            }
            std::sort(t.pos.begin() + int64_t(first), t.pos.end());
            t.seg.push_back(int32_t(t.pos.size()));
            cur.swap(next);
        }
        t.frag.push_back(int32_t(t.seg.size()) - 1);
    }
    return t;
}

// Where each piece set's tables go.
struct PieceSlots {
    int32_t n_ranges;
    int fwd_stat, bwd_stat;
    int fb, fe, fslot, frange, frs_slot, frs_cnt;
    int bb, be, bmid, bslot, brange, brs_slot, brs_cnt, brow;
};
constexpr PieceSlots kPieceSets[2] = {
    {CBM_XCD_RANGES, CBM_PS_FWD_PIECES, CBM_PS_BWD_PIECES, CBM_P_FPC_B, CBM_P_FPC_E, CBM_P_FPC_SLOT,
     CBM_P_FPC_RANGE, CBM_P_FRS_SLOT, CBM_P_FRS_CNT, CBM_P_BPC_B, CBM_P_BPC_E, CBM_P_BPC_MID, CBM_P_BPC_SLOT,
     CBM_P_BPC_RANGE, CBM_P_BRS_SLOT, CBM_P_BRS_CNT, CBM_P_BPC_ROW},
    {CBM_XCD_RANGES_WIDE, CBM_PS_FWD_PIECES4, CBM_PS_BWD_PIECES4, CBM_P_FPC4_B, CBM_P_FPC4_E, CBM_P_FPC4_SLOT,
     CBM_P_FPC4_RANGE, CBM_P_FRS4_SLOT, CBM_P_FRS4_CNT, CBM_P_BPC4_B, CBM_P_BPC4_E, CBM_P_BPC4_MID,
     CBM_P_BPC4_SLOT, CBM_P_BPC4_RANGE, CBM_P_BRS4_SLOT, CBM_P_BRS4_CNT, CBM_P_BPC4_ROW},
};

}  // namespace

extern "C" int cbm_plan_create(int32_t n, const int32_t *crow, const int32_t *col, const float *val, int64_t nnz,
                               const int32_t *tree_src, const int32_t *tree_dst, int64_t n_edges, int32_t n_threads,
                               cbm_plan **out) {
    if (!out || n < 1 || !crow || nnz < 0 || (nnz > 0 && (!col || !val)) || n_edges < 0 ||
        (n_edges > 0 && (!tree_src || !tree_dst)) || nnz > INT32_MAX) {
        set_error("plan: invalid arguments (n=%d nnz=%lld edges=%lld)", n, (long long)nnz, (long long)n_edges);
        return CBM_ERR_ARG;
    }
    *out = nullptr;
    if (crow[0] != 0 || int64_t(crow[n]) != nnz) {
        set_error("plan: row offsets run from %d to %d, expected 0 to %lld", crow[0], crow[n], (long long)nnz);
        return CBM_ERR_ARG;
    }
    for (int32_t r = 0; r < n; ++r)
        if (crow[r + 1] < crow[r]) {
            set_error("plan: row offsets decrease at row %d", r);
            return CBM_ERR_ARG;
        }
    const int saved_threads = omp_get_max_threads();
    if (n_threads > 0) omp_set_num_threads(n_threads);
    struct Restore {
        int t;
        ~Restore() { omp_set_num_threads(t); }
    } restore{saved_threads};
    try {
        auto plan = std::make_unique<cbm_plan>();
        // Parent array from the tree edges; a node may have at most one parent.
        std::vector<int32_t> parent(size_t(n), -1);
        for (int64_t e = 0; e < n_edges; ++e) {
            const int32_t u = tree_src[e], v = tree_dst[e];
            if (u < 0 || u >= n || v < 0 || v >= n) {
                set_error("plan: tree edge %lld = (%d -> %d) is outside 0..%d", (long long)e, u, v, n - 1);
                return CBM_ERR_ARG;
            }
            if (parent[size_t(v)] != -1 || u == v) {
                set_error("plan: node %d has more than one parent or is its own parent", v);
                return CBM_ERR_VALUE;
            }
            parent[size_t(v)] = u;
        }
        // Children by parent, ascending id.
        std::vector<int32_t> kptr(size_t(n) + 1, 0);
        std::vector<int32_t> kid(static_cast<size_t>(n_edges));
        for (int32_t v = 0; v < n; ++v)
            if (parent[size_t(v)] >= 0) ++kptr[size_t(parent[size_t(v)]) + 1];
        std::partial_sum(kptr.begin(), kptr.end(), kptr.begin());
        {
            std::vector<int32_t> fill(kptr.begin(), kptr.end() - 1);
            for (int32_t v = 0; v < n; ++v)
                if (parent[size_t(v)] >= 0) kid[size_t(fill[size_t(parent[size_t(v)])]++)] = v;
        }
        // Depth by BFS from the roots; nodes never reached sit on a cycle.
        std::vector<int32_t> depth(size_t(n), -1), queue;
        queue.reserve(size_t(n));
        for (int32_t v = 0; v < n; ++v)
            if (parent[size_t(v)] < 0) {
                depth[size_t(v)] = 0;
                queue.push_back(v);
            }
        for (size_t h = 0; h < queue.size(); ++h) {
            const int32_t u = queue[h];
            for (int32_t k = kptr[size_t(u)]; k < kptr[size_t(u) + 1]; ++k) {
                depth[size_t(kid[size_t(k)])] = depth[size_t(u)] + 1;
                queue.push_back(kid[size_t(k)]);
            }
        }
        if (int64_t(queue.size()) != n) {
            set_error("plan: the tree edges contain a cycle (%lld of %d nodes reachable from a root)",
                      (long long)queue.size(), n);
            return CBM_ERR_VALUE;
        }
        int32_t levels = 0;
        for (int32_t v = 0; v < n; ++v) levels = std::max(levels, depth[size_t(v)] + 1);
        plan->depth = levels;
        plan->stat[CBM_PS_DEPTH] = levels;

        // Level order: counting sort on (depth, short?) -- long rows first in
        // every level, node ids ascending within each part.
        auto &lvl_ptr = plan->arr[CBM_P_LVL_PTR];
        auto &lvl_node = plan->arr[CBM_P_LVL_NODE];
        auto &lvl_parent = plan->arr[CBM_P_LVL_PARENT];
        auto &lvl_nlong = plan->arr[CBM_P_LVL_NLONG];
        auto is_long = [&](int32_t v) { return crow[v + 1] - crow[v] > CBM_LONG_ROW; };
        std::vector<int32_t> key_ptr(size_t(2 * levels) + 1, 0);
        for (int32_t v = 0; v < n; ++v) ++key_ptr[size_t(2 * depth[size_t(v)] + (is_long(v) ? 0 : 1)) + 1];
        std::partial_sum(key_ptr.begin(), key_ptr.end(), key_ptr.begin());
        lvl_ptr.resize(size_t(levels) + 1);
        lvl_nlong.resize(size_t(levels));
        for (int32_t l = 0; l <= levels; ++l) lvl_ptr[size_t(l)] = key_ptr[size_t(2 * l)];
        for (int32_t l = 0; l < levels; ++l) lvl_nlong[size_t(l)] = key_ptr[size_t(2 * l + 1)] - key_ptr[size_t(2 * l)];
        lvl_node.resize(size_t(n));
        lvl_parent.resize(size_t(n));
        {
            std::vector<int32_t> fill(key_ptr.begin(), key_ptr.end() - 1);
            for (int32_t v = 0; v < n; ++v)
                lvl_node[size_t(fill[size_t(2 * depth[size_t(v)] + (is_long(v) ? 0 : 1))]++)] = v;
        }
        for (int32_t p = 0; p < n; ++p) lvl_parent[size_t(p)] = parent[size_t(lvl_node[size_t(p)])];
        std::vector<int32_t> pos_of_node(static_cast<size_t>(n));
        for (int32_t p = 0; p < n; ++p) pos_of_node[size_t(lvl_node[size_t(p)])] = p;
        {
            auto &ppos = plan->arr[CBM_P_LVL_PPOS];
            ppos.resize(size_t(n));
            for (int32_t p = 0; p < n; ++p)
                ppos[size_t(p)] = lvl_parent[size_t(p)] < 0 ? -1 : pos_of_node[size_t(lvl_parent[size_t(p)])];
        }
        // Forward tail: level 0 has no parent term, so a tail starts at 1 or below.
        {
            Tail tf = tree_tail(n, levels, depth, parent, kptr, kid, lvl_node, lvl_ptr, [](int32_t) { return true; },
                                pos_of_node, 1, 3.5, 4.0, 1.0);
            plan->stat[CBM_PS_TF_LEVEL] = tf.level;
            plan->arr[CBM_P_TF_FRAG].swap(tf.frag);
            plan->arr[CBM_P_TF_SEG].swap(tf.seg);
            plan->arr[CBM_P_TF_POS].swap(tf.pos);
        }

        // Forward Delta in level order with packed signs.
        auto &fptr = plan->arr[CBM_P_FWD_PTR];
        auto &fcol = plan->arr[CBM_P_FWD_COL];
        fptr.assign(size_t(n) + 1, 0);
        for (int32_t p = 0; p < n; ++p) {
            const int32_t v = lvl_node[size_t(p)];
            fptr[size_t(p) + 1] = fptr[size_t(p)] + (crow[v + 1] - crow[v]);
        }
        fcol.resize(size_t(nnz));
        int bad = -1;
#pragma omp parallel for schedule(dynamic, 4096) reduction(max : bad)
        for (int32_t p = 0; p < n; ++p) {
            const int32_t v = lvl_node[size_t(p)];
            int32_t o = fptr[size_t(p)];
            for (int32_t k = crow[v]; k < crow[v + 1]; ++k, ++o) {
                const int32_t c = col[k];
                const float x = val[k];
                if (c < 0 || c >= n || !(x != 0.0f) || x != x) bad = std::max(bad, k);
                fcol[size_t(o)] = int32_t(uint32_t(c) | (x < 0.0f ? kSignBit : 0u));
            }
        }
        if (bad >= 0) {
            set_error("plan: Delta entry %d has an invalid column (%d) or a zero/NaN value", bad, col[bad]);
            return CBM_ERR_VALUE;
        }

        // Forward chunk tables, one set per level.
        {
            Split f;
            auto &fck_ptr = plan->arr[CBM_P_FCK_PTR];
            auto &fred_ptr = plan->arr[CBM_P_FRED_PTR];
            fck_ptr.assign(size_t(levels) + 1, 0);
            fred_ptr.assign(size_t(levels) + 1, 0);
            int32_t max_slots = 0;
            for (int32_t l = 0; l < levels; ++l) {
                max_slots = std::max(max_slots, split_long_rows(fptr, lvl_ptr[size_t(l)], lvl_nlong[size_t(l)], f));
                fck_ptr[size_t(l) + 1] = int32_t(f.ck_pos.size());
                fred_ptr[size_t(l) + 1] = int32_t(f.red_pos.size());
            }
            // The same chunks with slots numbered over all levels, for a pass
            // that runs every level's rows in one launch: a row's chunks are
            // consecutive, so a multi-chunk row takes the next cnt slots.
            auto &gslot = plan->arr[CBM_P_FCK_GSLOT];
            auto &rgslot = plan->arr[CBM_P_FRED_GSLOT];
            gslot.assign(f.ck_pos.size(), -1);
            rgslot.assign(f.red_pos.size(), 0);
            int32_t next = 0;
            size_t r = 0;
            for (size_t i = 0; i < f.ck_pos.size();) {
                const int32_t cnt = int32_t(uint32_t(f.ck_meta[i]) >> 16);
                if (f.ck_slot[i] < 0) {
                    ++i;
                    continue;
                }
                if (r >= f.red_pos.size() || f.red_pos[r] != f.ck_pos[i] || f.red_cnt[r] != cnt) {
                    set_error("plan: forward chunk and reduction tables disagree at chunk %zu", i);
                    return CBM_ERR_INTERNAL;
                }
                rgslot[r++] = next;
                for (int32_t c = 0; c < cnt; ++c) gslot[i + size_t(c)] = next + c;
                i += size_t(cnt);
                next += cnt;
            }
            plan->stat[CBM_PS_FWD_GSLOTS] = next;
            plan->arr[CBM_P_FCK_POS].swap(f.ck_pos);
            plan->arr[CBM_P_FCK_META].swap(f.ck_meta);
            plan->arr[CBM_P_FCK_SLOT].swap(f.ck_slot);
            plan->arr[CBM_P_FRED_POS].swap(f.red_pos);
            plan->arr[CBM_P_FRED_SLOT].swap(f.red_slot);
            plan->arr[CBM_P_FRED_CNT].swap(f.red_cnt);
            plan->stat[CBM_PS_FWD_SLOTS] = max_slots;
        }
        if (n <= CBM_UNIT_MAX_ROWS) {
            Units un;
            if (build_units(n, levels, parent, kptr, kid, lvl_node, lvl_ptr, pos_of_node, fptr, un)) {
                plan->stat[CBM_PS_UNITS] = int64_t(un.ptr.size()) - 1;
                plan->stat[CBM_PS_UNIT_LEVELS] = un.levels;
                plan->arr[CBM_P_UN_PTR].swap(un.ptr);
                plan->arr[CBM_P_UN_ROW].swap(un.row);
                plan->arr[CBM_P_UN_LPAR].swap(un.lpar);
                plan->arr[CBM_P_UN_PC].swap(un.pc);
                plan->arr[CBM_P_UN_PCB].swap(un.pcb);
                plan->arr[CBM_P_UN_PCE].swap(un.pce);
                plan->arr[CBM_P_UN_SEG_PTR].swap(un.seg_ptr);
                plan->arr[CBM_P_UN_SEG].swap(un.seg);
                plan->arr[CBM_P_UN_OF].swap(un.of);
                plan->arr[CBM_P_UN_WAITED].swap(un.waited);
            }
        }
        const bool ranges = n <= CBM_RANGE_MAX_ROWS;
        if (ranges) {
            for (const PieceSlots &ps : kPieceSets) {
                Pieces pc;
                if (!range_pieces(n, fptr, fcol, nullptr, ps.n_ranges, pc)) pc = Pieces();
                plan->stat[ps.fwd_stat] = int64_t(pc.b.size());
                plan->arr[ps.fb].swap(pc.b);
                plan->arr[ps.fe].swap(pc.e);
                plan->arr[ps.fslot].swap(pc.slot);
                plan->arr[ps.frange].swap(pc.range_first);
                plan->arr[ps.frs_slot].swap(pc.rs_slot);
                plan->arr[ps.frs_cnt].swap(pc.rs_cnt);
            }
        }

        // Backward: Delta^T (the CSC of Delta), columns ordered long first,
        // then short, ascending within each part; rows ascending per column.
        {
            // Rows of leaves first, then rows of internal nodes: the adjoint
            // reads a leaf's term from G and an internal node's from Z, and
            // this split keeps both loops free of per-entry branches.
            auto internal = [&](int32_t r) { return kptr[size_t(r) + 1] > kptr[size_t(r)]; };
            // A parallel transpose that writes what a serial one would: the
            // rows are cut into T chunks of about nnz / T entries; chunk t
            // counts its leaf and internal entries per column, and its
            // entries of column j go after those of chunks 0 .. t-1 in each
            // part, so every column still lists its rows in ascending order.
            const int T = std::max(1, omp_get_max_threads());
            std::vector<int32_t> row_cut(size_t(T) + 1, n);
            row_cut[0] = 0;
            for (int t = 1; t < T; ++t)
                row_cut[size_t(t)] = int32_t(std::upper_bound(crow, crow + n + 1, int64_t(nnz) * t / T) - crow) - 1;
            for (int t = 1; t <= T; ++t) row_cut[size_t(t)] = std::max(row_cut[size_t(t)], row_cut[size_t(t) - 1]);
            // cnt[(2 t + part) n + j]: entries of column j in chunk t (part 0 leaf rows, 1 internal rows)
            std::vector<int32_t> cnt(size_t(2 * T) * size_t(n), 0);
#pragma omp parallel for schedule(static, 1)
            for (int t = 0; t < T; ++t) {
                int32_t *leaf = cnt.data() + size_t(2 * t) * size_t(n), *inner = leaf + n;
                for (int32_t r = row_cut[size_t(t)]; r < row_cut[size_t(t) + 1]; ++r) {
                    int32_t *c = internal(r) ? inner : leaf;
                    for (int32_t k = crow[r]; k < crow[r + 1]; ++k) ++c[size_t(col[k])];
                }
            }
            std::vector<int32_t> ccount(size_t(n), 0), cleaf(size_t(n), 0);
#pragma omp parallel for schedule(static)
            for (int32_t j = 0; j < n; ++j) {
                int32_t all = 0, lf = 0;
                for (int t = 0; t < T; ++t) {
                    lf += cnt[size_t(2 * t) * size_t(n) + size_t(j)];
                    all += cnt[size_t(2 * t) * size_t(n) + size_t(j)] + cnt[size_t(2 * t + 1) * size_t(n) + size_t(j)];
                }
                ccount[size_t(j)] = all;
                cleaf[size_t(j)] = lf;
            }
            auto &bnode = plan->arr[CBM_P_BWD_NODE];
            bnode.resize(size_t(n));
            int32_t n_long = 0;
            for (int32_t j = 0; j < n; ++j)
                if (ccount[size_t(j)] > CBM_LONG_ROW) bnode[size_t(n_long++)] = j;
            int32_t q = n_long;
            for (int32_t j = 0; j < n; ++j)
                if (ccount[size_t(j)] <= CBM_LONG_ROW) bnode[size_t(q++)] = j;
            std::vector<int32_t> pos_of(static_cast<size_t>(n));
            for (int32_t p = 0; p < n; ++p) pos_of[size_t(bnode[size_t(p)])] = p;
            auto &bptr = plan->arr[CBM_P_BWD_PTR];
            auto &brow = plan->arr[CBM_P_BWD_ROW];
            bptr.assign(size_t(n) + 1, 0);
            for (int32_t p = 0; p < n; ++p) bptr[size_t(p) + 1] = bptr[size_t(p)] + ccount[size_t(bnode[size_t(p)])];
            brow.resize(size_t(nnz));
            auto &bmid = plan->arr[CBM_P_BWD_MID];
            bmid.resize(size_t(n));
            for (int32_t p = 0; p < n; ++p) bmid[size_t(p)] = bptr[size_t(p)] + cleaf[size_t(bnode[size_t(p)])];
            // The counts become each chunk's first slot per column and part.
#pragma omp parallel for schedule(static)
            for (int32_t j = 0; j < n; ++j) {
                const int32_t p = pos_of[size_t(j)];
                int32_t at_leaf = bptr[size_t(p)], at_int = bmid[size_t(p)];
                for (int t = 0; t < T; ++t) {
                    int32_t &lf = cnt[size_t(2 * t) * size_t(n) + size_t(j)];
                    int32_t &in = cnt[size_t(2 * t + 1) * size_t(n) + size_t(j)];
                    const int32_t nl = lf, ni = in;
                    lf = at_leaf;
                    in = at_int;
                    at_leaf += nl;
                    at_int += ni;
                }
            }
#pragma omp parallel for schedule(static, 1)
            for (int t = 0; t < T; ++t) {
                int32_t *leaf = cnt.data() + size_t(2 * t) * size_t(n), *inner = leaf + n;
                for (int32_t r = row_cut[size_t(t)]; r < row_cut[size_t(t) + 1]; ++r) {
                    int32_t *at = internal(r) ? inner : leaf;
                    for (int32_t k = crow[r]; k < crow[r + 1]; ++k)
                        brow[size_t(at[size_t(col[k])]++)] = int32_t(uint32_t(r) | (val[k] < 0.0f ? kSignBit : 0u));
                }
            }
            std::vector<int32_t>().swap(cnt);
            Split b;
            plan->stat[CBM_PS_BWD_SLOTS] = split_long_rows(bptr, 0, n_long, b);
            plan->stat[CBM_PS_BWD_NLONG] = n_long;
            plan->arr[CBM_P_BCK_POS].swap(b.ck_pos);
            plan->arr[CBM_P_BCK_META].swap(b.ck_meta);
            plan->arr[CBM_P_BCK_SLOT].swap(b.ck_slot);
            plan->arr[CBM_P_BRED_POS].swap(b.red_pos);
            plan->arr[CBM_P_BRED_SLOT].swap(b.red_slot);
            plan->arr[CBM_P_BRED_CNT].swap(b.red_cnt);
            for (const PieceSlots &ps : kPieceSets) {
                if (!ranges || plan->stat[ps.fwd_stat] <= 0) continue;
                Pieces pc;
                if (!range_pieces(n, bptr, brow, &bmid, ps.n_ranges, pc)) pc = Pieces();
                plan->stat[ps.bwd_stat] = int64_t(pc.b.size());
                plan->arr[ps.bb].swap(pc.b);
                plan->arr[ps.be].swap(pc.e);
                plan->arr[ps.bmid].swap(pc.mid);
                plan->arr[ps.bslot].swap(pc.slot);
                plan->arr[ps.brange].swap(pc.range_first);
                plan->arr[ps.brs_slot].swap(pc.rs_slot);
                plan->arr[ps.brs_cnt].swap(pc.rs_cnt);
                plan->arr[ps.brow].swap(pc.order);
            }
        }

        // Internal nodes (those with children) by level, and their children;
        // a child that has children of its own carries the internal bit.
        {
            auto &rptr = plan->arr[CBM_P_REV_PTR];
            auto &rnode = plan->arr[CBM_P_REV_NODE];
            auto &cptr = plan->arr[CBM_P_CH_PTR];
            auto &cnode = plan->arr[CBM_P_CH_NODE];
            auto has_kids = [&](int32_t u) { return kptr[size_t(u) + 1] > kptr[size_t(u)]; };
            auto many_kids = [&](int32_t u) { return kptr[size_t(u) + 1] - kptr[size_t(u)] > CBM_LONG_ROW; };
            // Counting sort on (depth, short?): long child lists first in
            // every level, node ids ascending within each part.
            std::vector<int32_t> rkey(size_t(2 * levels) + 1, 0);
            for (int32_t v = 0; v < n; ++v)
                if (has_kids(v)) ++rkey[size_t(2 * depth[size_t(v)] + (many_kids(v) ? 0 : 1)) + 1];
            std::partial_sum(rkey.begin(), rkey.end(), rkey.begin());
            rptr.resize(size_t(levels) + 1);
            for (int32_t l = 0; l <= levels; ++l) rptr[size_t(l)] = rkey[size_t(2 * l)];
            auto &rnlong = plan->arr[CBM_P_REV_NLONG];
            rnlong.resize(size_t(levels));
            for (int32_t l = 0; l < levels; ++l) rnlong[size_t(l)] = rkey[size_t(2 * l + 1)] - rkey[size_t(2 * l)];
            rnode.resize(size_t(rptr[size_t(levels)]));
            {
                std::vector<int32_t> fill(rkey.begin(), rkey.end() - 1);
                for (int32_t v = 0; v < n; ++v)
                    if (has_kids(v))
                        rnode[size_t(fill[size_t(2 * depth[size_t(v)] + (many_kids(v) ? 0 : 1))]++)] = v;
            }
            auto &cmid = plan->arr[CBM_P_CH_MID];
            cptr.assign(rnode.size() + 1, 0);
            cmid.resize(rnode.size());
            for (size_t i = 0; i < rnode.size(); ++i) {
                const int32_t u = rnode[i];
                cptr[i + 1] = cptr[i] + (kptr[size_t(u) + 1] - kptr[size_t(u)]);
            }
            cnode.resize(size_t(n_edges));
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < int64_t(rnode.size()); ++i) {
                const int32_t u = rnode[size_t(i)];
                int32_t o = cptr[size_t(i)];
                for (int pass = 0; pass < 2; ++pass) {  // leaf children, then internal ones
                    if (pass == 1) cmid[size_t(i)] = o;
                    for (int32_t k = kptr[size_t(u)]; k < kptr[size_t(u) + 1]; ++k) {
                        const int32_t v = kid[size_t(k)];
                        if (has_kids(v) == (pass == 1)) cnode[size_t(o++)] = v;
                    }
                }
            }
            // Chunk and reduction tables of the long child lists, per level.
            Split r;
            auto &rck_ptr = plan->arr[CBM_P_RCK_PTR];
            auto &rred_ptr = plan->arr[CBM_P_RRED_PTR];
            rck_ptr.assign(size_t(levels) + 1, 0);
            rred_ptr.assign(size_t(levels) + 1, 0);
            int32_t max_slots = 0;
            for (int32_t l = 0; l < levels; ++l) {
                max_slots = std::max(max_slots, split_long_rows(cptr, rptr[size_t(l)], rnlong[size_t(l)], r));
                rck_ptr[size_t(l) + 1] = int32_t(r.ck_pos.size());
                rred_ptr[size_t(l) + 1] = int32_t(r.red_pos.size());
            }
            plan->arr[CBM_P_RCK_POS].swap(r.ck_pos);
            plan->arr[CBM_P_RCK_META].swap(r.ck_meta);
            plan->arr[CBM_P_RCK_SLOT].swap(r.ck_slot);
            plan->arr[CBM_P_RRED_POS].swap(r.red_pos);
            plan->arr[CBM_P_RRED_SLOT].swap(r.red_slot);
            plan->arr[CBM_P_RRED_CNT].swap(r.red_cnt);
            plan->stat[CBM_PS_REV_SLOTS] = max_slots;
        }
        // Reverse tail, below the deepest level with long child lists (those
        // take the chunk kernels of the per-level pass).
        {
            const auto &rptr = plan->arr[CBM_P_REV_PTR];
            const auto &rnode = plan->arr[CBM_P_REV_NODE];
            const auto &rnlong = plan->arr[CBM_P_REV_NLONG];
            std::vector<int32_t> rev_pos_of(size_t(n), -1);
            for (int32_t i = 0; i < rptr[size_t(levels)]; ++i) rev_pos_of[size_t(rnode[size_t(i)])] = i;
            int32_t min_level = 0;
            for (int32_t l = 0; l < levels; ++l)
                if (rnlong[size_t(l)] > 0) min_level = l + 1;
            Tail tr = tree_tail(n, levels, depth, parent, kptr, kid, lvl_node, lvl_ptr,
                                [&](int32_t v) { return kptr[size_t(v) + 1] > kptr[size_t(v)]; }, rev_pos_of,
                                min_level, 7.0, 5.0, 2.0);
            plan->stat[CBM_PS_TR_LEVEL] = tr.level;
            plan->arr[CBM_P_TR_FRAG].swap(tr.frag);
            plan->arr[CBM_P_TR_SEG].swap(tr.seg);
            plan->arr[CBM_P_TR_POS].swap(tr.pos);
        }
        *out = plan.release();
        return CBM_OK;
    } catch (const std::bad_alloc &) {
        set_error("plan: out of host memory");
        return CBM_ERR_NOMEM;
    }
}

extern "C" int64_t cbm_plan_size(const cbm_plan *p, int32_t which) {
    if (!p || which < 0 || which >= CBM_P_N_ARRAYS) return -1;
    return int64_t(p->arr[which].size());
}

extern "C" const int32_t *cbm_plan_array(const cbm_plan *p, int32_t which) {
    if (!p || which < 0 || which >= CBM_P_N_ARRAYS) return nullptr;
    return p->arr[which].data();
}

extern "C" int32_t cbm_plan_depth(const cbm_plan *p) { return p ? p->depth : -1; }

extern "C" int64_t cbm_plan_stat(const cbm_plan *p, int32_t which) {
    if (!p || which < 0 || which >= CBM_PS_N_STATS) return -1;
    return p->stat[which];
}

extern "C" void cbm_plan_free(cbm_plan *p) { delete p; }
