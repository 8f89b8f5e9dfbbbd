// HIP kernels of the CBM product for gfx950 (CDNA4, wave64).
//
// All work is one kind of pass over rows: gather a short list of rows of a
// dense N x F matrix with +-1 (times an optional per-row scale), optionally
// add one more scaled row, store an FP32 row.  The forward product runs it
// once per tree level (Delta row + parent row, parents finished by the
// previous level's launch); the adjoint runs it once per level from the
// deepest up over child lists, then once over Delta^T.
//
// Mapping: a group of LPR lanes (a power of two <= 64) owns one output row;
// each lane owns VE consecutive columns, loaded as one 16-byte vector (4 FP32
// or 8 BF16 values), so a row of F columns is covered by F/VE lanes and a
// wavefront holds 64/LPR rows.  The group's lanes fetch LPR packed indices in
// one coalesced load and broadcast them with cross-lane moves
// (ds_bpermute, or v_readlane when a row owns the whole wavefront); four
// gathers are kept in flight per lane.  Accumulation is FP32 with fma, so a
// +-1 entry is added exactly as an add or a subtract.
#include "row_pass.hpp"

namespace cbmk {
namespace {

using namespace rp;

// y[node_of[p]] += parscale[p] * y[parent_of[p]], the columns of lane l of
// a group of LPR lanes.
// The row's parent par (< 0: none), row and scale ps given.
template <int VE, int LPR>
__device__ inline void propagate_row_at(int32_t par, int32_t row, float ps, int l, float *y, int64_t ldy,
                                        int32_t ncols) {
    if (par < 0) return;
    for (int32_t c0 = l * VE; c0 < ncols; c0 += LPR * VE) {
        float a[VE], b[VE];
        Load<float, VE>::run(y + int64_t(par) * ldy + c0, a);
        Load<float, VE>::run(y + int64_t(row) * ldy + c0, b);
#pragma unroll
        for (int q = 0; q < VE; ++q) b[q] = fmaf(a[q], ps, b[q]);
        store_row<VE>(y + int64_t(row) * ldy + c0, b);
    }
}

template <int VE, int LPR>
__device__ inline void propagate_row(int32_t p, int l, const int32_t *__restrict__ node_of,
                                     const int32_t *__restrict__ parent_of, const float *__restrict__ parscale,
                                     float *y, int64_t ldy, int32_t ncols) {
    const int32_t par = parent_of[p];
    if (par < 0) return;
    propagate_row_at<VE, LPR>(par, node_of[p], parscale ? parscale[p] : 1.0f, l, y, ldy, ncols);
}

template <int VE, int LPR>
__device__ inline void propagate_row_comp(int32_t p, int l, const int32_t *__restrict__ node_of,
                                          const int32_t *__restrict__ parent_of, const float *__restrict__ parscale,
                                          const CompArgs &c, float *y, int64_t ldy, int32_t ncols);

template <int VE, int LPR>
__global__ __launch_bounds__(256) void k_propagate(int32_t p_begin, int32_t p_end, const int32_t *__restrict__ node_of,
                                                   const int32_t *__restrict__ parent_of,
                                                   const float *__restrict__ parscale, float *y, int64_t ldy,
                                                   int32_t ncols) {
    constexpr int kGroupsPerWave = 64 / LPR;
    const int lane = int(threadIdx.x & 63u);
    const int l = lane & (LPR - 1);
    const int64_t group =
        (int64_t(blockIdx.x) * int64_t(blockDim.x >> 6) + int64_t(threadIdx.x >> 6)) * kGroupsPerWave + lane / LPR;
    const int64_t p64 = int64_t(p_begin) + group;
    if (p64 >= int64_t(p_end)) return;
    propagate_row<VE, LPR>(int32_t(p64), l, node_of, parent_of, parscale, y, ldy, ncols);
}

template <int VE, int LPR>
__global__ __launch_bounds__(256) void k_propagate_comp(int32_t p_begin, int32_t p_end,
                                                        const int32_t *__restrict__ node_of,
                                                        const int32_t *__restrict__ parent_of,
                                                        const float *__restrict__ parscale, const CompArgs c, float *y,
                                                        int64_t ldy, int32_t ncols) {
    constexpr int kGroupsPerWave = 64 / LPR;
    const int lane = int(threadIdx.x & 63u);
    const int l = lane & (LPR - 1);
    const int64_t group =
        (int64_t(blockIdx.x) * int64_t(blockDim.x >> 6) + int64_t(threadIdx.x >> 6)) * kGroupsPerWave + lane / LPR;
    const int64_t p64 = int64_t(p_begin) + group;
    if (p64 >= int64_t(p_end)) return;
    propagate_row_comp<VE, LPR>(int32_t(p64), l, node_of, parent_of, parscale, c, y, ldy, ncols);
}

// The compensated form: (y, comp) of position p from the parent's y and
// comp rows; positions from c.base on own a comp row.
// The compensated form with position p's parent par, row, scale es and the
// parent's position pp given.
template <int VE, int LPR>
__device__ inline void propagate_row_comp_at(int32_t p, int32_t par, int32_t row, float es, int32_t pp, int l,
                                             const CompArgs &c, float *y, int64_t ldy, int32_t ncols) {
    for (int32_t c0 = l * VE; c0 < ncols; c0 += LPR * VE) {
        float comp[VE];
#pragma unroll
        for (int q = 0; q < VE; ++q) comp[q] = 0.0f;
        if (par >= 0) {
            float acc[VE], v[VE], cp[VE];
            Load<float, VE>::run(y + int64_t(row) * ldy + c0, acc);
            Load<float, VE>::run(y + int64_t(par) * ldy + c0, v);
            if (pp >= c.base) {
                Load<float, VE>::run(c.comp + int64_t(pp - c.base) * c.ldc + c0, cp);
            } else {
#pragma unroll
                for (int q = 0; q < VE; ++q) cp[q] = 0.0f;
            }
            comp_add<VE>(acc, es, v, cp, comp);
            store_row<VE>(y + int64_t(row) * ldy + c0, acc);
        }
        store_row<VE>(c.comp + int64_t(p - c.base) * c.ldc + c0, comp);
    }
}

template <int VE, int LPR>
__device__ inline void propagate_row_comp(int32_t p, int l, const int32_t *__restrict__ node_of,
                                          const int32_t *__restrict__ parent_of, const float *__restrict__ parscale,
                                          const CompArgs &c, float *y, int64_t ldy, int32_t ncols) {
    propagate_row_comp_at<VE, LPR>(p, parent_of[p], node_of[p], parscale ? parscale[p] : 1.0f, c.parent_pos[p], l, c,
                                   y, ldy, ncols);
}

// Rows of a tail fragment whose position, row, parent, scale and parent
// position the unstaged tail kernel holds in LDS (larger fragments read them
// from the plan's arrays level by level).
constexpr int kTailMetaRows = 512;

// The forward propagation of a tree tail: workgroup f finishes fragment f's
// segments, root first, one group of LPR lanes per row and a workgroup
// barrier between segments; positions from comp_first on take the
// compensated form.  The rows' indices and scales are read into LDS first,
// so a level waits only for its rows' loads and stores.
template <int VE, int LPR>
__global__ __launch_bounds__(256) void k_propagate_tail(const TailArgs t, const int32_t *__restrict__ node_of,
                                                        const int32_t *__restrict__ parent_of,
                                                        const float *__restrict__ parscale, const CompArgs c,
                                                        int32_t comp_first, float *y, int64_t ldy, int32_t ncols) {
    constexpr int kGroups = 256 / LPR;
    __shared__ int32_t m_pos[kTailMetaRows], m_row[kTailMetaRows], m_par[kTailMetaRows], m_pp[kTailMetaRows];
    __shared__ float m_es[kTailMetaRows];
    const int lane = int(threadIdx.x & 63u);
    const int l = lane & (LPR - 1);
    const int32_t group = int32_t(threadIdx.x / LPR);
    const int32_t f = t.order ? t.order[t.n_staged + int32_t(blockIdx.x)] : int32_t(blockIdx.x);
    const int32_t s0 = t.frag[f], s1 = t.frag[f + 1];
    const int32_t rb = t.seg[s0], n = t.seg[s1] - rb;
    if (n >= 0 && n <= kTailMetaRows) {
        for (int32_t i = int32_t(threadIdx.x); i < n; i += 256) {
            const int32_t p = t.pos[rb + i];
            m_pos[i] = p;
            m_row[i] = node_of[p];
            m_par[i] = parent_of[p];
            m_es[i] = parscale ? parscale[p] : 1.0f;
            m_pp[i] = p >= comp_first ? c.parent_pos[p] : -1;
        }
        __syncthreads();
        for (int32_t s = s0; s < s1; ++s) {
            const int32_t ie = t.seg[s + 1] - rb;
            for (int32_t i = t.seg[s] - rb + group; i < ie; i += kGroups) {
                const int32_t p = m_pos[i];
                if (p >= comp_first) {
                    propagate_row_comp_at<VE, LPR>(p, m_par[i], m_row[i], m_es[i], m_pp[i], l, c, y, ldy, ncols);
                } else {
                    propagate_row_at<VE, LPR>(m_par[i], m_row[i], m_es[i], l, y, ldy, ncols);
                }
            }
            __syncthreads();
        }
        return;
    }
    for (int32_t s = s0; s < s1; ++s) {
        for (int32_t i = t.seg[s] + group; i < t.seg[s + 1]; i += kGroups) {
            const int32_t p = t.pos[i];
            if (p >= comp_first) {
                propagate_row_comp<VE, LPR>(p, l, node_of, parent_of, parscale, c, y, ldy, ncols);
            } else {
                propagate_row<VE, LPR>(p, l, node_of, parent_of, parscale, y, ldy, ncols);
            }
        }
        __syncthreads();
    }
}

// The same propagation for the staged fragments, whose chains are long: a
// workgroup takes fragment t.order[blockIdx.x / n_slabs] and kTailSlab columns
// of it.  The fragment's rows (Delta-pass values), positions, parent slots
// and scales are loaded into LDS at once -- a row finds its parent's slot by
// a binary search, since the fragment's positions ascend (segments are tree
// levels, positions are in level order) -- so the levels then run from LDS
// alone, with a workgroup barrier between them, and the rows go out at the
// end.  A parent outside the fragment (the root's, finished by the launches
// before the tail) is read from y and comp.  Every value goes through the
// arithmetic of propagate_row / propagate_row_comp, so the bits are those of
// the unstaged tail.
template <int VE>
__global__ __launch_bounds__(256) void k_propagate_tail_staged(const TailArgs t, const int32_t *__restrict__ node_of,
                                                               const int32_t *__restrict__ parent_of,
                                                               const float *__restrict__ parscale, const CompArgs c,
                                                               int32_t comp_first, float *y, int64_t ldy,
                                                               int32_t ncols, int32_t n_slabs) {
    constexpr int LPR = kTailSlab / VE;
    constexpr int kGroups = 256 / LPR;
    constexpr int32_t kNoParent = -2, kOutside = -1;
    extern __shared__ int4 tail_lds[];
    const int32_t cap = t.staged_rows;
    // per row: x position, y parent's slot (kOutside, kNoParent), z parent's
    // position, w the scale's bits; then the rows, then their compensation rows
    int4 *s_meta = tail_lds;
    float *s_y = reinterpret_cast<float *>(s_meta + cap);
    float *s_c = s_y + int64_t(cap) * kTailSlab;  // (comp_first < INT32_MAX)
    const int32_t f = t.order[int32_t(blockIdx.x) / n_slabs];
    const int32_t slab = int32_t(blockIdx.x) % n_slabs;
    const int l = int(threadIdx.x) % LPR, group = int(threadIdx.x) / LPR;
    const int lc = l * VE;
    const int32_t c0 = slab * kTailSlab + lc;
    const bool active = c0 < ncols;
    const int32_t s0 = t.frag[f], s1 = t.frag[f + 1];
    const int32_t rb = t.seg[s0], n = t.seg[s1] - rb;
    if (n > cap || n < 0) return;  // the host stages only fragments that fit

    for (int32_t i = group; i < n; i += kGroups) {
        const int32_t p = t.pos[rb + i];
        if (l == 0) {
            const bool has = parent_of[p] >= 0;
            s_meta[i] = make_int4(p, has ? kOutside : kNoParent, has ? c.parent_pos[p] : -1,
                                  __float_as_int(parscale ? parscale[p] : 1.0f));
        }
        float v[VE];
        if (active) {
            Load<float, VE>::run(y + int64_t(node_of[p]) * ldy + c0, v);
        } else {
#pragma unroll
            for (int q = 0; q < VE; ++q) v[q] = 0.0f;
        }
#pragma unroll
        for (int q = 0; q < VE; ++q) s_y[i * kTailSlab + lc + q] = v[q];
    }
    __syncthreads();
    for (int32_t i = int32_t(threadIdx.x); i < n; i += 256) {
        const int32_t pp = s_meta[i].z;
        if (s_meta[i].y == kNoParent) continue;
        int32_t lo = 0, hi = i;  // a parent in the fragment precedes its child
        while (lo < hi) {
            const int32_t mid = (lo + hi) >> 1;
            if (s_meta[mid].x < pp) {
                lo = mid + 1;
            } else {
                hi = mid;
            }
        }
        if (lo < i && s_meta[lo].x == pp) s_meta[i].y = lo;
    }
    __syncthreads();

    for (int32_t s = s0; s < s1; ++s) {
        const int32_t ie = t.seg[s + 1] - rb;
        for (int32_t i = t.seg[s] - rb + group; i < ie; i += kGroups) {
            const int4 m = s_meta[i];
            const int32_t p = m.x, ps = m.y, pp = m.z;
            const bool comp_row = p >= comp_first;
            float *yi = s_y + i * kTailSlab + lc;
            if (ps == kNoParent) {
                if (comp_row)
#pragma unroll
                    for (int q = 0; q < VE; ++q) s_c[i * kTailSlab + lc + q] = 0.0f;
                continue;
            }
            const float es = __int_as_float(m.w);
            const bool parent_comp = comp_row && pp >= c.base;  // the parent has a compensation row
            float acc[VE], v[VE], cp[VE];
#pragma unroll
            for (int q = 0; q < VE; ++q) {
                acc[q] = yi[q];
                cp[q] = 0.0f;
            }
            if (ps >= 0) {
#pragma unroll
                for (int q = 0; q < VE; ++q) v[q] = s_y[ps * kTailSlab + lc + q];
                if (parent_comp)
#pragma unroll
                    for (int q = 0; q < VE; ++q) cp[q] = s_c[ps * kTailSlab + lc + q];
            } else if (active) {
                Load<float, VE>::run(y + int64_t(parent_of[p]) * ldy + c0, v);
                if (parent_comp) Load<float, VE>::run(c.comp + int64_t(pp - c.base) * c.ldc + c0, cp);
            } else {
#pragma unroll
                for (int q = 0; q < VE; ++q) v[q] = 0.0f;
            }
            if (comp_row) {
                float cm[VE];
                comp_add<VE>(acc, es, v, cp, cm);
#pragma unroll
                for (int q = 0; q < VE; ++q) s_c[i * kTailSlab + lc + q] = cm[q];
            } else {
#pragma unroll
                for (int q = 0; q < VE; ++q) acc[q] = fmaf(v[q], es, acc[q]);
            }
#pragma unroll
            for (int q = 0; q < VE; ++q) yi[q] = acc[q];
        }
        __syncthreads();
    }

    for (int32_t i = group; active && i < n; i += kGroups) {
        const int4 m = s_meta[i];
        float v[VE];
        if (m.y != kNoParent) {
#pragma unroll
            for (int q = 0; q < VE; ++q) v[q] = s_y[i * kTailSlab + lc + q];
            store_row<VE>(y + int64_t(node_of[m.x]) * ldy + c0, v);
        }
        if (m.x >= comp_first) {
#pragma unroll
            for (int q = 0; q < VE; ++q) v[q] = s_c[i * kTailSlab + lc + q];
            store_row<VE>(c.comp + int64_t(m.x - c.base) * c.ldc + c0, v);
        }
    }
}

__global__ __launch_bounds__(256) void k_branch_update(int32_t n_branches, int32_t col_blocks,
                                                       const int32_t *__restrict__ branches,
                                                       const int32_t *__restrict__ src, const int32_t *__restrict__ dst,
                                                       const float *__restrict__ mult, float *y, int64_t ldy,
                                                       int32_t ncols) {
    const int32_t b = int32_t(blockIdx.x) / col_blocks;
    const int32_t c = (int32_t(blockIdx.x) % col_blocks) * int32_t(blockDim.x) + int32_t(threadIdx.x);
    if (b >= n_branches || c >= ncols) return;
    for (int32_t e = branches[b]; e < branches[b + 1]; ++e) {
        const int32_t s = src[e], d = dst[e];
        const float ys = y[int64_t(s) * ldy + c];
        float *yd = y + int64_t(d) * ldy + c;
        if (mult) {
            const float t = (ys * mult[d]) / mult[s];
            *yd = *yd + t;
        } else {
            *yd = *yd + ys;
        }
    }
}

__global__ __launch_bounds__(256) void k_axpy(int32_t n_rows, int32_t ncols, float coef, const float *__restrict__ x,
                                              int64_t ldx, float *__restrict__ y, int64_t ldy) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= int64_t(n_rows) * ncols) return;
    const int64_t r = i / ncols, c = i % ncols;
    y[r * ldy + c] = fmaf(coef, x[r * ldx + c], y[r * ldy + c]);
}

__global__ __launch_bounds__(256) void k_dad_scales(int32_t n, const int32_t *__restrict__ node_of,
                                                    const int32_t *__restrict__ parent_of, const float *__restrict__ d,
                                                    float *__restrict__ ps) {
    const int32_t p = int32_t(blockIdx.x * blockDim.x + threadIdx.x);
    if (p >= n) return;
    const float dv = d[node_of[p]];
    const int32_t par = parent_of[p];
    const float dp = par >= 0 ? d[par] : 0.0f;
    ps[p] = dp != 0.0f ? dv / dp : 0.0f;
}

// The passes the library runs (anything else is rejected):
//   forward Delta pass:  g = X (FP32/BF16), extra none or parent (FP32 y)
//   operator spmm:       g = rhs (FP32), extra none or self (FP32 rhs)
//   adjoint reverse:     dual g = Z (FP32) / g2 = G, extra self (G)
//   adjoint Delta^T:     dual g = Z (FP32) / g2 = G, extra none
hipError_t dispatch(const RowPass &rp, const Launch &lk, hipStream_t stream) {
    if (!rp.dual) {
        if (!rp.g_bf16) {
            if (rp.extra == kExtraNone) return dispatch_ve<Pass<float, float, float, kExtraNone, false>>(rp, lk, stream);
            if (rp.e_bf16) return hipErrorInvalidValue;
            if (rp.extra == kExtraParent)
                return dispatch_ve<Pass<float, float, float, kExtraParent, false>>(rp, lk, stream);
            if (rp.extra == kExtraParentKahan)
                return dispatch_ve<Pass<float, float, float, kExtraParentKahan, false>>(rp, lk, stream);
            return dispatch_ve<Pass<float, float, float, kExtraSelf, false>>(rp, lk, stream);
        }
        if (rp.extra == kExtraNone) return dispatch_ve<Pass<bf16_t, bf16_t, float, kExtraNone, false>>(rp, lk, stream);
        if (rp.extra == kExtraParent && !rp.e_bf16)
            return dispatch_ve<Pass<bf16_t, bf16_t, float, kExtraParent, false>>(rp, lk, stream);
        if (rp.extra == kExtraParentKahan && !rp.e_bf16)
            return dispatch_ve<Pass<bf16_t, bf16_t, float, kExtraParentKahan, false>>(rp, lk, stream);
        return hipErrorInvalidValue;
    }
    if (rp.g_bf16) return hipErrorInvalidValue;  // the dual pass gathers FP32 Z
    if (rp.extra == kExtraSelf) {
        if (!rp.g2_bf16 && !rp.e_bf16) return dispatch_ve<Pass<float, float, float, kExtraSelf, true>>(rp, lk, stream);
        if (rp.g2_bf16 && rp.e_bf16) return dispatch_ve<Pass<float, bf16_t, bf16_t, kExtraSelf, true>>(rp, lk, stream);
        return hipErrorInvalidValue;
    }
    if (rp.extra == kExtraNone) {
        if (!rp.g2_bf16) return dispatch_ve<Pass<float, float, float, kExtraNone, true>>(rp, lk, stream);
        return dispatch_ve<Pass<float, bf16_t, float, kExtraNone, true>>(rp, lk, stream);
    }
    return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_row_pass(const RowPass &rp, hipStream_t stream) {
    return dispatch(rp, Launch{Kind::kPass, nullptr, nullptr}, stream);
}

hipError_t launch_row_chunks(const RowPass &rp, const ChunkArgs &c, hipStream_t stream) {
    return dispatch(rp, Launch{Kind::kChunk, &c, nullptr}, stream);
}

hipError_t launch_row_mixed(const RowPass &rp, const ChunkArgs &c, int *launched, hipStream_t stream) {
    Launch lk{Kind::kMixed, &c, nullptr};
    lk.launched = launched;
    return dispatch(rp, lk, stream);
}

hipError_t launch_row_reduce(const RowPass &rp, const ReduceArgs &r, hipStream_t stream) {
    return dispatch(rp, Launch{Kind::kReduce, nullptr, &r}, stream);
}

hipError_t launch_row_tail(const RowPass &rp, const TailArgs &t, hipStream_t stream) {
    RowPass r = rp;
    r.xcd_slabs = 1;  // one workgroup covers every column of its rows
    return dispatch(r, Launch{Kind::kTail, nullptr, nullptr, &t}, stream);
}

hipError_t launch_propagate_tail(const TailArgs &t, const int32_t *node_of, const int32_t *parent_of,
                                 const float *parscale, const CompArgs &c, int32_t comp_first, float *y, int64_t ldy,
                                 int32_t ncols, hipStream_t stream) {
    if (t.n_frag <= 0) return hipSuccess;
    const bool comp = comp_first < INT32_MAX;
    const bool vec = (ncols % 4) == 0 && aligned16(y, ldy, 4) && (!comp || aligned16(c.comp, c.ldc, 4));
    const int ve = vec ? 4 : 1;
    const int32_t n_staged = t.order ? t.n_staged : 0;
    if (n_staged < 0 || n_staged > t.n_frag) return hipErrorInvalidValue;
    if (n_staged > 0) {
        const int64_t lds = tail_staged_bytes(t.staged_rows, comp);
        const int32_t n_slabs = (ncols + kTailSlab - 1) / kTailSlab;
        if (t.staged_rows <= 0 || lds > kTailStagedMaxBytes || int64_t(n_staged) * n_slabs > int64_t(INT32_MAX))
            return hipErrorInvalidValue;
        const dim3 grid(static_cast<unsigned>(int64_t(n_staged) * n_slabs)), block(256);
        if (vec) {
            k_propagate_tail_staged<4><<<grid, block, size_t(lds), stream>>>(t, node_of, parent_of, parscale, c,
                                                                             comp_first, y, ldy, ncols, n_slabs);
        } else {
            k_propagate_tail_staged<1><<<grid, block, size_t(lds), stream>>>(t, node_of, parent_of, parscale, c,
                                                                             comp_first, y, ldy, ncols, n_slabs);
        }
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess || n_staged == t.n_frag) return e;
    }
    const unsigned rest = unsigned(t.n_frag - n_staged);
    const int lpr = lanes_per_row(ncols, ve);
#define CBM_TAIL_CASE(V, L)                                                                                          \
    if (ve == V && lpr == L) {                                                                                       \
        k_propagate_tail<V, L><<<dim3(rest), dim3(256), 0, stream>>>(t, node_of, parent_of, parscale, c, comp_first, \
                                                                     y, ldy, ncols);                                 \
        return hipGetLastError();                                                                                    \
    }
    CBM_TAIL_CASE(4, 1) CBM_TAIL_CASE(4, 2) CBM_TAIL_CASE(4, 4) CBM_TAIL_CASE(4, 8) CBM_TAIL_CASE(4, 16)
    CBM_TAIL_CASE(4, 32) CBM_TAIL_CASE(4, 64)
    CBM_TAIL_CASE(1, 1) CBM_TAIL_CASE(1, 2) CBM_TAIL_CASE(1, 4) CBM_TAIL_CASE(1, 8) CBM_TAIL_CASE(1, 16)
    CBM_TAIL_CASE(1, 32) CBM_TAIL_CASE(1, 64)
#undef CBM_TAIL_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_propagate(int32_t p_begin, int32_t p_end, const int32_t *node_of, const int32_t *parent_of,
                            const float *parscale, float *y, int64_t ldy, int32_t ncols, hipStream_t stream) {
    const int64_t n = int64_t(p_end) - p_begin;
    if (n <= 0) return hipSuccess;
    const bool vec = (ncols % 4) == 0 && aligned16(y, ldy, 4);
    const int ve = vec ? 4 : 1;
    const int lpr = lanes_per_row(ncols, ve);
    const int64_t rows_per_block = 4 * (64 / lpr);
    const dim3 grid(unsigned((n + rows_per_block - 1) / rows_per_block)), block(256);
#define CBM_PROP_CASE(V, L)                                                                                  \
    if (ve == V && lpr == L) {                                                                               \
        k_propagate<V, L><<<grid, block, 0, stream>>>(p_begin, p_end, node_of, parent_of, parscale, y, ldy, \
                                                      ncols);                                                \
        return hipGetLastError();                                                                            \
    }
    CBM_PROP_CASE(4, 1) CBM_PROP_CASE(4, 2) CBM_PROP_CASE(4, 4) CBM_PROP_CASE(4, 8) CBM_PROP_CASE(4, 16)
    CBM_PROP_CASE(4, 32) CBM_PROP_CASE(4, 64)
    CBM_PROP_CASE(1, 1) CBM_PROP_CASE(1, 2) CBM_PROP_CASE(1, 4) CBM_PROP_CASE(1, 8) CBM_PROP_CASE(1, 16)
    CBM_PROP_CASE(1, 32) CBM_PROP_CASE(1, 64)
#undef CBM_PROP_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_propagate_comp(int32_t p_begin, int32_t p_end, const int32_t *node_of, const int32_t *parent_of,
                                 const float *parscale, const CompArgs &c, float *y, int64_t ldy, int32_t ncols,
                                 hipStream_t stream) {
    const int64_t n = int64_t(p_end) - p_begin;
    if (n <= 0) return hipSuccess;
    if (p_begin < c.base) return hipErrorInvalidValue;  // every position needs its compensation row
    const bool vec = (ncols % 4) == 0 && aligned16(y, ldy, 4) && aligned16(c.comp, c.ldc, 4);
    const int ve = vec ? 4 : 1;
    const int lpr = lanes_per_row(ncols, ve);
    const int64_t rows_per_block = 4 * (64 / lpr);
    const dim3 grid(unsigned((n + rows_per_block - 1) / rows_per_block)), block(256);
#define CBM_PROPC_CASE(V, L)                                                                                     \
    if (ve == V && lpr == L) {                                                                                   \
        k_propagate_comp<V, L><<<grid, block, 0, stream>>>(p_begin, p_end, node_of, parent_of, parscale, c, y, \
                                                           ldy, ncols);                                          \
        return hipGetLastError();                                                                                \
    }
    CBM_PROPC_CASE(4, 1) CBM_PROPC_CASE(4, 2) CBM_PROPC_CASE(4, 4) CBM_PROPC_CASE(4, 8) CBM_PROPC_CASE(4, 16)
    CBM_PROPC_CASE(4, 32) CBM_PROPC_CASE(4, 64)
    CBM_PROPC_CASE(1, 1) CBM_PROPC_CASE(1, 2) CBM_PROPC_CASE(1, 4) CBM_PROPC_CASE(1, 8) CBM_PROPC_CASE(1, 16)
    CBM_PROPC_CASE(1, 32) CBM_PROPC_CASE(1, 64)
#undef CBM_PROPC_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_branch_update(int32_t n_branches, const int32_t *branches, const int32_t *src, const int32_t *dst,
                                const float *mult, float *y, int64_t ldy, int32_t ncols, hipStream_t stream) {
    if (n_branches <= 0 || ncols <= 0) return hipSuccess;
    const int32_t col_blocks = (ncols + 255) / 256;
    k_branch_update<<<dim3(unsigned(int64_t(n_branches) * col_blocks)), dim3(256), 0, stream>>>(
        n_branches, col_blocks, branches, src, dst, mult, y, ldy, ncols);
    return hipGetLastError();
}

hipError_t launch_axpy(int32_t n_rows, int32_t ncols, float coef, const float *x, int64_t ldx, float *y, int64_t ldy,
                       hipStream_t stream) {
    const int64_t total = int64_t(n_rows) * ncols;
    if (total <= 0) return hipSuccess;
    k_axpy<<<dim3(unsigned((total + 255) / 256)), dim3(256), 0, stream>>>(n_rows, ncols, coef, x, ldx, y, ldy);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_order_meta(int32_t n, const int32_t *__restrict__ order, int32_t n_rows,
                                                    const int32_t *__restrict__ ptr,
                                                    const int32_t *__restrict__ node_of, int4 *__restrict__ meta) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t p = order[i];
    if (p < 0 || p >= n_rows) {
        meta[i] = make_int4(-1, 0, 0, 0);
        return;
    }
    meta[i] = make_int4(p, ptr[p], ptr[p + 1], node_of ? node_of[p] : p);
}

hipError_t launch_order_meta(int32_t n, const int32_t *order, int32_t n_rows, const int32_t *ptr,
                             const int32_t *node_of, int32_t *meta, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    k_order_meta<<<dim3(unsigned((n + 255) / 256)), dim3(256), 0, stream>>>(n, order, n_rows, ptr, node_of,
                                                                            reinterpret_cast<int4 *>(meta));
    return hipGetLastError();
}

hipError_t launch_dad_scales(int32_t n, const int32_t *node_of, const int32_t *parent_of, const float *d, float *ps,
                             hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    k_dad_scales<<<dim3(unsigned((n + 255) / 256)), dim3(256), 0, stream>>>(n, node_of, parent_of, d, ps);
    return hipGetLastError();
}

}  // namespace cbmk