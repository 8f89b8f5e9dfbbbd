// Row-pass templates of the CBM product (gfx950, wave64): the device
// building blocks, the row-pass kernels and their launchers.  Included by
// cbm_kernels.hip (the dispatch over pass shapes) and by the rp_*.hip files,
// each of which instantiates the kernels of a few pass shapes, so that the
// shapes compile in parallel.  See cbm_kernels.hip for the mapping.
#pragma once

#include "cbm_kernels.hpp"

#include <algorithm>

namespace cbmk {
namespace rp {

using bf16_t = uint16_t;  // storage type of bfloat16 values

constexpr uint32_t kSign = 0x80000000u;
constexpr int32_t kIndexMask = 0x7fffffff;

__device__ inline float bf16_to_f32(uint32_t bits16) { return __uint_as_float(bits16 << 16); }

// Loads VE consecutive values starting at p as FP32.
template <typename T, int VE>
struct Load;

template <>
struct Load<float, 1> {
    __device__ static void run(const float *p, float (&v)[1]) { v[0] = *p; }
};
template <>
struct Load<float, 4> {
    __device__ static void run(const float *p, float (&v)[4]) {
        const float4 t = *reinterpret_cast<const float4 *>(p);
        v[0] = t.x;
        v[1] = t.y;
        v[2] = t.z;
        v[3] = t.w;
    }
};
template <>
struct Load<float, 8> {
    __device__ static void run(const float *p, float (&v)[8]) {
        const float4 a = reinterpret_cast<const float4 *>(p)[0];
        const float4 b = reinterpret_cast<const float4 *>(p)[1];
        v[0] = a.x;
        v[1] = a.y;
        v[2] = a.z;
        v[3] = a.w;
        v[4] = b.x;
        v[5] = b.y;
        v[6] = b.z;
        v[7] = b.w;
    }
};
template <>
struct Load<bf16_t, 1> {
    __device__ static void run(const bf16_t *p, float (&v)[1]) { v[0] = bf16_to_f32(*p); }
};
template <>
struct Load<bf16_t, 4> {
    __device__ static void run(const bf16_t *p, float (&v)[4]) {
        const uint2 t = *reinterpret_cast<const uint2 *>(p);
        v[0] = bf16_to_f32(t.x & 0xffffu);
        v[1] = bf16_to_f32(t.x >> 16);
        v[2] = bf16_to_f32(t.y & 0xffffu);
        v[3] = bf16_to_f32(t.y >> 16);
    }
};
template <>
struct Load<bf16_t, 8> {
    __device__ static void run(const bf16_t *p, float (&v)[8]) {
        const uint4 t = *reinterpret_cast<const uint4 *>(p);
        v[0] = bf16_to_f32(t.x & 0xffffu);
        v[1] = bf16_to_f32(t.x >> 16);
        v[2] = bf16_to_f32(t.y & 0xffffu);
        v[3] = bf16_to_f32(t.y >> 16);
        v[4] = bf16_to_f32(t.z & 0xffffu);
        v[5] = bf16_to_f32(t.z >> 16);
        v[6] = bf16_to_f32(t.w & 0xffffu);
        v[7] = bf16_to_f32(t.w >> 16);
    }
};

template <int VE>
__device__ inline void store_row(float *p, const float (&v)[VE]) {
    if constexpr (VE == 1) {
        *p = v[0];
    } else {
#pragma unroll
        for (int q = 0; q < VE; q += 4)
            *reinterpret_cast<float4 *>(p + q) = make_float4(v[q], v[q + 1], v[q + 2], v[q + 3]);
    }
}

// Value of lane k of the calling lane's group of LPR lanes.
template <int LPR>
__device__ inline int32_t group_bcast(int32_t v, int32_t k) {
    if constexpr (LPR == 1) {
        return v;
    } else if constexpr (LPR == 64) {
        return __builtin_amdgcn_readlane(v, k);
    } else {
        return __shfl(v, k, LPR);
    }
}

template <int LPR>
__device__ inline float group_bcast(float v, int32_t k) {
    return __int_as_float(group_bcast<LPR>(__float_as_int(v), k));
}

// +-scale: the packed index's bit 31 flips the sign of the scale.
__device__ inline float signed_scale(float scale, int32_t packed) {
    return __uint_as_float(__float_as_uint(scale) ^ (uint32_t(packed) & kSign));
}

// Compile-time shape of a row pass.
//   G   type of the gathered matrix (g)
//   G2  type of the second gathered matrix (g2), read when DUAL and the
//       packed index lacks the internal bit
//   E   type of the extra-term matrix (e)
template <typename G, typename G2, typename E, int EXTRA, bool DUAL>
struct Pass {
    using TG = G;
    using TG2 = G2;
    using TE = E;
    static constexpr int kExtra = EXTRA;
    static constexpr bool kDual = DUAL;
};

// acc += sum over k in [kb, ke) of sign_k * scale[j_k] * val_k * src[j_k, c0 .. c0+VE),
// by the LPR lanes of one group (all lanes of the group call it with the same
// kb, ke; lanes whose columns are past the end pass active = false but still
// take part in the index broadcasts).  scale and val may be null (1).
// SCALED: whether scale or val is given; without them an entry's factor is
// +-1 from the sign bit of the broadcast index itself, so only the index
// crosses lanes.
template <typename TS, int VE, int LPR, bool SCALED>
__device__ inline void gather_src_impl(const RowPass &a, const TS *__restrict__ src, int64_t ld,
                                       const float *__restrict__ scale, const float *__restrict__ val, int32_t kb,
                                       int32_t ke, int l, int32_t c0, bool active, float (&acc)[VE]) {
    if constexpr (LPR < 4) {
        // Groups narrower than four lanes fetch four entries per step, 4 / LPR
        // per lane (entry base + m * LPR + l), so that four gathers are still
        // in flight per lane.
        constexpr int M = 4 / LPR;
        for (int32_t base = kb; base < ke; base += 4) {
            int32_t mine[M];
            float mine_scale[M];
#pragma unroll
            for (int m = 0; m < M; ++m) {
                const int32_t e = base + m * LPR + l;
                mine[m] = 0;
                mine_scale[m] = 0.0f;
                if (e < ke) {
                    mine[m] = a.idx[e];
                    if constexpr (SCALED) {
                        float cs = scale ? scale[mine[m] & kIndexMask] : 1.0f;
                        if (val) cs *= val[e];
                        mine_scale[m] = signed_scale(cs, mine[m]);
                    }
                }
            }
            const int32_t cnt = min(4, ke - base);
            int32_t j[4];
            float sc[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int32_t raw = group_bcast<LPR>(mine[u / LPR], u % LPR);
                j[u] = raw & kIndexMask;
                if constexpr (SCALED) {
                    sc[u] = group_bcast<LPR>(mine_scale[u / LPR], u % LPR);
                } else {
                    sc[u] = signed_scale(1.0f, raw);
                }
            }
            if (active) {
                float v[4][VE];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    if (u < cnt) {
                        Load<TS, VE>::run(src + int64_t(j[u]) * ld + c0, v[u]);
                    } else {
#pragma unroll
                        for (int q = 0; q < VE; ++q) v[u][q] = 0.0f;
                    }
                }
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int q = 0; q < VE; ++q) acc[q] = fmaf(v[u][q], sc[u], acc[q]);
            }
        }
        return;
    }
    for (int32_t base = kb; base < ke; base += LPR) {
        int32_t mine = 0;
        float mine_scale = 0.0f;
        if (base + l < ke) {
            mine = a.idx[base + l];
            if constexpr (SCALED) {
                float cs = scale ? scale[mine & kIndexMask] : 1.0f;
                if (val) cs *= val[base + l];
                mine_scale = signed_scale(cs, mine);
            }
        }
        const int32_t cnt = min(LPR, ke - base);
        int32_t k = 0;
        for (; k + 4 <= cnt; k += 4) {
            int32_t j[4];
            float sc[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int32_t raw = group_bcast<LPR>(mine, k + u);
                j[u] = raw & kIndexMask;
                if constexpr (SCALED) {
                    sc[u] = group_bcast<LPR>(mine_scale, k + u);
                } else {
                    sc[u] = signed_scale(1.0f, raw);
                }
            }
            if (active) {
                float v[4][VE];
#pragma unroll
                for (int u = 0; u < 4; ++u) Load<TS, VE>::run(src + int64_t(j[u]) * ld + c0, v[u]);
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int q = 0; q < VE; ++q) acc[q] = fmaf(v[u][q], sc[u], acc[q]);
            }
        }
        if (k < cnt) {
            // The last one to three entries in one step, so that their
            // gathers are in flight together: the slots past the row load
            // nothing and add v = 0, which leaves acc's bits unchanged (acc
            // starts at +0 and a sum of fma results is never -0).
            int32_t j[4];
            float sc[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int32_t src_lane = min(k + u, LPR - 1);
                const int32_t raw = group_bcast<LPR>(mine, src_lane);
                j[u] = raw & kIndexMask;
                if constexpr (SCALED) {
                    sc[u] = group_bcast<LPR>(mine_scale, src_lane);
                } else {
                    sc[u] = signed_scale(1.0f, raw);
                }
            }
            if (active) {
                float v[4][VE];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    if (k + u < cnt) {
                        Load<TS, VE>::run(src + int64_t(j[u]) * ld + c0, v[u]);
                    } else {
#pragma unroll
                        for (int q = 0; q < VE; ++q) v[u][q] = 0.0f;
                    }
                }
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int q = 0; q < VE; ++q) acc[q] = fmaf(v[u][q], sc[u], acc[q]);
            }
        }
    }
}

template <typename TS, int VE, int LPR>
__device__ inline void gather_src(const RowPass &a, const TS *__restrict__ src, int64_t ld,
                                  const float *__restrict__ scale, const float *__restrict__ val, int32_t kb,
                                  int32_t ke, int l, int32_t c0, bool active, float (&acc)[VE]) {
    if (scale || val) {
        gather_src_impl<TS, VE, LPR, true>(a, src, ld, scale, val, kb, ke, l, c0, active, acc);
    } else {
        gather_src_impl<TS, VE, LPR, false>(a, src, ld, scale, val, kb, ke, l, c0, active, acc);
    }
}

// The gather of position p's entries in [kb, ke).  A dual pass reads the
// entries before mid[p] from g2 (scaled by g2scale) and the rest from g, as
// two branch-free loops.
template <class P, int VE, int LPR>
__device__ inline void gather_sum(const RowPass &a, int32_t p, int32_t kb, int32_t ke, int l, int32_t c0, bool active,
                                  float (&acc)[VE]) {
    using TG = typename P::TG;
    if constexpr (P::kDual) {
        using TG2 = typename P::TG2;
        const int32_t m = min(max(a.mid[p], kb), ke);
        gather_src<TG2, VE, LPR>(a, static_cast<const TG2 *>(a.g2), a.ldg2, a.g2scale, nullptr, kb, m, l, c0, active,
                                 acc);
        gather_src<TG, VE, LPR>(a, static_cast<const TG *>(a.g), a.ldg, nullptr, nullptr, m, ke, l, c0, active, acc);
    } else {
        gather_src<TG, VE, LPR>(a, static_cast<const TG *>(a.g), a.ldg, a.colscale, a.val, kb, ke, l, c0, active, acc);
    }
}

// acc += es * (v + cp) as the pair (acc, comp) with acc the rounded sum and
// comp its error: TwoProduct for es * v, TwoSum for the addition (Knuth),
// Fast2Sum to fold the error back.  Every operation is an explicitly rounded
// one, so contraction into fma cannot change the error terms.
template <int VE>
__device__ inline void comp_add(float (&acc)[VE], float es, const float (&v)[VE], const float (&cp)[VE],
                                float (&comp)[VE]) {
#pragma unroll
    for (int q = 0; q < VE; ++q) {
        const float t = __fmul_rn(es, v[q]);
        const float t_err = fmaf(es, v[q], -t);
        const float s = __fadd_rn(acc[q], t);
        const float bb = __fsub_rn(s, acc[q]);
        const float e = __fadd_rn(__fsub_rn(acc[q], __fsub_rn(s, bb)), __fsub_rn(t, bb));
        const float err = __fadd_rn(e, __fadd_rn(t_err, __fmul_rn(es, cp[q])));
        const float yq = __fadd_rn(s, err);
        comp[q] = __fsub_rn(err, __fsub_rn(yq, s));
        acc[q] = yq;
    }
}

// Row scale, the extra term, and the store of one finished row segment.
template <class P, int VE>
__device__ inline void finish_row(const RowPass &a, int32_t p, int32_t row, int32_t c0, float (&acc)[VE]) {
    const float rs = a.rowscale ? a.rowscale[row] : 1.0f;
#pragma unroll
    for (int q = 0; q < VE; ++q) acc[q] *= rs;
    using TE = typename P::TE;
    if constexpr (P::kExtra == kExtraParent) {
        const int32_t par = a.parent_of[p];
        if (par >= 0) {
            const float es = (a.escale ? a.escale[p] : 1.0f) * a.ecoef;
            float v[VE];
            Load<TE, VE>::run(static_cast<const TE *>(a.e) + int64_t(par) * a.lde + c0, v);
#pragma unroll
            for (int q = 0; q < VE; ++q) acc[q] = fmaf(v[q], es, acc[q]);
        }
    } else if constexpr (P::kExtra == kExtraSelf) {
        const float es = (a.escale ? a.escale[row] : 1.0f) * a.ecoef;
        float v[VE];
        Load<TE, VE>::run(static_cast<const TE *>(a.e) + int64_t(row) * a.lde + c0, v);
#pragma unroll
        for (int q = 0; q < VE; ++q) acc[q] = fmaf(v[q], es, acc[q]);
    } else if constexpr (P::kExtra == kExtraParentKahan) {
        // y = acc + es * y_par + es * c_par, kept as the pair (y, c) with
        // y the rounded sum and c its error: TwoProduct for es * y_par,
        // TwoSum for the addition (Knuth), Fast2Sum to fold the error back.
        float comp[VE];
#pragma unroll
        for (int q = 0; q < VE; ++q) comp[q] = 0.0f;
        const int32_t par = a.parent_of[p];
        if (par >= 0) {
            const float es = (a.escale ? a.escale[p] : 1.0f) * a.ecoef;
            float v[VE], cp[VE];
            Load<TE, VE>::run(static_cast<const TE *>(a.e) + int64_t(par) * a.lde + c0, v);
            const int32_t pp = a.parent_pos[p];
            if (pp >= a.comp_base) {
                Load<float, VE>::run(a.comp + int64_t(pp - a.comp_base) * a.ldc + c0, cp);
            } else {
#pragma unroll
                for (int q = 0; q < VE; ++q) cp[q] = 0.0f;
            }
            comp_add<VE>(acc, es, v, cp, comp);
        }
        store_row<VE>(a.comp + int64_t(p - a.comp_base) * a.ldc + c0, comp);
    }
    store_row<VE>(a.y + int64_t(row) * a.ldy + c0, acc);
}

// Short rows: one group of LPR lanes per row, 64 / LPR rows per wavefront
// (S = 1); or S = 64 / LPR sub-groups of LPR lanes per row, one row per
// wavefront, each sub-group summing a contiguous slice of the row's entries
// and the slices' sums added by a butterfly across the sub-groups (the same
// order on every lane, so every sub-group holds the same bits).
template <class P, int VE, int LPR, int S>
__device__ __forceinline__ void row_pass_block(const RowPass &a, uint32_t bid) {
    constexpr int kLanesPerRow = LPR * S;
    constexpr int kRowsPerWave = 64 / kLanesPerRow;
    const int lane = int(threadIdx.x & 63u);
    const int l = lane & (LPR - 1);
    const int sub = (lane / LPR) & (S - 1);
    // Workgroup b: slab b % xcd_slabs of row block b / xcd_slabs, or, with
    // XCD ranges, the range, row block and slab of the mapping in RowPass.
    int32_t slab, p_begin = a.p_begin, p_end = a.p_end;
    int64_t blk;
    if (a.ranges) {
        const uint32_t x = bid & 7u, per = 8u / uint32_t(a.range_count);
        const uint32_t k = x / per;
        const int64_t b8 = int64_t(bid >> 3);
        const int64_t phase = b8 / a.range_blocks;
        blk = b8 - phase * a.range_blocks;
        slab = int32_t(int64_t(x % per) * a.range_phases + phase);
        // Constant indices only: a runtime index into the argument's array
        // would copy the whole argument to scratch memory.
#pragma unroll
        for (uint32_t i = 0; i < 8; ++i) {
            if (i == k) {
                p_begin = a.range_first[i];
                p_end = a.range_first[i + 1];
            }
        }
    } else if (a.order) {
        const uint32_t x = bid & 7u, nb = uint32_t(a.order_blocks), base = nb >> 3, rem = nb & 7u;
        slab = 0;
        blk = int64_t(x) * base + min(x, rem) + int64_t(bid >> 3);
    } else {
        slab = int32_t(bid % uint32_t(a.xcd_slabs));
        blk = int64_t(bid / uint32_t(a.xcd_slabs));
    }
    const int64_t group = (blk * int64_t(blockDim.x >> 6) + int64_t(threadIdx.x >> 6)) * kRowsPerWave +
                          lane / kLanesPerRow;
    const int64_t p64 = int64_t(p_begin) + group;
    if (p64 >= int64_t(p_end)) return;  // a row's lanes leave together
    int32_t p, kb, ke, row;
    if (a.order_meta) {
        const int4 mt = reinterpret_cast<const int4 *>(a.order_meta)[p64];
        p = mt.x;
        kb = mt.y;
        ke = mt.z;
        row = mt.w;
        if (p < 0) return;
    } else {
        p = a.order ? a.order[p64] : int32_t(p64);
        if (p < 0) return;
        kb = a.ptr[p];
        ke = a.ptr_end ? a.ptr_end[p] : a.ptr[p + 1];
        row = a.node_of ? a.node_of[p] : p;
    }
    if (ke - kb > a.skip_above) return;  // a long row, done by the chunk kernel
    const int64_t len = int64_t(ke) - kb;
    const int32_t sb = kb + int32_t(len * sub / S), se = kb + int32_t(len * (sub + 1) / S);
    const int32_t cb = slab * a.slab_w, ce = min(cb + a.slab_w, a.ncols);
    constexpr int32_t kSpan = LPR * VE;
    const int32_t passes = (ce - cb + kSpan - 1) / kSpan;
    for (int32_t pass = 0; pass < passes; ++pass) {
        const int32_t c0 = cb + pass * kSpan + l * VE;
        const bool active = c0 < ce;
        float acc[VE];
#pragma unroll
        for (int q = 0; q < VE; ++q) acc[q] = 0.0f;
        gather_sum<P, VE, LPR>(a, p, sb, se, l, c0, active, acc);
        if constexpr (S > 1) {
#pragma unroll
            for (int off = LPR; off < kLanesPerRow; off <<= 1)
#pragma unroll
                for (int q = 0; q < VE; ++q) acc[q] += __shfl_xor(acc[q], off);
        }
        if (active && sub == 0) finish_row<P, VE>(a, p, row, c0, acc);
    }
}

template <class P, int VE, int LPR, int S>
__global__ __launch_bounds__(256) void k_row_pass(const RowPass a) {
    row_pass_block<P, VE, LPR, S>(a, blockIdx.x);
}

// A row pass over a tree tail: workgroup f finishes fragment f's segments,
// deepest first, one group of LPR lanes per row and a workgroup barrier
// between segments (the rows of one segment read those of the next deeper
// one, written by this workgroup).  Only the reverse pass runs here: each
// row's own term s_u G_u is loaded before its gathers, since it depends on
// nothing they produce and, loaded after them as finish_row does, it would
// be one more dependent wait on every level of the fragment; the arithmetic
// is finish_row's, so the bits are too.
template <class P, int VE, int LPR>
__global__ __launch_bounds__(256) void k_row_tail(const RowPass a, const TailArgs t) {
    static_assert(P::kDual && P::kExtra == kExtraSelf, "the reverse pass");
    using TE = typename P::TE;
    constexpr int kGroups = 256 / LPR;
    const int lane = int(threadIdx.x & 63u);
    const int l = lane & (LPR - 1);
    const int32_t group = int32_t(threadIdx.x / LPR);
    const int32_t f = int32_t(blockIdx.x);
    constexpr int32_t kSpan = LPR * VE;
    const int32_t passes = (a.ncols + kSpan - 1) / kSpan;
    for (int32_t s = t.frag[f + 1] - 1; s >= t.frag[f]; --s) {
        for (int32_t i = t.seg[s] + group; i < t.seg[s + 1]; i += kGroups) {
            const int32_t p = t.pos[i];
            const int32_t kb = a.ptr[p], ke = a.ptr_end ? a.ptr_end[p] : a.ptr[p + 1];
            const int32_t row = a.node_of ? a.node_of[p] : p;
            const float rs = a.rowscale ? a.rowscale[row] : 1.0f;
            const float es = (a.escale ? a.escale[row] : 1.0f) * a.ecoef;
            for (int32_t pass = 0; pass < passes; ++pass) {
                const int32_t c0 = pass * kSpan + l * VE;
                const bool active = c0 < a.ncols;
                // Lanes past the last column read column 0 (no branch around
                // the load) and store nothing.
                float ev[VE];
                Load<TE, VE>::run(static_cast<const TE *>(a.e) + int64_t(row) * a.lde + (active ? c0 : 0), ev);
                float acc[VE];
#pragma unroll
                for (int q = 0; q < VE; ++q) acc[q] = 0.0f;
                gather_sum<P, VE, LPR>(a, p, kb, ke, l, c0, active, acc);
                if (active) {
#pragma unroll
                    for (int q = 0; q < VE; ++q) acc[q] *= rs;
#pragma unroll
                    for (int q = 0; q < VE; ++q) acc[q] = fmaf(ev[q], es, acc[q]);
                    store_row<VE>(a.y + int64_t(row) * a.ldy + c0, acc);
                }
            }
        }
        __syncthreads();
    }
}

// Row p (output row `row`) is the sum of scratch rows s0 .. s0 + ns - 1 (ld =
// ncols), added in slot order (0 when ns is 0), then finished; by the LPR
// lanes of one group, every column.
template <class P, int VE, int LPR>
__device__ inline void sum_slots_finish(const RowPass &a, const float *__restrict__ scratch, int32_t p, int32_t row,
                                        int32_t s0, int32_t ns, int l) {
    for (int32_t c0 = l * VE; c0 < a.ncols; c0 += LPR * VE) {
        float acc[VE];
        if (ns > 0) {
            Load<float, VE>::run(scratch + int64_t(s0) * a.ncols + c0, acc);
        } else {
#pragma unroll
            for (int q = 0; q < VE; ++q) acc[q] = 0.0f;
        }
        int32_t s = 1;
        // Four slot loads in flight, added in slot order (the same bits as one
        // at a time).
        for (; s + 4 <= ns; s += 4) {
            float v[4][VE];
#pragma unroll
            for (int u = 0; u < 4; ++u) Load<float, VE>::run(scratch + int64_t(s0 + s + u) * a.ncols + c0, v[u]);
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int q = 0; q < VE; ++q) acc[q] += v[u][q];
        }
        for (; s < ns; ++s) {
            float v[VE];
            Load<float, VE>::run(scratch + int64_t(s0 + s) * a.ncols + c0, v);
#pragma unroll
            for (int q = 0; q < VE; ++q) acc[q] += v[q];
        }
        finish_row<P, VE>(a, p, row, c0, acc);
    }
}

// Long rows: one workgroup of BS threads per chunk of a row.  The chunk's
// entries are split into contiguous slices, one per group of LPR lanes; the
// groups' partial sums are added in group order through LDS.  A one-chunk
// row is finished here; otherwise the chunk's sum goes to its scratch slot.
template <class P, int VE, int LPR, int BS>
__device__ __forceinline__ void row_chunk_block(const RowPass &a, const ChunkArgs &c, uint32_t bid) {
    constexpr int kGroups = BS / LPR;
    __shared__ float part[BS * VE];
    const int32_t slab = int32_t(bid % uint32_t(a.xcd_slabs));
    const int32_t chunk = c.first + int32_t(bid / uint32_t(a.xcd_slabs));
    const int32_t p = c.pos[chunk];
    const uint32_t meta = uint32_t(c.meta[chunk]);
    const int64_t idx = meta & 0xFFFFu, cnt = meta >> 16;
    const int32_t kb = a.ptr[p], ke = a.ptr_end ? a.ptr_end[p] : a.ptr[p + 1];
    const int64_t len = int64_t(ke) - kb;
    const int64_t ck0 = kb + len * idx / cnt, ck1 = kb + len * (idx + 1) / cnt;
    const int g = int(threadIdx.x) / LPR, l = int(threadIdx.x) % LPR;
    const int32_t gb = int32_t(ck0 + (ck1 - ck0) * g / kGroups);
    const int32_t ge = int32_t(ck0 + (ck1 - ck0) * (g + 1) / kGroups);
    const int32_t row = a.node_of ? a.node_of[p] : p;
    const int32_t slot = c.slot[chunk];
    const int32_t cb = slab * a.slab_w, ce = min(cb + a.slab_w, a.ncols);
    constexpr int32_t kSpan = LPR * VE;
    const int32_t passes = (ce - cb + kSpan - 1) / kSpan;
    for (int32_t pass = 0; pass < passes; ++pass) {
        const int32_t c0 = cb + pass * kSpan + l * VE;
        const bool active = c0 < ce;
        float acc[VE];
#pragma unroll
        for (int q = 0; q < VE; ++q) acc[q] = 0.0f;
        gather_sum<P, VE, LPR>(a, p, gb, ge, l, c0, active, acc);
#pragma unroll
        for (int q = 0; q < VE; ++q) part[threadIdx.x * VE + q] = acc[q];
        __syncthreads();
        if (g == 0 && active) {
            for (int gg = 1; gg < kGroups; ++gg)
#pragma unroll
                for (int q = 0; q < VE; ++q) acc[q] += part[(gg * LPR + l) * VE + q];
            if (slot < 0) {
                finish_row<P, VE>(a, p, row, c0, acc);
            } else {
                store_row<VE>(c.scratch + int64_t(slot) * a.ncols + c0, acc);
            }
        }
        __syncthreads();
    }
    if (c.arrive != nullptr && slot >= 0 && threadIdx.x < 64) {
        // Wave 0 holds group 0, which stored this chunk's sums: it releases
        // them at device scope and counts the chunk in; the row's last chunk
        // to arrive acquires the others' sums, adds all of them in slot order
        // and finishes the row, then clears the counter for the next launch
        // (no other chunk of this launch touches it after its own arrival).
        const int32_t first = slot - int32_t(idx);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        uint32_t before = 0;
        if (threadIdx.x == 0)
            before = __hip_atomic_fetch_add(c.arrive + first, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        before = uint32_t(__builtin_amdgcn_readfirstlane(int32_t(before)));
        if (before == uint32_t(cnt - 1)) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            if (threadIdx.x == 0) __hip_atomic_store(c.arrive + first, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (g == 0) sum_slots_finish<P, VE, LPR>(a, c.scratch, p, row, first, int32_t(cnt), l);
        }
    }
}

template <class P, int VE, int LPR, int BS>
__global__ __launch_bounds__(BS) void k_row_chunk(const RowPass a, const ChunkArgs c) {
    row_chunk_block<P, VE, LPR, BS>(a, c, blockIdx.x);
}

// The long-row chunks and the short-row pass of one row range in one grid:
// workgroups below chunk_blocks are chunks (dispatched first, so the long
// rows, whose chains are the longest, start at once), the rest the short-row
// pass.  Each workgroup does exactly what it does in the kernel of its kind,
// so the bits are those of the two launches; only the wait between them is
// gone, which is the time of a range too small to fill the chip.
template <class P, int VE, int LPR, int S, int BS>
__global__ __launch_bounds__(BS) void k_row_mixed(const RowPass a, const ChunkArgs c, uint32_t chunk_blocks) {
    if (blockIdx.x < chunk_blocks) {
        row_chunk_block<P, VE, LPR, BS>(a, c, blockIdx.x);
    } else {
        row_pass_block<P, VE, LPR, S>(a, blockIdx.x - chunk_blocks);
    }
}

// Rows of several chunks: add the chunk sums in chunk order and finish.
template <class P, int VE, int LPR>
__global__ __launch_bounds__(256) void k_row_reduce(const RowPass a, const ReduceArgs r) {
    constexpr int kGroupsPerWave = 64 / LPR;
    const int lane = int(threadIdx.x & 63u);
    const int l = lane & (LPR - 1);
    const int64_t i64 =
        (int64_t(blockIdx.x) * int64_t(blockDim.x >> 6) + int64_t(threadIdx.x >> 6)) * kGroupsPerWave + lane / LPR;
    if (i64 >= int64_t(r.count)) return;
    const int32_t i = r.first + int32_t(i64);
    const int32_t p = r.pos ? r.pos[i] : i, s0 = r.slot[i], ns = r.cnt[i];
    const int32_t row = a.node_of ? a.node_of[p] : p;
    sum_slots_finish<P, VE, LPR>(a, r.scratch, p, row, s0, ns, l);
}

// Matrices of up to this many rows take the other lane width (dispatch_ve).
constexpr int32_t kAltLaneMaxRows = 65536;

inline int lanes_per_row(int32_t ncols, int ve) {
    const int32_t need = (ncols + ve - 1) / ve;
    int lpr = 1;
    while (lpr < need && lpr < 64) lpr <<= 1;
    return lpr;
}

inline bool aligned16(const void *p, int64_t ld, int elem) {
    return (reinterpret_cast<uintptr_t>(p) & 15u) == 0 && ((ld * elem) & 15) == 0;
}

enum class Kind { kPass, kChunk, kReduce, kTail, kMixed };

struct Launch {
    Kind kind;
    const ChunkArgs *chunk;
    const ReduceArgs *reduce;
    const TailArgs *tail = nullptr;
    int *launched = nullptr;  // kMixed: kernels launched
};

// The lanes a short-row pass gives each row (see RowPass::wave_rows): 64 (a
// wavefront per row, its entries in 64 / LPR slices), 16 (four rows per
// wavefront, 16 / LPR slices each; rows of fewer than 16 lanes only) or LPR
// (rows packed).  By rule: a wavefront per row while the rows cannot fill the
// device's wave slots packed and fit them in one round as wavefronts; 16
// lanes per row where more rows than slots would need several rounds of
// wavefronts; packed where packed rows fill the slots.
template <int LPR>
int row_lanes(const RowPass &rp) {
    // XCD slabs are for matrices that fill the chip; with them, rows stay packed.
    if (LPR >= 64 || rp.wave_rows == 0 || rp.xcd_slabs > 1) return LPR;
    if (rp.wave_rows == 1) return 64;
    if (rp.wave_rows == 2) return LPR < 16 ? 16 : LPR;
    if (rp.rows_hint <= 0 || int64_t(rp.rows_hint) * LPR >= 64 * kWaveSlots) return LPR;
    if (LPR < 16 && int64_t(rp.rows_hint) > kWaveSlots) return 16;
    return 64;
}

// Whether one grid of k_row_mixed keeps as many waves per CU as the chunk
// kernel and the short-row pass it stands for (it needs the registers of the
// larger part), asked of the runtime once per instantiation.
template <class P, int VE, int LPR, int S, int BS>
bool mixed_keeps_occupancy() {
    static const bool keeps = [] {
        int m = 0, c = 0, p = 0;
        const bool ok =
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&m, k_row_mixed<P, VE, LPR, S, BS>, BS, 0) == hipSuccess &&
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&c, k_row_chunk<P, VE, LPR, BS>, BS, 0) == hipSuccess &&
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&p, k_row_pass<P, VE, LPR, S>, 256, 0) == hipSuccess;
        (void)hipGetLastError();
        return ok && m * BS >= c * BS && m * BS >= p * 256;
    }();
    return keeps;
}

template <class P, int VE, int LPR>
hipError_t launch_one(const RowPass &rp, const Launch &lk, hipStream_t stream) {
    constexpr int64_t kRowsPerBlock = 4 * (64 / LPR);
    if (lk.kind == Kind::kPass && rp.ranges) {
        // One grid for all ranges and phases: 8 workgroups (one per XCD) per
        // row block, all row blocks of a phase before the next phase's.
        int64_t most = 0;
        for (int k = 0; k < rp.range_count; ++k)
            most = std::max<int64_t>(most, int64_t(rp.range_first[k + 1]) - rp.range_first[k]);
        if (most <= 0) return hipSuccess;
        RowPass r = rp;
        r.range_blocks = int32_t((most + kRowsPerBlock - 1) / kRowsPerBlock);
        k_row_pass<P, VE, LPR, 1><<<dim3(unsigned(int64_t(r.range_blocks) * r.range_phases * 8)), dim3(256), 0,
                                    stream>>>(r);
    } else if (lk.kind == Kind::kMixed) {
        if (rp.ranges) return hipErrorInvalidValue;
        const int64_t n = std::max<int64_t>(0, int64_t(rp.p_end) - rp.p_begin);
        const int64_t ns = rp.xcd_slabs;
        const int64_t chunk_blocks = int64_t(lk.chunk->count) * ns;
        // The chunk kernel's workgroup size (16 groups of LPR lanes, at least
        // 256 threads) for the whole grid; the short-row part packs its rows
        // into workgroups of that size.
        constexpr int kBs = (16 * LPR > 256) ? 16 * LPR : 256;
        constexpr int64_t kWaves = kBs / 64;
        constexpr int kWaveS = LPR < 64 ? 64 / LPR : 1;
        constexpr int kMidS = LPR < 16 ? 16 / LPR : 1;
        constexpr int64_t kRows = kWaves * (64 / LPR);
        const int lanes = row_lanes<LPR>(rp);
        const bool wave = lanes == 64;
        const bool mid = lanes == 16 && LPR < 16;
        const int64_t per_block = wave ? kWaves : mid ? kWaves * 4 : kRows;
        const int64_t pass_blocks = (n + per_block - 1) / per_block * ns;
        *lk.launched = 0;
        if (chunk_blocks + pass_blocks == 0) return hipSuccess;
        // One grid gains the wait between the two launches and lets the
        // short rows fill the chip while the chunks' longer chains finish.
        // It runs with the registers of the larger part, so a grid made
        // mostly of chunks (of a product that fills the chip) takes it only
        // if that costs the chunks no waves.
        const bool small = rp.rows_hint > 0 && int64_t(rp.rows_hint) * LPR < 64 * kWaveSlots;
        const bool one =
            rp.mixed_grid > 0 ||
            (rp.mixed_grid < 0 && (small || chunk_blocks <= pass_blocks ||
                                   (wave  ? mixed_keeps_occupancy<P, VE, LPR, kWaveS, kBs>()
                                    : mid ? mixed_keeps_occupancy<P, VE, LPR, kMidS, kBs>()
                                          : mixed_keeps_occupancy<P, VE, LPR, 1, kBs>())));
        if (!one) {
            const hipError_t e = launch_one<P, VE, LPR>(rp, Launch{Kind::kChunk, lk.chunk, nullptr}, stream);
            if (e != hipSuccess) return e;
            *lk.launched = n > 0 ? 2 : 1;
            return launch_one<P, VE, LPR>(rp, Launch{Kind::kPass, nullptr, nullptr}, stream);
        }
        *lk.launched = 1;
        RowPass r = rp;
        r.order_blocks = int32_t(pass_blocks);
        if constexpr (LPR < 64) {
            if (wave) {
                k_row_mixed<P, VE, LPR, kWaveS, kBs><<<dim3(unsigned(chunk_blocks + pass_blocks)), dim3(kBs), 0,
                                                       stream>>>(r, *lk.chunk, uint32_t(chunk_blocks));
                return hipGetLastError();
            }
        }
        if constexpr (LPR < 16) {
            if (mid) {
                k_row_mixed<P, VE, LPR, kMidS, kBs><<<dim3(unsigned(chunk_blocks + pass_blocks)), dim3(kBs), 0,
                                                      stream>>>(r, *lk.chunk, uint32_t(chunk_blocks));
                return hipGetLastError();
            }
        }
        k_row_mixed<P, VE, LPR, 1, kBs><<<dim3(unsigned(chunk_blocks + pass_blocks)), dim3(kBs), 0, stream>>>(
            r, *lk.chunk, uint32_t(chunk_blocks));
    } else if (lk.kind == Kind::kPass) {
        const int64_t n = int64_t(rp.p_end) - rp.p_begin;
        if (n <= 0) return hipSuccess;
        const int64_t ns = rp.xcd_slabs;
        RowPass r = rp;
        const int lanes = row_lanes<LPR>(rp);
        if constexpr (LPR < 64) {
            if (lanes == 64) {
                r.order_blocks = int32_t((n + 3) / 4 * ns);
                k_row_pass<P, VE, LPR, 64 / LPR><<<dim3(unsigned(r.order_blocks)), dim3(256), 0, stream>>>(r);
                return hipGetLastError();
            }
        }
        if constexpr (LPR < 16) {
            if (lanes == 16) {
                r.order_blocks = int32_t((n + 15) / 16 * ns);
                k_row_pass<P, VE, LPR, 16 / LPR><<<dim3(unsigned(r.order_blocks)), dim3(256), 0, stream>>>(r);
                return hipGetLastError();
            }
        }
        r.order_blocks = int32_t((n + kRowsPerBlock - 1) / kRowsPerBlock * ns);
        k_row_pass<P, VE, LPR, 1><<<dim3(unsigned(r.order_blocks)), dim3(256), 0, stream>>>(r);
    } else if (lk.kind == Kind::kTail) {
        if constexpr (P::kDual && P::kExtra == kExtraSelf) {  // the reverse pass only
            if (lk.tail->n_frag <= 0) return hipSuccess;
            k_row_tail<P, VE, LPR><<<dim3(unsigned(lk.tail->n_frag)), dim3(256), 0, stream>>>(rp, *lk.tail);
        } else {
            return hipErrorInvalidValue;
        }
    } else if (lk.kind == Kind::kChunk) {
        if (lk.chunk->count <= 0) return hipSuccess;
        constexpr int kBs = (16 * LPR > 256) ? 16 * LPR : 256;
        k_row_chunk<P, VE, LPR, kBs>
            <<<dim3(unsigned(int64_t(lk.chunk->count) * rp.xcd_slabs)), dim3(kBs), 0, stream>>>(rp, *lk.chunk);
    } else {
        const int64_t n = lk.reduce->count;
        if (n <= 0) return hipSuccess;
        k_row_reduce<P, VE, LPR>
            <<<dim3(unsigned((n + kRowsPerBlock - 1) / kRowsPerBlock)), dim3(256), 0, stream>>>(rp, *lk.reduce);
    }
    return hipGetLastError();
}

template <class P, int VE>
hipError_t dispatch_lpr(const RowPass &rp, const Launch &lk, int lpr, hipStream_t stream) {
    switch (lpr) {
        case 1: return launch_one<P, VE, 1>(rp, lk, stream);
        case 2: return launch_one<P, VE, 2>(rp, lk, stream);
        case 4: return launch_one<P, VE, 4>(rp, lk, stream);
        case 8: return launch_one<P, VE, 8>(rp, lk, stream);
        case 16: return launch_one<P, VE, 16>(rp, lk, stream);
        case 32: return launch_one<P, VE, 32>(rp, lk, stream);
        case 64: return launch_one<P, VE, 64>(rp, lk, stream);
        default: return hipErrorInvalidValue;
    }
}

// The vector path needs every operand 16-byte aligned with row strides that
// keep it so; anything else runs one column per lane.
template <class P>
hipError_t dispatch_ve(const RowPass &rp, const Launch &lk, hipStream_t stream) {
    using TG = typename P::TG;
    // 16-byte loads of the narrowest gathered type (a dual pass over FP32 z
    // and BF16 g reads 8 columns per lane: 16 bytes of g, 32 of z).
    constexpr int kNarrow = (P::kDual && sizeof(typename P::TG2) < sizeof(TG)) ? int(sizeof(typename P::TG2))
                                                                                  : int(sizeof(TG));
    constexpr int kVe = 16 / kNarrow;
    bool vec = (rp.ncols % kVe) == 0 && aligned16(rp.g, rp.ldg, int(sizeof(TG))) && aligned16(rp.y, rp.ldy, 4);
    if (P::kExtra != kExtraNone) vec = vec && aligned16(rp.e, rp.lde, int(sizeof(typename P::TE)));
    if (P::kDual) vec = vec && aligned16(rp.g2, rp.ldg2, int(sizeof(typename P::TG2)));
    if (P::kExtra == kExtraParentKahan) vec = vec && aligned16(rp.comp, rp.ldc, 4);
    // The other lane width of a single-source pass over a small matrix
    // (kAlt columns per lane, 0 for none), where the pass is bound by its
    // rows' chains of dependent gathers rather than by bandwidth: FP32 rows
    // of 64 or more columns take 32 bytes per lane, so a row has half the
    // lanes and a wave twice the rows (with a wavefront per row, each row's
    // entries split over twice the sub-groups: chains half as long); BF16
    // rows narrower than 64 bytes take 8 bytes per lane, so that such a row
    // keeps four lanes.  N = 19,717: FP32 F = 128 step 0.074 -> 0.043 ms,
    // BF16 F = 16 0.034 -> 0.025 ms.  On matrices that fill the device the
    // usual widths measured faster (N = 169,343, FP32 F = 128: 0.268 against
    // 0.291 ms).
    constexpr int kAlt = P::kDual ? 0 : (sizeof(TG) == 4 ? 8 : 4);
    const bool small = rp.rows_hint > 0 && rp.rows_hint <= kAltLaneMaxRows;
    bool alt = false;
    if constexpr (kAlt == 8) alt = small && vec && rp.ncols % 8 == 0 && rp.ncols >= 64;
    if constexpr (kAlt == 4) alt = small && vec && rp.ncols < 32;
    const int ve = vec ? (alt ? kAlt : kVe) : 1;
    // Slab width: a multiple of the vector width and at least 16 vectors, so
    // that every slab keeps 16 or more lanes per row.  Rows that wide (and
    // the unslabbed rows, which are wider) give the chunk kernel 16 groups
    // either way, so its entry slices, and with them the bits, do not
    // change; narrower gathers than 16 vectors also measured slower.
    RowPass r = rp;
    int32_t w;
    if (rp.ranges) {
        // The range mapping needs exactly xcd_slabs slabs per range (a slab
        // left empty on narrow rows is fine: the others cover the columns).
        w = (rp.ncols + rp.xcd_slabs - 1) / rp.xcd_slabs;
        w = (w + ve - 1) / ve * ve;
    } else {
        const int32_t want = std::min<int32_t>(std::max<int32_t>(1, rp.xcd_slabs),
                                               std::max<int32_t>(1, rp.ncols / (16 * ve)));
        w = (rp.ncols + want - 1) / want;
        w = (w + ve - 1) / ve * ve;
        r.xcd_slabs = (rp.ncols + w - 1) / w;
    }
    r.slab_w = w;
    if constexpr (kAlt > 0) {
        if (alt) return dispatch_lpr<P, kAlt>(r, lk, lanes_per_row(w, kAlt), stream);
    }
    if (vec) return dispatch_lpr<P, kVe>(r, lk, lanes_per_row(w, kVe), stream);
    return dispatch_lpr<P, 1>(r, lk, lanes_per_row(w, 1), stream);
}

// The pass shapes the library runs (see dispatch in cbm_kernels.hip); each
// is instantiated in exactly one rp_*.hip file.
using PassF32 = Pass<float, float, float, kExtraNone, false>;
using PassF32Parent = Pass<float, float, float, kExtraParent, false>;
using PassF32Kahan = Pass<float, float, float, kExtraParentKahan, false>;
using PassF32Self = Pass<float, float, float, kExtraSelf, false>;
using PassBf16 = Pass<bf16_t, bf16_t, float, kExtraNone, false>;
using PassBf16Parent = Pass<bf16_t, bf16_t, float, kExtraParent, false>;
using PassBf16Kahan = Pass<bf16_t, bf16_t, float, kExtraParentKahan, false>;
using PassDualSelfF32 = Pass<float, float, float, kExtraSelf, true>;
using PassDualSelfBf16 = Pass<float, bf16_t, bf16_t, kExtraSelf, true>;
using PassDualF32 = Pass<float, float, float, kExtraNone, true>;
using PassDualBf16 = Pass<float, bf16_t, float, kExtraNone, true>;

#define CBM_RP_DECL(P) extern template hipError_t dispatch_ve<P>(const RowPass &, const Launch &, hipStream_t);
CBM_RP_DECL(PassF32) CBM_RP_DECL(PassF32Parent) CBM_RP_DECL(PassF32Kahan) CBM_RP_DECL(PassF32Self)
CBM_RP_DECL(PassBf16) CBM_RP_DECL(PassBf16Parent) CBM_RP_DECL(PassBf16Kahan)
CBM_RP_DECL(PassDualSelfF32) CBM_RP_DECL(PassDualSelfBf16) CBM_RP_DECL(PassDualF32) CBM_RP_DECL(PassDualBf16)
#undef CBM_RP_DECL

}  // namespace rp
}  // namespace cbmk
