// The forward product in one launch: Delta pass and tree propagation of
// every level, unit by unit (the plan's CBM_P_UN_* tables).
//
// A unit is a set of connected pieces of the tree ("components") of at most
// CBM_UNIT_ROWS rows.  One workgroup takes one unit and one 64-column slab:
//   1. its rows' Delta entries, cut into pieces of at most CBM_UNIT_PIECE
//      entries, are gathered by the workgroup's lane groups, one piece each
//      (pieces of one row run side by side), into LDS;
//   2. level by level within the unit (a workgroup barrier between levels),
//      each row adds its pieces in order, applies the row scale, and adds
//      its parent's finished row -- from LDS when the parent is in the unit,
//      otherwise from the output, once the unit holding it has set its flag;
//   3. the unit's rows are written out, then its flag is set.
// Units are taken in ticket order (an atomic counter), and a unit's outside
// parents lie in units of lower index, so every unit a workgroup waits on is
// held by a workgroup that is already running: the waits always end.  Each
// wait is still bounded (spin_limit polls); a wait that gives up counts into
// the error word, which the host reads with cbm_fused_errors.
//
// The counters live in device memory and reset themselves: the last
// workgroup to finish clears the ticket and done counters and advances the
// epoch, and a flag means "finished in this call" when it holds epoch + 1.
// So a call needs no memset and can be replayed from a HIP graph.
#include "row_pass.hpp"

namespace cbmk {
namespace {

using namespace rp;

constexpr int kFusedThreads = 512;
constexpr int kSlabCols = 64;

template <typename TX, int VE>
__global__ __launch_bounds__(kFusedThreads) void k_fused_forward(const FusedArgs a) {
    constexpr int LPR = VE == 1 ? 64 : kSlabCols / VE;  // lanes per row segment of one slab
    constexpr int kGroups = kFusedThreads / LPR;
    __shared__ float lds[CBM_UNIT_PIECES * kSlabCols];
    __shared__ uint32_t s_ticket, s_epoch;
    if (a.ordered) {
        if (threadIdx.x == 0) {
            s_epoch = __hip_atomic_load(a.state + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            s_ticket = __hip_atomic_fetch_add(a.state, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
    } else if (threadIdx.x == 0) {
        s_epoch = 0;
        s_ticket = blockIdx.x;
    }
    __syncthreads();
    const uint32_t ticket = s_ticket, want = s_epoch + 1u;
    const int32_t u = int32_t(ticket / uint32_t(a.n_slabs)), slab = int32_t(ticket % uint32_t(a.n_slabs));
    const int group = int(threadIdx.x) / LPR, l = int(threadIdx.x) % LPR;
    const int32_t c0 = slab * kSlabCols + l * VE;  // first column of this lane
    const bool active = c0 < a.ncols;
    const int lc = l * VE;                         // the lane's column within the slab
    const int32_t r0 = a.un_ptr[u], r1 = a.un_ptr[u + 1];
    int32_t pc0 = a.un_pc[r0], pc1 = a.un_pc[r1];
    // A unit beyond the LDS tables (a damaged descriptor) computes nothing
    // and counts as an error, but still sets its flag below.
    const bool fits = r1 >= r0 && r1 - r0 <= CBM_UNIT_ROWS && pc1 >= pc0 && pc1 - pc0 <= CBM_UNIT_PIECES;
    if (!fits) {
        if (threadIdx.x == 0) __hip_atomic_fetch_add(a.state + 3, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        pc1 = pc0;
    }
    RowPass rp;
    rp.idx = a.idx;
    const TX *x = static_cast<const TX *>(a.x);

    // 1. pieces into LDS
    for (int32_t pi = pc0 + group; pi < pc1; pi += kGroups) {
        float acc[VE];
#pragma unroll
        for (int q = 0; q < VE; ++q) acc[q] = 0.0f;
        gather_src<TX, VE, LPR>(rp, x, a.ldx, a.colscale, nullptr, a.un_pcb[pi], a.un_pce[pi], l, c0, active, acc);
        float *dst = lds + (pi - pc0) * kSlabCols + lc;
#pragma unroll
        for (int q = 0; q < VE; ++q) dst[q] = acc[q];
    }
    __syncthreads();

    // 2. rows level by level; a row's value ends in its first piece's slot
    const int32_t sg1 = fits ? a.un_seg_ptr[u + 1] : 0;
    for (int32_t k = fits ? a.un_seg_ptr[u] : 0; k < sg1; ++k) {
        const int32_t sb = a.un_seg[k], se = k + 1 < sg1 ? a.un_seg[k + 1] : r1;
        for (int32_t i = sb + group; i < se; i += kGroups) {
            const int32_t p = a.un_row[i];
            const int32_t f0 = a.un_pc[i] - pc0, f1 = a.un_pc[i + 1] - pc0;
            float *slot = lds + f0 * kSlabCols + lc;
            float acc[VE];
#pragma unroll
            for (int q = 0; q < VE; ++q) acc[q] = slot[q];
            for (int32_t f = f0 + 1; f < f1; ++f)
#pragma unroll
                for (int q = 0; q < VE; ++q) acc[q] += lds[f * kSlabCols + lc + q];
            const int32_t row = a.node_of[p];
            const float rs = a.rowscale ? a.rowscale[row] : 1.0f;
#pragma unroll
            for (int q = 0; q < VE; ++q) acc[q] *= rs;
            const int32_t lp = a.un_lpar[i];
            if (lp != -1) {
                const float es = a.escale ? a.escale[p] : 1.0f;
                float v[VE];
                if (lp >= 0 && lp < r1 - r0) {
                    const float *src = lds + (a.un_pc[r0 + lp] - pc0) * kSlabCols + lc;
#pragma unroll
                    for (int q = 0; q < VE; ++q) v[q] = src[q];
                } else {
                    // the parent's unit, this slab: wait for its flag (only
                    // units of lower index are waited on)
                    const int32_t pu = a.un_of[a.parent_pos[p]];
                    const bool earlier = pu >= 0 && pu < u;
                    if (!earlier) __hip_atomic_fetch_add(a.state + 3, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    const uint32_t *flag = a.flags + int64_t(earlier ? pu : u) * a.n_slabs + slab;
                    uint32_t polls = 0;
                    while (earlier && __hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != want) {
                        __builtin_amdgcn_s_sleep(2);
                        if (++polls > a.spin_limit) {
                            __hip_atomic_fetch_add(a.state + 3, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                            break;
                        }
                    }
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                    if (active) {
                        Load<float, VE>::run(a.y + int64_t(a.parent_of[p]) * a.ldy + c0, v);
                    } else {
#pragma unroll
                        for (int q = 0; q < VE; ++q) v[q] = 0.0f;
                    }
                }
#pragma unroll
                for (int q = 0; q < VE; ++q) acc[q] = fmaf(v[q], es, acc[q]);
            }
#pragma unroll
            for (int q = 0; q < VE; ++q) slot[q] = acc[q];
        }
        __syncthreads();
    }

    // 3. rows out, then the flag (release: every wave's stores are made
    // visible device-wide before the barrier that precedes the flag store)
    for (int32_t i = r0 + group; fits && i < r1; i += kGroups) {
        if (!active) continue;
        const float *src = lds + (a.un_pc[i] - pc0) * kSlabCols + lc;
        float v[VE];
#pragma unroll
        for (int q = 0; q < VE; ++q) v[q] = src[q];
        store_row<VE>(a.y + int64_t(a.node_of[a.un_row[i]]) * a.ldy + c0, v);
    }
    // Only a unit that a later unit waits on publishes.  Every wave releases
    // its own stores at device scope (a single fence by one thread after the
    // barrier let waiters read stale rows), then one thread sets the flag.
    const bool waited = a.ordered && a.un_waited[u] != 0;
    if (waited) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    if (!a.ordered) return;  // no flags, no counters
    __syncthreads();
    if (threadIdx.x == 0) {
        if (waited)
            __hip_atomic_store(a.flags + int64_t(u) * a.n_slabs + slab, want, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t total = uint32_t(a.n_units) * uint32_t(a.n_slabs);
        if (__hip_atomic_fetch_add(a.state + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == total - 1u) {
            // every workgroup of this call has taken its ticket, read the
            // epoch and finished: reset for the next call (kernel boundaries
            // order these stores before the next call's reads)
            __hip_atomic_store(a.state, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(a.state + 1, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(a.state + 2, want, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

}  // namespace

hipError_t launch_fused_forward(const FusedArgs &a, bool x_bf16, hipStream_t stream) {
    if (a.n_units <= 0 || a.n_slabs <= 0 || a.ncols <= 0) return hipSuccess;
    const int64_t total = int64_t(a.n_units) * a.n_slabs;
    if (total > int64_t(UINT32_MAX)) return hipErrorInvalidValue;
    const int elem = x_bf16 ? 2 : 4;
    const int ve_vec = 16 / elem;
    const bool vec = (a.ncols % ve_vec) == 0 && aligned16(a.x, a.ldx, elem) && aligned16(a.y, a.ldy, 4);
    const dim3 grid(static_cast<unsigned>(total)), block(kFusedThreads);
    if (x_bf16) {
        if (vec) {
            k_fused_forward<bf16_t, 8><<<grid, block, 0, stream>>>(a);
        } else {
            k_fused_forward<bf16_t, 1><<<grid, block, 0, stream>>>(a);
        }
    } else if (vec) {
        k_fused_forward<float, 4><<<grid, block, 0, stream>>>(a);
    } else {
        k_fused_forward<float, 1><<<grid, block, 0, stream>>>(a);
    }
    return hipGetLastError();
}

}  // namespace cbmk
