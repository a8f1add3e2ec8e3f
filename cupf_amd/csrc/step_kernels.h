// step_kernels.h -- the persistent step and fix-up kernels (single simulation
// and batched), instantiated once per (K, gather variant) by the step_k*_g*.hip
// files so the large template instantiations compile in parallel.
#pragma once

#include <algorithm>

#include "apt_device.h"

namespace cupf {

namespace {

// APTrhs_psi (3DgrainMPI/APTPhaseField.cu:178-354) over every interior point:
// a persistent grid of step_blocks workgroups takes work items one at a time,
// the edge items first (direct path, slowest per point), then the listed tile
// chunks.  Every workgroup leaves once the items are exhausted; nothing waits
// on another workgroup.
// OVF: the variant for steps whose neighbourhoods overflow K (in-tile keep
// test, fix-up records; 7-point gather only); the host picks it per launch
// (StepArgs::ovf_variant).  Its edge items are apt_edge's, on a second stream.
template <int K, bool G27, bool OVF>
__global__ __launch_bounds__(kBlock) void apt_step_tiled(StepArgs a)
{
    __shared__ unsigned s_q;
    const float a_s = 1.0f + 3.0f * a.kin_delta;
    const float epsv = -4.0f * a.kin_delta / a_s;
    const unsigned ne = OVF ? 0u : (unsigned)(a.nzchunk * a.edge_items);
    const unsigned nh = __atomic_load_n(a.wl_count, __ATOMIC_RELAXED);
    const unsigned n = ne + nh + __atomic_load_n(a.wl_light, __ATOMIC_RELAXED);
    const unsigned cap = (unsigned)(a.nzchunk * a.items_chunk);
    unsigned nres = 0u;  // overflowed points this wave resolved in its tiles
    for (;;) {
        // the next item is taken only when this workgroup is free, which keeps
        // the tail of the list balanced
        if (threadIdx.x == 0) s_q = atomicAdd(a.wl_count + 1, 1u);
        __syncthreads();
        // uniform over the workgroup: the item's fields come through scalar loads
        const unsigned q = __builtin_amdgcn_readfirstlane(s_q);
        __syncthreads();
        if (q >= n) break;
        if (q < ne) {
            const int zc = (int)q / a.edge_items;
            const int e = (int)q - zc * a.edge_items;
            const int k0 = 1 + zc * a.kchunk;
            const int k1 = min(k0 + a.kchunk, a.fnz - 1);
#pragma unroll 1
            for (int r = 0; r < kEdgePerLane; ++r)
                edge_point<K, G27>(a, e * kEdgePoints + r * kBlock + threadIdx.x, k0, k1, a_s, epsv);
        } else {
            // heavy chunks from the front of the list, then light ones from the back
            const unsigned w = q - ne < nh ? q - ne : cap - 1u - (q - ne - nh);
            const int g = a.wl_item[w];
            const int zc = g / a.items_chunk;
            const int t = g - zc * a.items_chunk;
            const int ty = t / a.tiles_x;
            const int tx = t - ty * a.tiles_x;
            const int k0 = 1 + zc * a.kchunk;
            const int k1 = min(k0 + a.kchunk, a.fnz - 1);
            tile_march<K, G27, OVF>(a, tx, ty, k0, k1, a.wl_need[w], a.wl_act[w], a_s, epsv, nres);
        }
    }
    if ((threadIdx.x & 63) == 0 && nres) atomicAdd(a.overflow, (unsigned long long)nres);
}

// The edge items (the +x strip outside the tiles) of an overflow step (7-point
// gather): one workgroup per item; a one-column strip as edge tiles (keep
// test and records), other strips on the direct path.  Out of the step kernel
// so its tile march keeps registers (206 against 240 VGPRs, fewer scalar
// spills), and on a second stream that forks after the plan kernel and joins
// before the fix-up, so its workgroups run beside the step kernel's.  The two
// write disjoint points and activity words and meet only in the fix-up
// list's atomics.  (Steps without overflow keep their edge items in the step
// kernel: there the second launch and its idle workgroups, each holding its
// LDS, cost more than they save.)
template <int K>
__global__ __launch_bounds__(kBlock) void apt_edge(StepArgs a)
{
    const float a_s = 1.0f + 3.0f * a.kin_delta;
    const float epsv = -4.0f * a.kin_delta / a_s;
    const int q = blockIdx.x;
    const int zc = q / a.edge_items;
    const int e = q - zc * a.edge_items;
    const int k0 = 1 + zc * a.kchunk;
    const int k1 = min(k0 + a.kchunk, a.fnz - 1);
    if (edge_strip_is_column(a)) {
        const unsigned nres = edge_tile<K>(a, e, k0, k1, a_s, epsv);
        if ((threadIdx.x & 63) == 0 && nres) atomicAdd(a.overflow, (unsigned long long)nres);
    } else {
#pragma unroll 1
        for (int r = 0; r < kEdgePerLane; ++r)
            edge_point<K, false>(a, e * kEdgePoints + r * kBlock + threadIdx.x, k0, k1, a_s, epsv);
    }
}

// Fix-up pass for the points of the last apt_step launch whose neighbourhood
// overflowed K slots; adds their number to the overflow diagnostic.  The list
// lives in this step's counter set, which the next-but-one plan kernel
// clears, so nothing here has to wait for the other workgroups.
template <int K, bool G27>
__device__ __forceinline__ void fixup_run(const StepArgs& a_arg, int blk, int nblk)
{
    // The arguments are read from an LDS copy: with them in scalar registers
    // the compiler keeps 229 VGPRs live (2 waves/SIMD); from LDS it needs 92
    // for K=8 (5 waves/SIMD; 120 and 4 with the neighbour entries loaded one
    // neighbour ahead), and each entry is a long chain of dependent loads that
    // only more waves in flight can overlap (more workgroups per CU made the
    // start-up transient of the 256x256x512 run 14 % faster than 2).
    __shared__ StepArgs s_a;
    constexpr int kWords = (int)(sizeof(StepArgs) / 4);
    static_assert(sizeof(StepArgs) % 4 == 0 && kWords <= kFixLanes, "StepArgs staging");
    if ((int)threadIdx.x < kWords)
        reinterpret_cast<unsigned*>(&s_a)[threadIdx.x] =
            reinterpret_cast<const unsigned*>(&a_arg)[threadIdx.x];
    __syncthreads();
    const StepArgs& a = s_a;
    const unsigned n = __atomic_load_n(a.fix_count, __ATOMIC_RELAXED);
    if (blk == 0 && threadIdx.x == 0 && a.seen) {  // for the host's variant choice
        if (n >= kOvfHeavy) a.seen[1] = a.launch_index;
        a.seen[0] = a.launch_index;
    }
    if (n == 0u) return;
    if (blk == 0 && threadIdx.x == 0) atomicAdd(a.overflow, (unsigned long long)n);
    const float a_s = 1.0f + 3.0f * a.kin_delta;
    const float epsv = -4.0f * a.kin_delta / a_s;
#pragma unroll 1
    for (unsigned q = blk * kFixLanes + threadIdx.x; q < n; q += nblk * kFixLanes) {
        int i, j, k;
        bool rec;
        fixup_point(a, q, i, j, k, rec);
        if (!G27 && rec) apt_point_overflow_rec<K>(a, i, j, k, (int)q, a_s, epsv);
        else apt_point_overflow<K, G27>(a, i, j, k, a_s, epsv);
    }
}

template <int K, bool G27>
__global__ __launch_bounds__(kFixLanes) void apt_fixup(StepArgs a)
{
    fixup_run<K, G27>(a, (int)blockIdx.x, (int)gridDim.x);
}

// The fix-up of one step and the work-list plan of the next in one launch:
// blocks [0, fa.fix_blocks) run the fix-up with fa, the others the plan with
// pa.  The plan reads only activity words, which the fix-up does not write
// (the step kernel marks an overflowed point's tile), and the fix-up reads
// its own step's counter set while the plan fills the next one and clears
// the one after (three sets in turn), so the two halves need no order
// between them.  Saves one launch and its dispatch gap per step.
template <int K, bool G27>
__global__ __launch_bounds__(kBlock) void apt_fixplan(StepArgs fa, StepArgs pa)
{
    static_assert(kFixLanes == kBlock, "one workgroup size for both halves");
    if ((int)blockIdx.x < fa.fix_blocks) fixup_run<K, G27>(fa, (int)blockIdx.x, fa.fix_blocks);
    else plan_block(pa, (int)blockIdx.x - fa.fix_blocks);
}

// Index of the simulation whose range [base_i, base_{i+1}) holds work index q
// (bases ascending; sims with an empty range are skipped over).
__device__ __forceinline__ int batch_find(const BatchSim* sims, int n, int q, bool edge)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        const int base = edge ? sims[mid].edge_base : sims[mid].plan_base;
        if (base <= q) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Batched APTrhs_psi: the persistent grid takes the edge items of every
// simulation, then the tile chunks the batched plan kernel listed, each with
// its own simulation's step arguments.
template <int K, bool G27, bool OVF>
__global__ __launch_bounds__(kBlock) void batch_step(BatchArgs b)
{
    __shared__ unsigned s_q;
    // The members' records do not change during the kernel (batch_advance
    // rewrites them after it): read through the constant address space with a
    // workgroup-uniform index, every step-argument field is a scalar load.
    typedef const __attribute__((address_space(4))) BatchSim CSim;
    const CSim* sims = (const CSim*)b.sims;
    const unsigned* cnt = b.cnt + 4 * __atomic_load_n(b.cnt + 8, __ATOMIC_RELAXED);
    const unsigned ne = (unsigned)b.edge_total;
    const unsigned nh = __atomic_load_n(cnt, __ATOMIC_RELAXED);
    const unsigned n = ne + nh + __atomic_load_n(cnt + 2, __ATOMIC_RELAXED);
    const unsigned cap = (unsigned)b.wl_cap;
    for (;;) {
        if (threadIdx.x == 0) s_q = atomicAdd((unsigned*)cnt + 1, 1u);
        __syncthreads();
        const unsigned q = __builtin_amdgcn_readfirstlane(s_q);
        if (q >= n) break;
        const bool edge = q < ne;
        // heavy chunks from the front of the list, then light ones from the back
        const unsigned w = q - ne < nh ? q - ne : cap - 1u - (q - ne - nh);
        const int sim =
            __builtin_amdgcn_readfirstlane(edge ? batch_find(b.sims, b.n, (int)q, true) : b.wl_sim[w]);
        const StepArgs& a = *(const StepArgs*)&sims[sim].a;
        const float a_s = 1.0f + 3.0f * a.kin_delta;
        const float epsv = -4.0f * a.kin_delta / a_s;
        if (edge) {
            const int e0 = (int)q - sims[sim].edge_base;
            const int zc = e0 / a.edge_items;
            const int e = e0 - zc * a.edge_items;
            const int k0 = 1 + zc * a.kchunk;
            const int k1 = min(k0 + a.kchunk, a.fnz - 1);
            if (!G27 && edge_strip_is_column(a)) {
                const unsigned nres = edge_tile<K, OVF>(a, e, k0, k1, a_s, epsv);
                if ((threadIdx.x & 63) == 0 && nres) atomicAdd(a.overflow, (unsigned long long)nres);
            } else {
#pragma unroll 1
                for (int r = 0; r < kEdgePerLane; ++r)
                    edge_point<K, G27>(a, e * kEdgePoints + r * kBlock + threadIdx.x, k0, k1, a_s, epsv);
            }
        } else {
            const int g = b.wl_item[w];
            const int zc = g / a.items_chunk;
            const int t = g - zc * a.items_chunk;
            const int ty = t / a.tiles_x;
            const int tx = t - ty * a.tiles_x;
            const int k0 = 1 + zc * a.kchunk;
            const int k1 = min(k0 + a.kchunk, a.fnz - 1);
            unsigned nres = 0u;
            tile_march<K, G27, OVF>(a, tx, ty, k0, k1, b.wl_need[w], b.wl_act[w], a_s, epsv, nres);
            if ((threadIdx.x & 63) == 0 && nres) atomicAdd(a.overflow, (unsigned long long)nres);
        }
        __syncthreads();  // s_q is rewritten for the next item
    }
}

// Batched fix-up: the overflow lists of all simulations form one index space
// (prefix sums of their lengths, in LDS) that the grid strides over, so long
// and short lists alike are spread over every workgroup.
template <int K, bool G27>
__global__ __launch_bounds__(kFixLanes) void batch_fixup(BatchArgs b)
{
    __shared__ unsigned s_pre[kMaxBatch + 1];
    for (int t = threadIdx.x; t < b.n; t += kFixLanes)
        s_pre[t + 1] = __atomic_load_n(b.sims[t].a.fix_count, __ATOMIC_RELAXED);
    __syncthreads();
    if (threadIdx.x == 0) {
        s_pre[0] = 0u;
        for (int t = 0; t < b.n; ++t) {
            const unsigned c = s_pre[t + 1];
            if (blockIdx.x == 0 && c) atomicAdd(b.sims[t].a.overflow, (unsigned long long)c);
            s_pre[t + 1] = s_pre[t] + c;
        }
    }
    __syncthreads();
    const unsigned total = s_pre[b.n];
    if (blockIdx.x == 0 && threadIdx.x == 0 && b.seen) {  // for the host's variant choice
        const unsigned long long idx = ++b.cnt[9];
        if (total >= kOvfHeavy) b.seen[1] = idx;
        b.seen[0] = idx;
    }
    for (unsigned g = blockIdx.x * kFixLanes + threadIdx.x; g < total; g += gridDim.x * kFixLanes) {
        int lo = 0, hi = b.n - 1;  // last simulation whose range starts at or before g
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (s_pre[mid] <= g) lo = mid;
            else hi = mid - 1;
        }
        const StepArgs& a = b.sims[lo].a;
        const unsigned q = g - s_pre[lo];
        int i, j, k;
        bool rec;
        fixup_point(a, q, i, j, k, rec);
        const float a_s = 1.0f + 3.0f * a.kin_delta;
        const float epsv = -4.0f * a.kin_delta / a_s;
        if (!G27 && rec) apt_point_overflow_rec<K>(a, i, j, k, (int)q, a_s, epsv);
        else apt_point_overflow<K, G27>(a, i, j, k, a_s, epsv);
    }
}

template <int K, bool G27>
hipError_t launch_step_t(const StepArgs& a, hipStream_t stream, const EdgeLaunch* edge,
                         hipEvent_t ev_begin, hipEvent_t ev_planned, hipEvent_t ev_mid,
                         hipEvent_t ev_end, bool planned, const StepArgs* next)
{
    dim3 grid(a.step_blocks), block(kBlock), fgrid(a.fix_blocks), fblock(kFixLanes);
    // overflow steps (7-point gather) run their edge items as apt_edge
    const int ne = (!G27 && a.ovf_variant) ? a.nzchunk * a.edge_items : 0;
    if (ne > 0 && !edge) return hipErrorInvalidValue;
    if (ev_begin) (void)hipEventRecord(ev_begin, stream);
    hipError_t e;
    if (!planned && (e = launch_plan(a, stream)) != hipSuccess) return e;
    if (ev_planned) (void)hipEventRecord(ev_planned, stream);
    if (ne > 0) {
        if ((e = hipEventRecord(edge->ready, stream)) != hipSuccess ||
            (e = hipStreamWaitEvent(edge->stream, edge->ready, 0)) != hipSuccess)
            return e;
        hipLaunchKernelGGL((apt_edge<K>), dim3(ne), block, 0, edge->stream, a);
        if ((e = hipEventRecord(edge->done, edge->stream)) != hipSuccess) return e;
    }
    if (!G27 && a.ovf_variant)
        hipLaunchKernelGGL((apt_step_tiled<K, G27, !G27>), grid, block, 0, stream, a);
    else
        hipLaunchKernelGGL((apt_step_tiled<K, G27, false>), grid, block, 0, stream, a);
    if (ne > 0 && (e = hipStreamWaitEvent(stream, edge->done, 0)) != hipSuccess) return e;
    if (ev_mid) (void)hipEventRecord(ev_mid, stream);
    if (next) {  // this step's fix-up and the next step's plan in one launch
        const int pblocks = max(1, cdiv(plan_waves(next->kchunk, next->nzchunk, next->items_chunk), kBlock / 64));
        hipLaunchKernelGGL((apt_fixplan<K, G27>), dim3(a.fix_blocks + pblocks), block, 0, stream, a, *next);
    } else {
        hipLaunchKernelGGL((apt_fixup<K, G27>), fgrid, fblock, 0, stream, a);
    }
    if (ev_end) (void)hipEventRecord(ev_end, stream);
    return hipGetLastError();
}

template <int K, bool G27>
hipError_t launch_batch_t(const BatchArgs& b, hipStream_t stream)
{
    dim3 grid(b.step_blocks), block(kBlock);
    hipError_t e = launch_batch_plan(b, stream);
    if (e != hipSuccess) return e;
    if (!G27 && b.ovf_variant)
        hipLaunchKernelGGL((batch_step<K, G27, !G27>), grid, block, 0, stream, b);
    else
        hipLaunchKernelGGL((batch_step<K, G27, false>), grid, block, 0, stream, b);
    hipLaunchKernelGGL((batch_fixup<K, G27>), dim3(b.fix_blocks), dim3(kFixLanes), 0, stream, b);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_batch_advance(b, stream);
}

// Workgroups that fit the device at once: which = 0 step, 1 batched step,
// 2 fix-up, 3 batched fix-up.
template <int K, bool G27>
int slots_t(int device, int which)
{
    int ncu = 0, per = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        return 0;
    hipError_t e;
    switch (which) {
    case 0: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, apt_step_tiled<K, G27, !G27>, kBlock, 0); break;
    case 1: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, batch_step<K, G27, !G27>, kBlock, 0); break;
    case 2: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, apt_fixup<K, G27>, kFixLanes, 0); break;
    default: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, batch_fixup<K, G27>, kFixLanes, 0); break;
    }
    if (e != hipSuccess || per < 1) return 0;
    return per * ncu;
}

}  // namespace

}  // namespace cupf
