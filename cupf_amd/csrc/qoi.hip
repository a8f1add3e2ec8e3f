// qoi.hip -- device-side quantities of interest on the collected grain-id
// field (alpha, fnx*fny*fnz with its halo, as pf_collect leaves it): per-grain
// volumes (QOI::calculateQoIs, 3DgrainMPI/QOI.cpp:198-231), the 3-D
// junction search (QOI3D::searchJunctionsOnImage, 3DgrainMPI/QOI.cpp:233-311)
// and the line QoIs' tip scan and grain areas (QOI::calculateLineQoIs,
// 3DgrainMPI/QOI.cpp:38-95).  Only the results cross PCIe, not the field.
#include "apt_device.h"

namespace cupf {

namespace {

// Interior point s (scan order: k-major, then j, then i) to its field index.
// The context keeps K*fnx*fny*fnz below 2^31, so 32-bit arithmetic holds.
__device__ __forceinline__ int interior_index(unsigned s, int nx, int ny, int fnx, int fny, int* k_out)
{
    const unsigned i = s % (unsigned)nx, r = s / (unsigned)nx;
    const unsigned j = r % (unsigned)ny, k = r / (unsigned)ny;
    *k_out = (int)k + 1;
    return ((int)k + 1) * fnx * fny + ((int)j + 1) * fnx + (int)i + 1;
}

// Per-grain point counts and first occurrence over interior points: counts[a]
// += 1 and first[a] = min scan index for 1 <= a <= max_id, where the scan
// index runs k-major, then j, then i (the reference's loop order, so sorting
// ids by first[] reproduces its first-seen order).  Points with a larger id
// are counted in *above.  The 64 points of a wave are consecutive in scan
// order and usually hold one or two ids, so the lanes of one id add their
// count with a single atomic (per-point atomics on the few ids of a thin
// solid layer serialise on those addresses); the lowest such lane holds the
// smallest scan index.
__global__ __launch_bounds__(kBlock) void qoi_volume_kernel(const int* __restrict__ alpha, int fnx,
                                                            int fny, int fnz, int max_id,
                                                            int* counts, int* first,
                                                            unsigned long long* above)
{
    const int nx = fnx - 2, ny = fny - 2, nz = fnz - 2;
    const unsigned n = (unsigned)nx * ny * nz;
    const int lane = threadIdx.x & 63;
    unsigned long long over = 0ull;
    // wave-uniform trip count: every lane reaches the ballots
    for (unsigned base = blockIdx.x * kBlock + (threadIdx.x & ~63u); base < n;
         base += gridDim.x * kBlock) {
        const unsigned s = base + lane;
        int a = 0, k;
        if (s < n) a = alpha[interior_index(s, nx, ny, fnx, fny, &k)];
        over += a > max_id ? 1ull : 0ull;
        const bool valid = a > 0 && a <= max_id;
        unsigned long long pend = __ballot(valid);
        while (pend) {
            const int leader = __ffsll((long long)pend) - 1;
            const int id = __shfl(a, leader);
            const unsigned long long same = __ballot(valid && a == id) & pend;
            if (lane == leader) {
                atomicAdd(&counts[id], (int)__popcll(same));
                atomicMin(&first[id], (int)s);
            }
            pend &= ~same;
        }
    }
    if (over) atomicAdd(above, over);
}

// 3-D junctions: a solid interior point whose 19-point neighbourhood (offsets
// with di^2 + dj^2 + dk^2 <= 2, the point itself included) holds no liquid
// (alpha 0) and exactly 4 distinct ids, none more than 10 times.  Its row of
// 8 ints: i, j, k, largest count, the 4 ids ascending.
__device__ __forceinline__ bool junction_at(const int* __restrict__ alpha, int fnx, int fny, int nx,
                                            int ny, unsigned s, int (&o)[8])
{
    const int A = fnx * fny;
    const unsigned r = s / (unsigned)nx;
    const int i = 1 + (int)(s - r * (unsigned)nx);
    const int j = 1 + (int)(r % (unsigned)ny), k = 1 + (int)(r / (unsigned)ny);
    const int C = k * A + j * fnx + i;
    if (alpha[C] == 0) return false;
    int id[4] = {0, 0, 0, 0}, cnt[4] = {0, 0, 0, 0};
    int nd = 0;
    bool ok = true;
#pragma unroll
    for (int dk = -1; dk <= 1; ++dk)
#pragma unroll
        for (int dj = -1; dj <= 1; ++dj)
#pragma unroll
            for (int di = -1; di <= 1; ++di) {
                if (di * di + dj * dj + dk * dk > 2) continue;
                const int a = alpha[C + dk * A + dj * fnx + di];
                ok = ok && a != 0;
                bool hit = false;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const bool h = q < nd && id[q] == a;
                    cnt[q] += h ? 1 : 0;
                    hit = hit || h;
                }
                if (!hit) {
                    ok = ok && nd < 4;
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const bool t = q == nd;
                        id[q] = t ? a : id[q];
                        cnt[q] = t ? 1 : cnt[q];
                    }
                    nd = nd < 4 ? nd + 1 : nd;
                }
            }
    if (!ok || nd != 4) return false;
    int mx = cnt[0];
#pragma unroll
    for (int q = 1; q < 4; ++q) mx = cnt[q] > mx ? cnt[q] : mx;
    if (mx > 10) return false;
    // 4-element sorting network on the ids
    auto cs = [](int& x, int& y) {
        const int lo = x < y ? x : y, hi = x < y ? y : x;
        x = lo;
        y = hi;
    };
    cs(id[0], id[1]);
    cs(id[2], id[3]);
    cs(id[0], id[2]);
    cs(id[1], id[3]);
    cs(id[1], id[2]);
    o[0] = i;
    o[1] = j;
    o[2] = k;
    o[3] = mx;
    o[4] = id[0];
    o[5] = id[1];
    o[6] = id[2];
    o[7] = id[3];
    return true;
}

// The rows come out in the reference's scan order (k, then j, then i) without
// a sort: workgroup b tests the kJChunk consecutive scan indices from
// b * kJChunk, in kJPass passes of kBlock; a count pass, an exclusive scan of
// the workgroup counts, then a write pass that ranks each hit within its pass
// (wave ballots) after the hits of earlier passes and workgroups.
constexpr int kJPass = 8;
constexpr int kJChunk = kJPass * kBlock;
constexpr int kJScan = 1024;  // threads of the single-workgroup scan

__global__ __launch_bounds__(kBlock) void qoi_junction_count(const int* __restrict__ alpha, int fnx,
                                                             int fny, int fnz, unsigned* blk_cnt)
{
    __shared__ unsigned s_tot;
    const int nx = fnx - 2, ny = fny - 2;
    const unsigned n = (unsigned)nx * ny * (unsigned)(fnz - 2);
    if (threadIdx.x == 0) s_tot = 0u;
    __syncthreads();
    unsigned c = 0u;
    for (int p = 0; p < kJPass; ++p) {
        const unsigned s = blockIdx.x * (unsigned)kJChunk + p * kBlock + threadIdx.x;
        int o[8];
        c += (s < n && junction_at(alpha, fnx, fny, nx, ny, s, o)) ? 1u : 0u;
    }
    if (c) atomicAdd(&s_tot, c);
    __syncthreads();
    if (threadIdx.x == 0) blk_cnt[blockIdx.x] = s_tot;
}

__global__ __launch_bounds__(kJScan) void qoi_junction_scan(const unsigned* __restrict__ blk_cnt,
                                                            unsigned* blk_off, int nblk, unsigned* total)
{
    __shared__ unsigned part[kJScan];
    const int per = (nblk + kJScan - 1) / kJScan;
    const int b0 = min((int)threadIdx.x * per, nblk), b1 = min(b0 + per, nblk);
    unsigned own = 0u;
    for (int b = b0; b < b1; ++b) own += blk_cnt[b];
    part[threadIdx.x] = own;
    __syncthreads();
    for (int d = 1; d < kJScan; d <<= 1) {  // inclusive scan over the threads' sums
        const unsigned v = (int)threadIdx.x >= d ? part[threadIdx.x - d] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    unsigned run = part[threadIdx.x] - own;
    for (int b = b0; b < b1; ++b) {
        blk_off[b] = run;
        run += blk_cnt[b];
    }
    if (threadIdx.x == kJScan - 1) *total = part[kJScan - 1];
}

__global__ __launch_bounds__(kBlock) void qoi_junction_write(const int* __restrict__ alpha, int fnx,
                                                             int fny, int fnz,
                                                             const unsigned* __restrict__ blk_off,
                                                             int* rows, int cap)
{
    constexpr int kWaves = kBlock / 64;
    __shared__ unsigned s_w[kWaves];
    const int nx = fnx - 2, ny = fny - 2;
    const unsigned n = (unsigned)nx * ny * (unsigned)(fnz - 2);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    unsigned base = blk_off[blockIdx.x];
    for (int p = 0; p < kJPass; ++p) {
        const unsigned s = blockIdx.x * (unsigned)kJChunk + p * kBlock + threadIdx.x;
        int o[8];
        const bool hit = s < n && junction_at(alpha, fnx, fny, nx, ny, s, o);
        const unsigned long long m = __ballot(hit);
        if (lane == 0) s_w[w] = (unsigned)__popcll(m);
        __syncthreads();
        unsigned before = 0u, all = 0u;
#pragma unroll
        for (int u = 0; u < kWaves; ++u) {
            before += u < w ? s_w[u] : 0u;
            all += s_w[u];
        }
        if (hit) {
            const unsigned q = base + before + (unsigned)__popcll(m & ((1ull << lane) - 1ull));
            if (q < (unsigned)cap) {
                int* r = rows + 8 * (size_t)q;
#pragma unroll
                for (int u = 0; u < 8; ++u) r[u] = o[u];
            }
        }
        base += all;
        __syncthreads();  // s_w is rewritten by the next pass
    }
}

// Planes whose interior holds a liquid point (alpha 0): liq[k] = 1.  Every
// writer stores the same value, so plain stores suffice; one lane per plane
// present in the wave stores.
__global__ __launch_bounds__(kBlock) void qoi_liquid_planes_kernel(const int* __restrict__ alpha,
                                                                   int fnx, int fny, int fnz,
                                                                   unsigned* liq)
{
    const int nx = fnx - 2, ny = fny - 2, nz = fnz - 2;
    const unsigned n = (unsigned)nx * ny * nz;
    const int lane = threadIdx.x & 63;
    for (unsigned base = blockIdx.x * kBlock + (threadIdx.x & ~63u); base < n;
         base += gridDim.x * kBlock) {
        const unsigned s = base + lane;
        int k = 0;
        bool wet = false;
        if (s < n) wet = alpha[interior_index(s, nx, ny, fnx, fny, &k)] == 0;
        unsigned long long pend = __ballot(wet);
        while (pend) {
            const int leader = __ffsll((long long)pend) - 1;
            const int kl = __shfl(k, leader);
            const unsigned long long same = __ballot(wet && k == kl) & pend;
            if (lane == leader) liq[kl] = 1u;
            pend &= ~same;
        }
    }
}

// Line QoIs (QOI::calculateLineQoIs, 3DgrainMPI/QOI.cpp:71-90): per grain id
// 1..ng, interior points (total) and those above the tip plane (extra), with
// one atomic per id present in a wave as in qoi_volume_kernel.
__global__ __launch_bounds__(kBlock) void qoi_area_kernel(const int* __restrict__ alpha, int fnx,
                                                          int fny, int fnz, int ng, int tip,
                                                          int* total, int* extra)
{
    const int nx = fnx - 2, ny = fny - 2, nz = fnz - 2;
    const unsigned n = (unsigned)nx * ny * nz;
    const int lane = threadIdx.x & 63;
    for (unsigned base = blockIdx.x * kBlock + (threadIdx.x & ~63u); base < n;
         base += gridDim.x * kBlock) {
        const unsigned s = base + lane;
        int a = 0, k = 0;
        if (s < n) a = alpha[interior_index(s, nx, ny, fnx, fny, &k)];
        const bool valid = a > 0 && a <= ng;
        const unsigned long long up = __ballot(valid && k > tip);
        unsigned long long pend = __ballot(valid);
        while (pend) {
            const int leader = __ffsll((long long)pend) - 1;
            const int id = __shfl(a, leader);
            const unsigned long long same = __ballot(valid && a == id) & pend;
            if (lane == leader) {
                atomicAdd(&total[id - 1], (int)__popcll(same));
                const int e = (int)__popcll(same & up);
                if (e) atomicAdd(&extra[id - 1], e);
            }
            pend &= ~same;
        }
    }
}

// first[] starts at INT_MAX (no occurrence).
__global__ __launch_bounds__(kBlock) void qoi_fill_kernel(int* p, int n, int v)
{
    for (int q = blockIdx.x * kBlock + threadIdx.x; q < n; q += gridDim.x * kBlock) p[q] = v;
}

}  // namespace

hipError_t launch_qoi_volume(const int* alpha, int fnx, int fny, int fnz, int max_id, int* counts,
                             int* first, unsigned long long* above, hipStream_t stream)
{
    const int nb = max(1, min(2048, cdiv(max_id + 1, kBlock)));
    hipLaunchKernelGGL(qoi_fill_kernel, dim3(nb), dim3(kBlock), 0, stream, counts, max_id + 1, 0);
    hipLaunchKernelGGL(qoi_fill_kernel, dim3(nb), dim3(kBlock), 0, stream, first, max_id + 1,
                       0x7fffffff);
    hipLaunchKernelGGL(qoi_fill_kernel, dim3(1), dim3(kBlock), 0, stream, (int*)above, 2, 0);
    const long long n = (long long)(fnx - 2) * (fny - 2) * (fnz - 2);
    hipLaunchKernelGGL(qoi_volume_kernel, dim3(max(1, min(8192, cdiv(n, kBlock)))), dim3(kBlock), 0,
                       stream, alpha, fnx, fny, fnz, max_id, counts, first, above);
    return hipGetLastError();
}

hipError_t launch_qoi_liquid_planes(const int* alpha, int fnx, int fny, int fnz, unsigned* liq,
                                   hipStream_t stream)
{
    hipLaunchKernelGGL(qoi_fill_kernel, dim3(max(1, cdiv(fnz, kBlock))), dim3(kBlock), 0, stream,
                       (int*)liq, fnz, 0);
    const long long n = (long long)(fnx - 2) * (fny - 2) * (fnz - 2);
    hipLaunchKernelGGL(qoi_liquid_planes_kernel, dim3(max(1, min(8192, cdiv(n, kBlock)))),
                       dim3(kBlock), 0, stream, alpha, fnx, fny, fnz, liq);
    return hipGetLastError();
}

hipError_t launch_qoi_area(const int* alpha, int fnx, int fny, int fnz, int ng, int tip, int* total,
                           int* extra, hipStream_t stream)
{
    const int nb = max(1, min(2048, cdiv(ng, kBlock)));
    hipLaunchKernelGGL(qoi_fill_kernel, dim3(nb), dim3(kBlock), 0, stream, total, ng, 0);
    hipLaunchKernelGGL(qoi_fill_kernel, dim3(nb), dim3(kBlock), 0, stream, extra, ng, 0);
    const long long n = (long long)(fnx - 2) * (fny - 2) * (fnz - 2);
    hipLaunchKernelGGL(qoi_area_kernel, dim3(max(1, min(8192, cdiv(n, kBlock)))), dim3(kBlock), 0,
                       stream, alpha, fnx, fny, fnz, ng, tip, total, extra);
    return hipGetLastError();
}

int qoi_junction_blocks(int fnx, int fny, int fnz)
{
    return max(1, cdiv((long long)(fnx - 2) * (fny - 2) * (fnz - 2), kJChunk));
}

hipError_t launch_qoi_junctions(const int* alpha, int fnx, int fny, int fnz, unsigned* blk_cnt,
                                unsigned* blk_off, int* rows, int cap, unsigned* count,
                                hipStream_t stream)
{
    const int nblk = qoi_junction_blocks(fnx, fny, fnz);
    hipLaunchKernelGGL(qoi_junction_count, dim3(nblk), dim3(kBlock), 0, stream, alpha, fnx, fny, fnz,
                       blk_cnt);
    hipLaunchKernelGGL(qoi_junction_scan, dim3(1), dim3(kJScan), 0, stream, blk_cnt, blk_off, nblk,
                       count);
    if (cap > 0)
        hipLaunchKernelGGL(qoi_junction_write, dim3(nblk), dim3(kBlock), 0, stream, alpha, fnx, fny,
                           fnz, blk_off, rows, cap);
    return hipGetLastError();
}

}  // namespace cupf
