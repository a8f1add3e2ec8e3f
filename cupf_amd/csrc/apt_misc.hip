// apt_misc.hip -- the plan kernels, the batch advance kernel and the
// non-step kernels of the APT path (halo fill, collect, init, plane sums,
// moving frame, nucleation, activity map, copy), with their launchers.
#include "apt_device.h"

namespace cupf {

namespace {

// A plan workgroup appends all its listed chunks with one atomic (the list
// counter is a single word every workgroup hits): 4 waves for one simulation,
// 16 for a batch, whose plan has many more waves.
constexpr int kBatchPlanBlock = 1024;

// Work list of one step (runs before apt_step_tiled): waves over (z-chunk
// group, tiles), see plan_wave.  A tile chunk is listed only if some plane of
// it needs processing, with its plane masks.  Edge items are not listed: the
// first nzchunk * edge_items work indices are edge items, always taken (their
// points decide one by one), or the edge kernel's in overflow steps.
__global__ __launch_bounds__(kBlock) void apt_plan(StepArgs a)
{
    plan_block(a, (int)blockIdx.x);
}

// Batched work list: every simulation's plan waves in turn; plan_base counts
// waves.
__global__ __launch_bounds__(kBatchPlanBlock) void batch_plan(BatchArgs b)
{
    constexpr int kPlanWaves = kBatchPlanBlock / 64;
    __shared__ unsigned s_cnt[2 * kPlanWaves + 2];
    // wave-uniform index and member records read through the constant address
    // space (batch_advance rewrites them only after this kernel): the search
    // and the step-argument fields are scalar loads
    const int gw = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * kBatchPlanBlock + threadIdx.x) >> 6));
    typedef const __attribute__((address_space(4))) BatchSim CSim;
    const CSim* sims = (const CSim*)b.sims;
    unsigned* cnt = b.cnt + 4 * __atomic_load_n(b.cnt + 8, __ATOMIC_RELAXED);
    const int lane = threadIdx.x & 63;
    bool put = false, heavy = false;
    int item = 0, sim = 0;
    uint64_t need = 0ull, act = 0ull;
    if (gw < b.plan_total) {
        int lo = 0, hi = b.n - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (sims[mid].plan_base <= gw) lo = mid;
            else hi = mid - 1;
        }
        sim = lo;
        const StepArgs& a = *(const StepArgs*)&sims[sim].a;
        plan_wave(a, gw - sims[sim].plan_base, lane, put, item, need, act);
        heavy = plan_heavy(need, a.kchunk);
    }
    const unsigned q = plan_slot<kPlanWaves>(put, heavy, cnt, cnt + 2, (unsigned)b.wl_cap, s_cnt);
    if (put) {
        b.wl_sim[q] = sim;
        b.wl_item[q] = item;
        b.wl_need[q] = need;
        b.wl_act[q] = act;
    }
}

// After a batched step: every simulation moves to its next step the way the
// host does between single launches (buffers swap, the destination takes the
// step's epoch, the step index and t advance, the next of the member's three
// counter sets is cleared for the next step); the batch's own two list
// counter sets alternate.
__global__ __launch_bounds__(64) void batch_advance(BatchArgs b)
{
    const int s = blockIdx.x * 64 + threadIdx.x;
    if (s == 0) {
        const unsigned p = b.cnt[8];
        b.cnt[4 * (1 - p)] = 0u;
        b.cnt[4 * (1 - p) + 1] = 0u;
        b.cnt[4 * (1 - p) + 2] = 0u;
        b.cnt[8] = 1u - p;
    }
    if (s >= b.n) return;
    BatchSim& m = b.sims[s];
    m.ep[1 - m.cur] = m.a.ep_new;
    m.cur = 1 - m.cur;
    m.step += 1;
    m.parity = (m.parity + 1) % 3;
    unsigned* next = m.cnt + 4 * m.parity;
    next[0] = 0u;
    next[1] = 0u;
    next[2] = 0u;
    next[3] = 0u;
    StepArgs& a = m.a;
    a.ph = m.ph[m.cur];
    a.arg = m.arg[m.cur];
    a.ph_new = m.ph[1 - m.cur];
    a.arg_new = m.arg[1 - m.cur];
    a.act_src = m.act[m.cur];
    a.act_dst = m.act[1 - m.cur];
    a.ep_src = m.ep[m.cur];
    a.ep_dst = m.ep[1 - m.cur];
    m.ep_ctr += 1u;
    a.ep_new = m.ep_ctr;
    a.fix_count = next;
    a.cnt_next = m.cnt + 4 * ((m.parity + 1) % 3);
    a.wl_count = next + 2;
    a.wl_light = next + 1;
    // t_cur_step = (2*kt+1)*dt*tau0 in float (3DgrainMPI/APTPhaseField.cu:863,919)
    a.t = (float)m.step * a.dt * m.tau0;
}

// Halo fill (APTsetBC3D, 3DgrainMPI/APTPhaseField.cu:93-174).  Each halo cell
// is written exactly once, from its per-axis mapped interior point, so the
// result is deterministic at edges and corners (the reference races there)
// and equals a z-then-y-then-x sequential face fill.
__global__ __launch_bounds__(kBlock) void bc_fill(float* ph, int* arg, int fnx, int fny, int fnz,
                                                  int L, int bcx, int bcy, int bcz, int nhalo)
{
    int h = blockIdx.x * kBlock + threadIdx.x;
    if (h >= nhalo) return;
    const int s = blockIdx.y;
    const int A = fnx * fny;
    int i, j, k;
    if (h < 2 * A) {
        k = h < A ? 0 : fnz - 1;
        const int r = h < A ? h : h - A;
        j = r / fnx;
        i = r - j * fnx;
    } else {
        h -= 2 * A;
        const int By = fnx * (fnz - 2);
        if (h < 2 * By) {
            j = h < By ? 0 : fny - 1;
            const int r = h < By ? h : h - By;
            k = 1 + r / fnx;
            i = r - (k - 1) * fnx;
        } else {
            h -= 2 * By;
            const int Bx = (fny - 2) * (fnz - 2);
            i = h < Bx ? 0 : fnx - 1;
            const int r = h < Bx ? h : h - Bx;
            k = 1 + r / (fny - 2);
            j = 1 + (r - (k - 1) * (fny - 2));
        }
    }
    const int src = bc_map(i, fnx, bcx) + bc_map(j, fny, bcy) * fnx + bc_map(k, fnz, bcz) * A + s * L;
    const int dst = i + j * fnx + k * A + s * L;
    ph[dst] = ph[src];
    arg[dst] = arg[src];
}

// APTcollect_PF (3DgrainMPI/APTPhaseField.cu:55-75): argmax over slots with a
// strict > from slot 0; written only if the max exceeds LS (double compare).
__global__ __launch_bounds__(kBlock) void collect_kernel(const float* __restrict__ ph,
                                                         const int* __restrict__ arg,
                                                         float* __restrict__ phi,
                                                         int* __restrict__ alpha, int L, int K,
                                                         int clear_alpha)
{
    const int C = blockIdx.x * kBlock + threadIdx.x;
    if (C >= L) return;
    int am = 0;
    float mv = ph[C];
    for (int s = 1; s < K; ++s) {
        const float v = ph[C + s * L];
        if (v > mv) {
            mv = v;
            am = s;
        }
    }
    if ((double)mv > kLS) {
        phi[C] = mv;
        alpha[C] = arg[C + am * L];
    } else if (clear_alpha) {
        alpha[C] = 0;
    }
}

// set_minus1 + memset(-1) + APTini_PF on both buffers
// (3DgrainMPI/APTPhaseField.cu:78-89,727-728,817-818).
__global__ __launch_bounds__(kBlock) void ini_kernel(float* ph0, int* arg0, float* ph1, int* arg1,
                                                     const float* __restrict__ phi0,
                                                     const int* __restrict__ alpha0, int L, int K)
{
    const int q = blockIdx.x * kBlock + threadIdx.x;
    if (q >= L * K) return;
    float v = -1.0f;
    int id = -1;
    if (q < L && (double)phi0[q] > kLS) {
        v = phi0[q];
        id = alpha0[q];
    }
    ph0[q] = v;
    arg0[q] = id;
    ph1[q] = v;
    arg1[q] = id;
}

// ave_x (3DgrainMPI/PhaseField.cu:366-378) as an exact fixed-point sum:
// every phi is rounded to a multiple of 2^-40 and summed in int64, which is
// independent of summation order (the reference's float atomics are not).
__global__ __launch_bounds__(kBlock) void plane_sum_kernel(const float* __restrict__ ph,
                                                           long long* sums, int A, int L, int K,
                                                           int k_from)
{
    __shared__ long long red[kBlock];
    const int k = k_from + (int)blockIdx.y;
    long long acc = 0;
    for (int s = 0; s < K; ++s)
        for (int q = blockIdx.x * kBlock + threadIdx.x; q < A; q += gridDim.x * kBlock)
            acc += (long long)rint((double)ph[s * L + k * A + q] * 1099511627776.0);
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd((unsigned long long*)&sums[k], (unsigned long long)red[0]);
}

// APTmove_frame, first half (3DgrainMPI/APTPhaseField.cu:357-391): shift every
// slot of the interior down one plane into the other buffer (top plane
// becomes -1/-1), shifted z into zbuf, archive alpha plane 1, count loss area.
__global__ __launch_bounds__(kBlock) void move_frame_kernel(
    const float* __restrict__ ph, const int* __restrict__ arg, float* __restrict__ phb,
    int* __restrict__ argb, const float* __restrict__ z, float* __restrict__ zb,
    const int* __restrict__ alpha, int* __restrict__ alpha_full, int* __restrict__ loss,
    int move_count, int fnx, int fny, int fnz, int L, int K, int num_theta)
{
    const long long q = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (q >= (long long)L * K) return;
    const int s = (int)(q / L);
    const int C = (int)(q - (long long)s * L);
    const int A = fnx * fny;
    const int k = C / A;
    const int r = C - k * A;
    const int j = r / fnx;
    const int i = r - j * fnx;
    if (i == 0 && j == 0 && s == 0) {
        if (k > 0 && k < fnz - 2) zb[k] = z[k + 1];
        if (k == fnz - 2) zb[k] = 2 * z[fnz - 2] - z[fnz - 3];
    }
    if (i > 0 && i < fnx - 1 && j > 0 && j < fny - 1) {
        if (k > 0 && k < fnz - 2) {
            phb[q] = ph[q + A];
            argb[q] = arg[q + A];
        } else if (k == fnz - 2) {
            phb[q] = -1.0f;
            argb[q] = -1;
        }
    }
    if (k == 1 && s == 0) {
        const int al = alpha[C];
        alpha_full[(long long)move_count * A + C] = al;
        if (al >= 1 && al <= num_theta) atomicAdd(loss + al - 1, 1);
    }
}

// APTcopy_frame (3DgrainMPI/APTPhaseField.cu:393-410).
__global__ __launch_bounds__(kBlock) void copy_frame_kernel(
    const float* __restrict__ phb, const int* __restrict__ argb, const float* __restrict__ zb,
    float* __restrict__ ph, int* __restrict__ arg, float* __restrict__ z, int fnx, int fny,
    int fnz, int L, int K)
{
    const long long q = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (q >= (long long)L * K) return;
    const int s = (int)(q / L);
    const int C = (int)(q - (long long)s * L);
    const int A = fnx * fny;
    const int k = C / A;
    const int r = C - k * A;
    const int j = r / fnx;
    const int i = r - j * fnx;
    if (i == 0 && j == 0 && s == 0 && k > 0 && k < fnz - 1) z[k] = zb[k];
    if (i > 0 && i < fnx - 1 && j > 0 && j < fny - 1 && k > 0 && k < fnz - 1) {
        ph[q] = phb[q];
        arg[q] = argb[q];
    }
}

// init_nucl_status (3DgrainMPI/APTPhaseField.cu:420-444), block centre taken
// without the halo offset exactly as the reference does.
__global__ __launch_bounds__(kBlock) void nucl_status_kernel(const float* __restrict__ phi0,
                                                             int* status, int cnx, int cny,
                                                             int cnz, int p, int fnx, int fny)
{
    const int C = blockIdx.x * kBlock + threadIdx.x;
    if (C >= cnx * cny * cnz) return;
    const int k = C / (cnx * cny);
    const int r = C - k * cnx * cny;
    const int j = r / cnx;
    const int i = r - j * cnx;
    const int w = 2 * p + 1;
    const int g = (w * k + p) * fnx * fny + (w * j + p) * fnx + (w * i + p);
    status[C] = ((double)phi0[g] > kLS) ? 1 : 0;
}

// splitmix64 finaliser (Steele, Lea & Flood, OOPSLA 2014).
__device__ __forceinline__ uint64_t mix64(uint64_t v)
{
    v = (v ^ (v >> 30)) * 0xbf58476d1ce4e5b9ULL;
    v = (v ^ (v >> 27)) * 0x94d049bb133111ebULL;
    return v ^ (v >> 31);
}

// nuncl_possibility (3DgrainMPI/APTPhaseField.cu:446-455).
__device__ __forceinline__ float nucl_prob(const NuclArgs& a, float delT, float d_delT, float nmax)
{
    float slope = -0.5f * (delT - a.undcool_mean) * (delT - a.undcool_mean) / a.undcool_std
                / a.undcool_std;
    slope = expf(slope);
    float density = nmax / (sqrtf((float)(2.0f * M_PI)) * a.undcool_std) * slope * d_delT;
    return 8.0f * a.nuc_rad * a.nuc_rad * a.nuc_rad * density;
}

// add_nucl (3DgrainMPI/APTPhaseField.cu:459-536) with a counter-based draw per
// (seed, coarse cell, call) instead of per-cell cuRAND state.
__global__ __launch_bounds__(kBlock) void nucl_kernel(NuclArgs a)
{
    const int C = blockIdx.x * kBlock + threadIdx.x;
    if (C >= a.cnx * a.cny * a.cnz) return;
    if (a.status[C] != 0) return;
    const int k = C / (a.cnx * a.cny);
    const int r = C - k * a.cnx * a.cny;
    const int j = r / a.cnx;
    const int i = r - j * a.cnx;
    const int w = 2 * a.p + 1;
    const int gi = w * i + a.p + 1, gj = w * j + a.p + 1, gk = w * k + a.p + 1;
    const int A = a.fnx * a.fny;
    const int gC = gk * A + gj * a.fnx + gi;
    int st = 0;
    for (int s = 0; s < a.K; ++s)
        if ((double)a.ph[gC + s * a.L] > kLS) st = 1;
    if (st) {
        a.status[C] = 1;
        return;
    }
    float posb;
    if (a.use_iuc) {
        float delT = a.uc0 * 1e6 * (a.t + a.dt);
        float d_delT = a.uc0 * 1e6 * a.dt;
        posb = nucl_prob(a, delT, d_delT, a.nuc_Nmax0);
    } else {
        const float xx = a.x[gi] - a.X0_mac, yy = a.y[gj] - a.Y0_mac, zz = a.z[gk] - a.Z0_mac;
        float Tc = interp4(a.T3d, xx, yy, zz, a.t - a.t0_mac, a.Nx, a.Ny, a.Nz, a.Nt, a.Dx_mac,
                           a.Dt_mac);
        float Tcd = interp4(a.T3d, xx, yy, zz, a.t + a.dt - a.t0_mac, a.Nx, a.Ny, a.Nz, a.Nt,
                            a.Dx_mac, a.Dt_mac);
        posb = nucl_prob(a, -Tcd, Tc - Tcd, a.nuc_Nmax);
    }
    uint64_t h = mix64(a.seed + 0x9e3779b97f4a7c15ULL);
    h = mix64(h ^ ((uint64_t)C * 0xd1b54a32d192ed03ULL));
    h = mix64(h ^ (a.call * 0x8cb92ba72f3d8dd7ULL + 0x632be59bd9b4e019ULL));
    const float u = (float)((h >> 40) + 1) * (1.0f / 16777216.0f);
    const int id = 1 + (int)(((h & 0xffffffffULL) * 0x7ffffffeULL) >> 32);
    if (u < posb) {
        int q = 0;
        for (int lk = -a.p; lk <= a.p; ++lk)
            for (int lj = -a.p; lj <= a.p; ++lj)
                for (int li = -a.p; li <= a.p; ++li) {
                    const int o = gC + lk * A + lj * a.fnx + li;
                    const float v = a.ball[q++];
                    a.ph[o] = v;
                    if (v > -0.9999f) a.arg[o] = id;
                    a.act[(gk + lk) * a.ntiles + act_tile_x(gi + li) + act_tile_y(gj + lj) * a.ntx] =
                        a.ep;
                }
        a.status[C] = 1;
        atomicAdd(a.events, 1ull);
    }
}

// Activity map from a buffer's contents (after it was written other than by
// a step): one lane per interior point, K slots each.
__global__ __launch_bounds__(kBlock) void act_rebuild_kernel(const float* __restrict__ ph,
                                                             const int* __restrict__ arg,
                                                             unsigned* act, unsigned ep, int fnx,
                                                             int fny, int fnz, int K, int ntx,
                                                             int ntiles)
{
    const int nx = fnx - 2, ny = fny - 2;
    const long long q = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (q >= (long long)nx * ny * (fnz - 2)) return;
    const int k = 1 + (int)(q / ((long long)nx * ny));
    const int r = (int)(q - (long long)(k - 1) * nx * ny);
    const int j = 1 + r / nx;
    const int i = 1 + r - (j - 1) * nx;
    const int L = fnx * fny * fnz;
    const int C = i + j * fnx + k * fnx * fny;
    bool on = false;
    for (int s = 0; s < K; ++s) on = on || arg[C + s * L] != -1 || ph[C + s * L] != -1.0f;
    if (on) act[k * ntiles + act_tile_x(i) + act_tile_y(j) * ntx] = ep;
}

// Plain streaming copy (16 B per lane per iteration), the reference point for
// HBM bandwidth and for scaling the FETCH_SIZE / WRITE_SIZE counters.
__global__ __launch_bounds__(kBlock) void copy_kernel(const float4* __restrict__ src,
                                                      float4* __restrict__ dst, long long n)
{
    for (long long q = (long long)blockIdx.x * kBlock + threadIdx.x; q < n;
         q += (long long)gridDim.x * kBlock)
        dst[q] = src[q];
}

}  // namespace

hipError_t launch_plan(const StepArgs& a, hipStream_t stream)
{
    dim3 pgrid(max(1, cdiv(plan_waves(a.kchunk, a.nzchunk, a.items_chunk), kBlock / 64)));
    hipLaunchKernelGGL(apt_plan, pgrid, dim3(kBlock), 0, stream, a);
    return hipGetLastError();
}

hipError_t launch_batch_plan(const BatchArgs& b, hipStream_t stream)
{
    hipLaunchKernelGGL(batch_plan, dim3(max(1, cdiv(b.plan_total, kBatchPlanBlock / 64))),
                       dim3(kBatchPlanBlock), 0,
                       stream, b);
    return hipGetLastError();
}

hipError_t launch_batch_advance(const BatchArgs& b, hipStream_t stream)
{
    hipLaunchKernelGGL(batch_advance, dim3(max(1, cdiv(b.n, 64))), dim3(64), 0, stream, b);
    return hipGetLastError();
}

hipError_t launch_step(const StepArgs& a, int K, int gather27, hipStream_t stream,
                       const EdgeLaunch* edge, hipEvent_t ev_begin, hipEvent_t ev_planned,
                       hipEvent_t ev_mid, hipEvent_t ev_end, bool planned, const StepArgs* next)
{
    if (K == 5)
        return gather27 ? launch_step_k5_g27(a, stream, edge, ev_begin, ev_planned, ev_mid, ev_end, planned, next)
                        : launch_step_k5_g7(a, stream, edge, ev_begin, ev_planned, ev_mid, ev_end, planned, next);
    if (K == 8)
        return gather27 ? launch_step_k8_g27(a, stream, edge, ev_begin, ev_planned, ev_mid, ev_end, planned, next)
                        : launch_step_k8_g7(a, stream, edge, ev_begin, ev_planned, ev_mid, ev_end, planned, next);
    return hipErrorInvalidValue;
}

hipError_t launch_batch_step(const BatchArgs& b, int K, int gather27, hipStream_t stream)
{
    if (K == 5) return gather27 ? launch_batch_k5_g27(b, stream) : launch_batch_k5_g7(b, stream);
    if (K == 8) return gather27 ? launch_batch_k8_g27(b, stream) : launch_batch_k8_g7(b, stream);
    return hipErrorInvalidValue;
}

static int slots(int K, int gather27, int device, int which)
{
    if (K == 5) return gather27 ? step_slots_k5_g27(device, which) : step_slots_k5_g7(device, which);
    if (K == 8) return gather27 ? step_slots_k8_g27(device, which) : step_slots_k8_g7(device, which);
    return 0;
}

int step_kernel_slots(int K, int gather27, int device) { return slots(K, gather27, device, 0); }

int batch_kernel_slots(int K, int gather27, int device) { return slots(K, gather27, device, 1); }

int fixup_kernel_slots(int K, int gather27, int device, bool batch)
{
    return slots(K, gather27, device, batch ? 3 : 2);
}

hipError_t launch_act_rebuild(const float* ph, const int* arg, unsigned* act, unsigned ep, int fnx,
                              int fny, int fnz, int K, int ntx, int ntiles, hipStream_t stream)
{
    const long long n = (long long)(fnx - 2) * (fny - 2) * (fnz - 2);
    hipLaunchKernelGGL(act_rebuild_kernel, dim3(cdiv(n, kBlock)), dim3(kBlock), 0, stream, ph, arg,
                       act, ep, fnx, fny, fnz, K, ntx, ntiles);
    return hipGetLastError();
}

hipError_t launch_copy(const void* src, void* dst, size_t bytes, hipStream_t stream)
{
    const long long n = (long long)(bytes / 16);
    hipLaunchKernelGGL(copy_kernel, dim3(256 * 16), dim3(kBlock), 0, stream,
                       (const float4*)src, (float4*)dst, n);
    return hipGetLastError();
}

hipError_t launch_bc_fill(float* ph, int* arg, int fnx, int fny, int fnz, int K, int bcx, int bcy,
                          int bcz, hipStream_t stream)
{
    const int nhalo = 2 * fnx * fny + 2 * fnx * (fnz - 2) + 2 * (fny - 2) * (fnz - 2);
    dim3 grid(cdiv(nhalo, kBlock), K);
    hipLaunchKernelGGL(bc_fill, grid, dim3(kBlock), 0, stream, ph, arg, fnx, fny, fnz,
                       fnx * fny * fnz, bcx, bcy, bcz, nhalo);
    return hipGetLastError();
}

hipError_t launch_collect(const float* ph, const int* arg, float* phi, int* alpha, int L, int K,
                          int clear_alpha, hipStream_t stream)
{
    hipLaunchKernelGGL(collect_kernel, dim3(cdiv(L, kBlock)), dim3(kBlock), 0, stream, ph, arg,
                       phi, alpha, L, K, clear_alpha);
    return hipGetLastError();
}

hipError_t launch_ini(float* ph0, int* arg0, float* ph1, int* arg1, const float* phi0,
                      const int* alpha0, int L, int K, hipStream_t stream)
{
    hipLaunchKernelGGL(ini_kernel, dim3(cdiv((long long)L * K, kBlock)), dim3(kBlock), 0, stream,
                       ph0, arg0, ph1, arg1, phi0, alpha0, L, K);
    return hipGetLastError();
}

hipError_t launch_plane_sums(const float* ph, long long* sums, int fnx, int fny, int fnz, int K,
                             int k_from, int k_to, hipStream_t stream)
{
    const int A = fnx * fny;
    if (k_from < 0 || k_to > fnz || k_to <= k_from) return hipErrorInvalidValue;
    hipError_t e = hipMemsetAsync(sums + k_from, 0, sizeof(long long) * (k_to - k_from), stream);
    if (e != hipSuccess) return e;
    dim3 grid(min(cdiv(A, kBlock), 64), k_to - k_from);
    hipLaunchKernelGGL(plane_sum_kernel, grid, dim3(kBlock), 0, stream, ph, sums, A, A * fnz, K,
                       k_from);
    return hipGetLastError();
}

hipError_t launch_move_frame(float* ph_old, int* arg_old, float* ph_new, int* arg_new, float* z,
                             float* zbuf, const int* alpha, int* alpha_full, int* loss_area,
                             int move_count, int fnx, int fny, int fnz, int K, int num_theta,
                             hipStream_t stream)
{
    const int L = fnx * fny * fnz;
    dim3 grid(cdiv((long long)L * K, kBlock));
    hipLaunchKernelGGL(move_frame_kernel, grid, dim3(kBlock), 0, stream, ph_old, arg_old, ph_new,
                       arg_new, z, zbuf, alpha, alpha_full, loss_area, move_count, fnx, fny, fnz,
                       L, K, num_theta);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(copy_frame_kernel, grid, dim3(kBlock), 0, stream, ph_new, arg_new, zbuf,
                       ph_old, arg_old, z, fnx, fny, fnz, L, K);
    return hipGetLastError();
}

hipError_t launch_nucl_status(const float* phi0, int* status, int cnx, int cny, int cnz, int p,
                              int fnx, int fny, hipStream_t stream)
{
    const int n = cnx * cny * cnz;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(nucl_status_kernel, dim3(cdiv(n, kBlock)), dim3(kBlock), 0, stream, phi0,
                       status, cnx, cny, cnz, p, fnx, fny);
    return hipGetLastError();
}

hipError_t launch_nucl(const NuclArgs& a, hipStream_t stream)
{
    const int n = a.cnx * a.cny * a.cnz;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(nucl_kernel, dim3(cdiv(n, kBlock)), dim3(kBlock), 0, stream, a);
    return hipGetLastError();
}

}  // namespace cupf
