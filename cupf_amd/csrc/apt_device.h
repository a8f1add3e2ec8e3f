// apt_device.h -- device code of the APT step shared by the kernel translation
// units (included by apt_misc.hip and step_k*.hip; internal linkage).
//
// Field layout (3DgrainMPI/devicefunc.cu_inl:24-29): SoA slot planes,
// element (i,j,k,slot) at i + j*fnx + k*fnx*fny + slot*L, phi float32 and
// grain id int32 in separate arrays, -1 = empty slot.
//
// Arithmetic follows the reference source operation by operation (build with
// -ffp-contract=off): float everywhere except the reference's implicit double
// promotions (the 1e6 thermal factors, LS=-0.995 and the -0.9999 kill test).
#pragma once

#include "apt_internal.h"

namespace cupf {

namespace {

constexpr int kBlock = 256;        // 4 wave64 per workgroup
constexpr int kFixLanes = 256;     // fix-up workgroup

// Halo index -> interior source index along one axis (APTsetBC3D,
// 3DgrainMPI/APTPhaseField.cu:111-112,134-135,157-158): no-flux mirrors
// out[0]=in[2], out[n-1]=in[n-3]; periodic wraps out[0]=in[n-2], out[n-1]=in[1].
__device__ __forceinline__ int bc_map(int q, int n, int periodic)
{
    if (q == 0) return periodic ? n - 2 : 2;
    if (q == n - 1) return periodic ? 1 : n - 3;
    return q;
}

// kine_ani (3DgrainMPI/APTPhaseField.cu:36-52) with a row-major rotation;
// each row is (m0*ux + m1*uy) + m2*uz, which equals the reference's angle
// form bit for bit (a - b*c == a + (-b)*c; the "+ 0" only flips signs of zeros
// that are squared away).
__device__ __forceinline__ float kine_ani(float ux, float uy, float uz, float m0, float m1,
                                          float m2, float m3, float m4, float m5, float m6,
                                          float m7, float m8, float a_s, float epsv, float eps)
{
    float ux2 = m0 * ux + m1 * uy + m2 * uz;
    ux2 = ux2 * ux2;
    float uy2 = m3 * ux + m4 * uy + m5 * uz;
    uy2 = uy2 * uy2;
    float uz2 = m6 * ux + m7 * uy + m8 * uz;
    uz2 = uz2 * uz2;
    float mag = ux2 + uy2 + uz2;
    float mag2 = mag * mag;
    if (mag > eps) return a_s * (1.0f + epsv * (ux2 * ux2 + uy2 * uy2 + uz2 * uz2) / mag2);
    return 1.0f;
}

// interp3Dtemperature (3DgrainMPI/devicefunc.cu_inl:85-125), indices clamped
// into the macro grid (the reference reads out of bounds there).
__device__ __forceinline__ float interp3(const float* __restrict__ u, float x, float y, float z,
                                         int Nx, int Ny, int Nz, float Dx)
{
    int kx = (int)(x / Dx), ky = (int)(y / Dx), kz = (int)(z / Dx);
    float dx = x / Dx - kx, dy = y / Dx - ky, dz = z / Dx - kz;
    if (kx == Nx - 1) { kx = Nx - 2; dx = 1.0f; }
    if (ky == Ny - 1) { ky = Ny - 2; dy = 1.0f; }
    if (kz == Nz - 1) { kz = Nz - 2; dz = 1.0f; }
    kx = min(max(kx, 0), Nx - 2);
    ky = min(max(ky, 0), Ny - 2);
    kz = min(max(kz, 0), Nz - 2);
    int o = kx + ky * Nx + kz * Nx * Ny;
    int on = kx + ky * Nx + (kz + 1) * Nx * Ny;
    return ((1.0f - dx) * (1.0f - dy) * u[o] + (1.0f - dx) * dy * u[o + Nx]
            + dx * (1.0f - dy) * u[o + 1] + dx * dy * u[o + Nx + 1]) * (1.0f - dz)
         + ((1.0f - dx) * (1.0f - dy) * u[on] + (1.0f - dx) * dy * u[on + Nx]
            + dx * (1.0f - dy) * u[on + 1] + dx * dy * u[on + Nx + 1]) * dz;
}

// interp4Dtemperature (3DgrainMPI/devicefunc.cu_inl:32-83), clamped likewise.
__device__ __forceinline__ float interp4(const float* __restrict__ u, float x, float y, float z,
                                         float t, int Nx, int Ny, int Nz, int Nt, float Dx,
                                         float Dt)
{
    int kx = (int)(x / Dx), ky = (int)(y / Dx), kz = (int)(z / Dx), kt = (int)(t / Dt);
    float dx = x / Dx - kx, dy = y / Dx - ky, dz = z / Dx - kz, dt = t / Dt - kt;
    if (kx == Nx - 1) { kx = Nx - 2; dx = 1.0f; }
    if (ky == Ny - 1) { ky = Ny - 2; dy = 1.0f; }
    if (kz == Nz - 1) { kz = Nz - 2; dz = 1.0f; }
    if (kt == Nt - 1) { kt = Nt - 2; dt = 1.0f; }
    kx = min(max(kx, 0), Nx - 2);
    ky = min(max(ky, 0), Ny - 2);
    kz = min(max(kz, 0), Nz - 2);
    kt = min(max(kt, 0), Nt - 2);
    const int NN = Nx * Ny * Nz;
    int o = kx + ky * Nx + kz * Nx * Ny + kt * NN;
    int on = kx + ky * Nx + (kz + 1) * Nx * Ny + kt * NN;
    float up = ((1.0f - dx) * (1.0f - dy) * u[o] + (1.0f - dx) * dy * u[o + Nx]
                + dx * (1.0f - dy) * u[o + 1] + dx * dy * u[o + Nx + 1]) * (1.0f - dz)
             + ((1.0f - dx) * (1.0f - dy) * u[on] + (1.0f - dx) * dy * u[on + Nx]
                + dx * (1.0f - dy) * u[on + 1] + dx * dy * u[on + Nx + 1]) * dz;
    o += NN;
    on += NN;
    float un = ((1.0f - dx) * (1.0f - dy) * u[o] + (1.0f - dx) * dy * u[o + Nx]
                + dx * (1.0f - dy) * u[o + 1] + dx * dy * u[o + Nx + 1]) * (1.0f - dz)
             + ((1.0f - dx) * (1.0f - dy) * u[on] + (1.0f - dx) * dy * u[on + Nx]
                + dx * (1.0f - dy) * u[on + 1] + dx * dy * u[on + Nx + 1]) * dz;
    return (1.0f - dt) * up + dt * un;
}

// Thermal value of one point at coordinates (xi, yj, zk)
// (3DgrainMPI/APTPhaseField.cu:278-327; type 5: 3Dgrain/APTPhaseField.cu:303).
// Returns true where the reference `continue`s past the slot, leaving the
// destination untouched.
__device__ __forceinline__ bool thermal_Tv(const StepArgs& a, float xi, float yj, float zk, float& T)
{
    const float t = a.t;
    if (a.use_iuc) {
        T = -a.uc0 * 1e6 * t;
        return false;
    }
    switch (a.thermal_type) {
    case 1:
        T = a.G * (zk - a.z0) - a.uc * 1e6 * t;
        return false;
    case 2: {
        float dist = sqrtf((yj - 0.5f * a.lyd) * (yj - 0.5f * a.lyd)
                           + (zk - a.z0 - a.lzd) * (zk - a.z0 - a.lzd)) - a.r0;
        T = -a.G * dist - a.uc * 1e6 * t;
        return false;
    }
    case 3: {
        float lm = (a.r0 - a.z0) / a.tan_angle;
        float x_start = 0;
        float xl = xi - x_start;
        float r0_x = a.z0 + (a.r0 - a.z0) * xl / lm;
        float z_tilt = a.z0 + a.lzd - zk;
        float dist = a.cos_angle * (sqrtf((yj - 0.5f * a.lyd) * (yj - 0.5f * a.lyd)
                                          + z_tilt * z_tilt) - r0_x);
        if (xl > lm + a.V * t * 1e6) return true;
        T = -a.G * dist - a.uc * 1e6 * t;
        return false;
    }
    case 4: {
        float xq = xi - a.V * t * 1e6;
        if (xq > a.mp_len + a.x0 || xq < 0.0f) return true;
        float v = interp3(a.psi_mac, xq - a.X0_mac, yj - a.Y0_mac, zk - a.Z0_mac, a.Nx, a.Ny,
                          a.Nz, a.Dx_mac);
        T = (a.order == 100) ? -v : -a.G * v;
        return false;
    }
    case 5:
        T = a.G * (zk - a.R * 1e6 * t - 2);
        return false;
    default:
        T = 0.0f;
        return false;
    }
}

// The same at grid point (i, j, k), coordinates read from the x/y/z arrays.
__device__ __forceinline__ bool thermal_T(const StepArgs& a, int i, int j, int k, float& T)
{
    const int tt = a.thermal_type;
    const bool nx = !a.use_iuc && (tt == 3 || tt == 4);
    const bool ny = !a.use_iuc && (tt >= 2 && tt <= 4);
    return thermal_Tv(a, nx ? a.x[i] : 0.0f, ny ? a.y[j] : 0.0f, a.z[k], T);
}

// Gather state of one point.  la: slot ids in first-appearance order (no
// holes, n filled).  pm[s]: 4-bit nibble q holds the neighbour slot p whose id
// landed in slot q at stencil row s (D,B,L,C,R,T,U), 15 = none; the last
// matching p wins, as the reference's overwrite does.
template <int K>
struct Gather {
    int la[K];
    uint32_t pm[7];
    int n;
    bool ovf;
};

template <int K>
__device__ __forceinline__ void gather_init(Gather<K>& g)
{
#pragma unroll
    for (int q = 0; q < K; ++q) g.la[q] = -1;
#pragma unroll
    for (int s = 0; s < 7; ++s) g.pm[s] = 0xffffffffu;
    g.n = 0;
    g.ovf = false;
}

// Pass A for one neighbour point (3DgrainMPI/APTPhaseField.cu:215-232): load
// its K ids and probe-insert them.  la has no holes, so the probe returns the
// matching slot or slot n.  S is the stencil row (-1: 27-point extra point).
template <int K, int S>
__device__ __forceinline__ void scan_point(const int* __restrict__ arg, int base, int L,
                                           Gather<K>& g)
{
    // the K loads go out together: a load inside the probe loop would sit
    // behind the previous probe's branches and wait out the memory latency
    // once per slot
    int ids[K];
#pragma unroll
    for (int p = 0; p < K; ++p) ids[p] = arg[base + p * L];
#pragma unroll
    for (int p = 0; p < K; ++p) {
        const int id = ids[p];
        if (id == -1) continue;
        int pos = g.n;
#pragma unroll
        for (int q = 0; q < K; ++q) pos = (g.la[q] == id) ? q : pos;
        if (pos >= K) {
            g.ovf = true;
            continue;
        }
        if (pos == g.n) {  // a new id takes the next free slot
#pragma unroll
            for (int q = 0; q < K; ++q) g.la[q] = (q == pos) ? id : g.la[q];
            ++g.n;
        }
        if (S >= 0) {
            const uint32_t sh = 4u * (uint32_t)pos;
            g.pm[S] = (g.pm[S] & ~(15u << sh)) | ((uint32_t)p << sh);
        }
    }
    // one neighbour point's loads in flight per lane; occupancy supplies the
    // rest of the memory parallelism
    __builtin_amdgcn_sched_barrier(0);
}

// Neighbour bases of an interior point through the per-axis halo maps.
struct Nbr {
    int xs[3], ys[3], zs[3];  // (minus, centre, plus) per axis, already mapped
};

__device__ __forceinline__ Nbr make_nbr(const StepArgs& a, int i, int j, int k)
{
    Nbr n;
    n.xs[0] = bc_map(i - 1, a.fnx, a.bcx); n.xs[1] = i; n.xs[2] = bc_map(i + 1, a.fnx, a.bcx);
    n.ys[0] = bc_map(j - 1, a.fny, a.bcy); n.ys[1] = j; n.ys[2] = bc_map(j + 1, a.fny, a.bcy);
    n.zs[0] = bc_map(k - 1, a.fnz, a.bcz); n.zs[1] = k; n.zs[2] = bc_map(k + 1, a.fnz, a.bcz);
    return n;
}

__device__ __forceinline__ int nbase(const StepArgs& a, const Nbr& nb, int xi, int yi, int zi)
{
    return nb.xs[xi] + nb.ys[yi] * a.fnx + nb.zs[zi] * a.fnx * a.fny;
}

// Base of z-major neighbour position q = xi + 3*yi + 9*zi (0..26) chosen at
// run time, mapped from the centre point (nb.xs[1], nb.ys[1], nb.zs[1]).
// Arithmetic only: a run-time pick among the Nbr fields compiles to an
// indexed load from a private copy of the struct (scratch memory).
__device__ __forceinline__ int nbase_q(const StepArgs& a, const Nbr& nb, int q)
{
    const int xi = q % 3, yi = (q / 3) % 3, zi = q / 9;
    const int x = bc_map(nb.xs[1] + xi - 1, a.fnx, a.bcx);
    const int y = bc_map(nb.ys[1] + yi - 1, a.fny, a.bcy);
    const int z = bc_map(nb.zs[1] + zi - 1, a.fnz, a.bcz);
    return x + y * a.fnx + z * a.fnx * a.fny;
}

// Rolled runs of z-major positions [from, to) that carry ids only (direct path).
template <int K>
__device__ __forceinline__ void scan_point_run(const StepArgs& a, const Nbr& nb, int from, int to,
                                               Gather<K>& g);

// Stencil row bases D,B,L,C,R,T,U (stencil = -3..3, 3DgrainMPI/APTPhaseField.cu:211-214).
__device__ __forceinline__ void stencil_bases(const StepArgs& a, const Nbr& nb, int (&b)[7])
{
    b[0] = nbase(a, nb, 1, 1, 0);
    b[1] = nbase(a, nb, 1, 0, 1);
    b[2] = nbase(a, nb, 0, 1, 1);
    b[3] = nbase(a, nb, 1, 1, 1);
    b[4] = nbase(a, nb, 2, 1, 1);
    b[5] = nbase(a, nb, 1, 2, 1);
    b[6] = nbase(a, nb, 1, 1, 2);
}

template <int K>
__device__ __forceinline__ void scan_point_run(const StepArgs& a, const Nbr& nb, int from, int to,
                                               Gather<K>& g)
{
#pragma unroll 1
    for (int q = from; q < to; ++q) scan_point<K, -1>(a.arg, nbase_q(a, nb, q), a.L, g);
}

// Pass A over the whole neighbourhood in the reference's scan order.
template <int K, bool G27>
__device__ __forceinline__ void scan_all(const StepArgs& a, const Nbr& nb, Gather<K>& g)
{
    if (!G27) {
        scan_point<K, 0>(a.arg, nbase(a, nb, 1, 1, 0), a.L, g);
        scan_point<K, 1>(a.arg, nbase(a, nb, 1, 0, 1), a.L, g);
        scan_point<K, 2>(a.arg, nbase(a, nb, 0, 1, 1), a.L, g);
        scan_point<K, 3>(a.arg, nbase(a, nb, 1, 1, 1), a.L, g);
        scan_point<K, 4>(a.arg, nbase(a, nb, 2, 1, 1), a.L, g);
        scan_point<K, 5>(a.arg, nbase(a, nb, 1, 2, 1), a.L, g);
        scan_point<K, 6>(a.arg, nbase(a, nb, 1, 1, 2), a.L, g);
    } else {
        // z-major 27-point order (3Dgrain/APTPhaseField.cu:241-268), position
        // q = xi + 3*yi + 9*zi.  The stencil rows keep phi: D=4 B=10 L=12 C=13
        // R=14 T=16 U=22; the other 20 positions only contribute ids and run as
        // rolled loops (same order) to keep the code size down.
        scan_point_run<K>(a, nb, 0, 4, g);
        scan_point<K, 0>(a.arg, nbase_q(a, nb, 4), a.L, g);
        scan_point_run<K>(a, nb, 5, 10, g);
        scan_point<K, 1>(a.arg, nbase_q(a, nb, 10), a.L, g);
        scan_point_run<K>(a, nb, 11, 12, g);
        scan_point<K, 2>(a.arg, nbase_q(a, nb, 12), a.L, g);
        scan_point<K, 3>(a.arg, nbase_q(a, nb, 13), a.L, g);
        scan_point<K, 4>(a.arg, nbase_q(a, nb, 14), a.L, g);
        scan_point_run<K>(a, nb, 15, 16, g);
        scan_point<K, 5>(a.arg, nbase_q(a, nb, 16), a.L, g);
        scan_point_run<K>(a, nb, 17, 22, g);
        scan_point<K, 6>(a.arg, nbase_q(a, nb, 22), a.L, g);
        scan_point_run<K>(a, nb, 23, 27, g);
    }
}

// Orientation rows of a slot: the grain's rotation where its centre phi
// exceeds LS (double compare), the identity elsewhere
// (3DgrainMPI/APTPhaseField.cu:258-273).
struct Rot {
    float m0, m1, m2, m3, m4, m5, m6, m7, m8;
};

__device__ __forceinline__ Rot slot_rot(const StepArgs& a, int gid, float phC)
{
    Rot r{1.0f, 0.0f, 0.0f, 0.0f, 1.0f, 0.0f, 0.0f, 0.0f, 1.0f};
    if (phC > kLSf) {
        const int tid = gid <= a.num_theta ? gid : (gid % a.num_theta) + 1;
        const float4* p = reinterpret_cast<const float4*>(a.R12 + 12 * tid);
        const float4 r0 = p[0], r1 = p[1], r2 = p[2];
        r = Rot{r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w, r2.x};
    }
    return r;
}

// Right-hand side and explicit Euler update of one filled slot
// (3DgrainMPI/APTPhaseField.cu:250-340) with its orientation rows.  repul is
// the ascending sum of 0.25*(phiC_p+1)^2 over the other filled slots (empty
// ones add exactly +0).
__device__ __forceinline__ float slot_update_r(const StepArgs& a, const Rot& m, float phD,
                                               float phB, float phL, float phC, float phR,
                                               float phT, float phU, float repul, float Up,
                                               float a_s, float epsv)
{
    const float phxn = (phR - phL) * 0.5f;
    const float phyn = (phT - phB) * 0.5f;
    const float phzn = (phU - phD) * 0.5f;
    const float A2 = kine_ani(phxn, phyn, phzn, m.m0, m.m1, m.m2, m.m3, m.m4, m.m5, m.m6, m.m7,
                              m.m8, a_s, epsv, a.eps);
    const float diff = phR + phL + phT + phB + phU + phD - 6.0f * phC;
    const float omc = 1.0f - phC * phC;
    const float rhs = diff * a.hi * a.hi + omc * phC
                    - a.lamd * Up * (omc * omc - 6.0f * (phC + 1.0f) * repul);
    const float dphi = rhs / A2;
    return phC + a.dt * dphi;
}

__device__ __forceinline__ float slot_update(const StepArgs& a, int gid, float phD, float phB,
                                             float phL, float phC, float phR, float phT,
                                             float phU, float repul, float Up, float a_s,
                                             float epsv)
{
    return slot_update_r(a, slot_rot(a, gid, phC), phD, phB, phL, phC, phR, phT, phU, repul, Up,
                         a_s, epsv);
}

// Pass B: right-hand side and explicit Euler update of every slot of one point
// (3DgrainMPI/APTPhaseField.cu:238-351).  phi_at(S, p) returns the phi of
// neighbour slot p at stencil row S; only the filled slots' values are read.
template <int K, class PhiAt>
__device__ __forceinline__ void rhs_point(const StepArgs& a, int i, int j, int k, int C,
                                          const Gather<K>& g, float a_s, float epsv, PhiAt&& phi_at)
{
    float* __restrict__ phn = a.ph_new;
    int* __restrict__ agn = a.arg_new;
    const int L = a.L;
    if (g.n == 0) {
#pragma unroll
        for (int q = 0; q < K; ++q) {
            agn[C + q * L] = -1;
            phn[C + q * L] = -1.0f;
        }
        return;
    }
    float T;
    const bool skip = thermal_T(a, i, j, k, T);
    const float Up = T / a.L_cp;
    // centre phi of every slot first: the repulsion sums over the other slots.
    // Reads are unconditional (an absent slot reads slot 0 and is replaced by
    // -1 afterwards) so each group of them goes out together instead of one
    // memory latency per read.
    float pc[K];
#pragma unroll
    for (int q = 0; q < K; ++q) {
        const uint32_t p = (g.pm[3] >> (4 * q)) & 15u;
        const float x = phi_at(3, p < (uint32_t)K ? p : 0u);
        pc[q] = (q < g.n && p != 15u) ? x : -1.0f;
    }
#pragma unroll
    for (int q = 0; q < K; ++q) {
        const int o = C + q * L;
        if (q >= g.n) {
            agn[o] = -1;
            phn[o] = -1.0f;
            continue;
        }
        if (skip) continue;
        float v[7];
#pragma unroll
        for (int s = 0; s < 7; ++s) {
            if (s == 3) {
                v[s] = pc[q];
            } else {
                const uint32_t p = (g.pm[s] >> (4 * q)) & 15u;
                const float x = phi_at(s, p < (uint32_t)K ? p : 0u);
                v[s] = (p != 15u) ? x : -1.0f;
            }
        }
        const int gid = g.la[q];
        float repul = 0.0f;
#pragma unroll
        for (int p = 0; p < K; ++p) {
            if (p != q && p < g.n) {
                const float w = pc[p] + 1.0f;
                repul += 0.25f * w * w;
            }
        }
        const float nv = slot_update(a, gid, v[0], v[1], v[2], v[3], v[4], v[5], v[6], repul, Up,
                                     a_s, epsv);
        phn[o] = nv;
        agn[o] = (nv < kKillf) ? -1 : gid;
        __builtin_amdgcn_sched_barrier(0);
    }
}

// Pass B into registers, in two phases so that the memory reads of a point
// (orientation rows, the paraboloid field) can be issued before the next
// plane's prefetch: slot_prep evaluates the thermal field, the centre phi of
// every slot and the rows; slot_finish computes the new slot values into
// out_id/out_ph.  wr is false only for slots the reference leaves untouched
// (the thermal `continue`).  The caller then issues one straight-line block
// of stores.
template <int K>
struct SlotPrep {
    float pc[K];
    float Up;
    bool skip;
};

template <int K, class PhiAt>
__device__ __forceinline__ void slot_prep(const StepArgs& a, const Gather<K>& g, float xi, float yj,
                                          float zk, PhiAt&& phi_at, SlotPrep<K>& sp)
{
    float T;
    sp.skip = thermal_Tv(a, xi, yj, zk, T);
    sp.Up = T / a.L_cp;
#pragma unroll
    for (int q = 0; q < K; ++q) {
        const uint32_t p = (g.pm[3] >> (4 * q)) & 15u;
        sp.pc[q] = (q < g.n && p != 15u) ? phi_at(3, p) : -1.0f;

    }
}

template <int K, class PhiAt>
__device__ __forceinline__ void slot_finish(const StepArgs& a, const Gather<K>& g,
                                            const SlotPrep<K>& sp, float a_s, float epsv,
                                            PhiAt&& phi_at, int (&out_id)[K], float (&out_ph)[K],
                                            bool& wr)
{
#pragma unroll
    for (int q = 0; q < K; ++q) {
        out_id[q] = -1;
        out_ph[q] = -1.0f;
    }
    wr = true;
    if (g.n == 0) return;
    wr = !sp.skip;
    if (sp.skip) return;
    // 0.25*(phiC_p+1)^2 of every filled slot, once per point
    float rw[K];
#pragma unroll
    for (int p = 0; p < K; ++p) {
        const float w = sp.pc[p] + 1.0f;
        rw[p] = 0.25f * w * w;
    }
#pragma unroll
    for (int q = 0; q < K; ++q) {
        if (q >= g.n) continue;
        float v[7];
#pragma unroll
        for (int s = 0; s < 7; ++s) {
            if (s == 3) {
                v[s] = sp.pc[q];
            } else {
                const uint32_t p = (g.pm[s] >> (4 * q)) & 15u;
                v[s] = (p != 15u) ? phi_at(s, p) : -1.0f;
            }
        }
        const int gid = g.la[q];
        float repul = 0.0f;
#pragma unroll
        for (int p = 0; p < K; ++p)
            if (p != q && p < g.n) repul += rw[p];
        const float nv = slot_update_r(a, slot_rot(a, gid, sp.pc[q]), v[0], v[1], v[2], v[3], v[4], v[5], v[6], repul,
                                       sp.Up, a_s, epsv);
        out_ph[q] = nv;
        out_id[q] = (nv < kKillf) ? -1 : gid;
    }
}

// Direct-path pass B: phi straight from global memory.
template <int K>
__device__ __forceinline__ void rhs_point(const StepArgs& a, int i, int j, int k,
                                          const int (&b)[7], const Gather<K>& g, float a_s,
                                          float epsv)
{
    rhs_point<K>(a, i, j, k, b[3], g, a_s, epsv,
                 [&](int s, uint32_t p) { return a.ph[b[s] + (int)p * a.L]; });
}

// Overflow policy for a point whose neighbourhood holds more than K distinct
// ids (undefined behaviour in the reference, 3DgrainMPI/APTPhaseField.cu:222-231).
// Kept set: the K ids ranked first by (largest phi over their entries, then
// earliest entry reaching it).  Computed in one streaming pass over the
// entries with a K-candidate list: a newcomer replaces the worst candidate
// only if its phi is strictly larger.  The worst candidate's key never
// decreases, so no id left outside can outrank it and the list ends as the
// exact top K.  Kept ids then take slots in first-appearance order (pass A
// again, filtered).
template <int K>
struct TopK {
    int id[K];
    float mx[K];
    int pos[K];
    int n;
};

template <int K>
__device__ __forceinline__ void topk_push(TopK<K>& c, int id, float v, int pos)
{
    // kept id: raise its maximum (empty entries hold id -1, which never
    // reaches here)
    bool hit = false;
#pragma unroll
    for (int q = 0; q < K; ++q) {
        const bool h = c.id[q] == id;
        const bool u = h && v > c.mx[q];
        c.mx[q] = u ? v : c.mx[q];
        c.pos[q] = u ? pos : c.pos[q];
        hit |= h;
    }
    if (hit) return;
    // new id: append while there is room, else replace the worst candidate
    // (smallest phi, latest position among equals) if v beats it
    int w = 0;
    float wm = c.mx[0];
    int wp = c.pos[0];
#pragma unroll
    for (int q = 1; q < K; ++q) {
        const bool worse = c.mx[q] < wm || (c.mx[q] == wm && c.pos[q] > wp);
        w = worse ? q : w;
        wm = worse ? c.mx[q] : wm;
        wp = worse ? c.pos[q] : wp;
    }
    const bool full = c.n >= K;
    const int tgt = full ? w : c.n;
    const bool upd = !full || v > wm;
#pragma unroll
    for (int q = 0; q < K; ++q) {
        const bool t = upd && q == tgt;
        c.id[q] = t ? id : c.id[q];
        c.mx[q] = t ? v : c.mx[q];
        c.pos[q] = t ? pos : c.pos[q];
    }
    c.n += full ? 0 : 1;
}

// Candidates of one neighbour point, entry p at arg/ph[base + p * L]; P is
// the point's position in the scan order.
template <int K>
__device__ __forceinline__ void topk_point(const int* __restrict__ arg, const float* __restrict__ ph,
                                           int base, int L, int P, TopK<K>& c)
{
    int ids[K];  // all 2K loads first, as in scan_point
    float phs[K];
#pragma unroll
    for (int p = 0; p < K; ++p) {
        ids[p] = arg[base + p * L];
        phs[p] = ph[base + p * L];
    }
#pragma unroll
    for (int p = 0; p < K; ++p) {
        if (ids[p] == -1) continue;
        topk_push(c, ids[p], phs[p], P * K + p);
    }
    __builtin_amdgcn_sched_barrier(0);
}

// The K (id, phi) entries of one neighbour point, loaded one neighbour ahead
// of their top-K pushes so that the fix-up's chain of neighbour loads
// overlaps the pushes instead of taking one memory round trip per neighbour.
template <int K>
struct NbrEntries {
    int id[K];
    float ph[K];
};

template <int K>
__device__ __forceinline__ void entries_load(const int* __restrict__ arg, const float* __restrict__ ph,
                                             int base, int L, NbrEntries<K>& e)
{
#pragma unroll
    for (int p = 0; p < K; ++p) {
        e.id[p] = arg[base + p * L];
        e.ph[p] = ph[base + p * L];
    }
}

// topk_point's pushes for entries loaded earlier; P: the point's position in
// the scan order.
template <int K>
__device__ __forceinline__ void entries_push(const NbrEntries<K>& e, int P, TopK<K>& c)
{
#pragma unroll
    for (int p = 0; p < K; ++p) {
        if (e.id[p] == -1) continue;
        topk_push(c, e.id[p], e.ph[p], P * K + p);
    }
}

// Pass A restricted to the kept ids, on one neighbour's K ids loaded earlier.
template <int K, int S>
__device__ __forceinline__ void ids_kept(const int (&ids)[K], Gather<K>& g, const TopK<K>& kept)
{
#pragma unroll
    for (int p = 0; p < K; ++p) {
        const int id = ids[p];
        if (id == -1) continue;
        bool ok = false;
#pragma unroll
        for (int u = 0; u < K; ++u) ok |= (kept.id[u] == id);
        if (!ok) continue;
        int pos = g.n;
#pragma unroll
        for (int q = 0; q < K; ++q) pos = (g.la[q] == id) ? q : pos;
#pragma unroll
        for (int q = 0; q < K; ++q) g.la[q] = (q == pos) ? id : g.la[q];
        g.n += (pos == g.n) ? 1 : 0;
        if (S >= 0) {
            const uint32_t sh = 4u * (uint32_t)pos;
            g.pm[S] = (g.pm[S] & ~(15u << sh)) | ((uint32_t)p << sh);
        }
    }
}

template <int K>
__device__ __forceinline__ void ids_load(const int* __restrict__ arg, int base, int L, int (&ids)[K])
{
#pragma unroll
    for (int p = 0; p < K; ++p) ids[p] = arg[base + p * L];
}

// Pass A restricted to the kept ids.
template <int K, int S>
__device__ __forceinline__ void scan_point_kept(const int* __restrict__ arg, int base, int L,
                                                Gather<K>& g, const TopK<K>& kept)
{
    int ids[K];  // the K loads first, as in scan_point
    ids_load<K>(arg, base, L, ids);
    ids_kept<K, S>(ids, g, kept);
    __builtin_amdgcn_sched_barrier(0);
}

// Pass A restricted to the kept ids over the 7 stencil points D,B,L,C,R,T,U
// (point u at arg[bases[u] + p * L]), each point's ids loaded while the
// previous point's are scanned.
template <int K>
__device__ __forceinline__ void kept_pass7(const int* __restrict__ arg, const int (&bases)[7], int L,
                                           Gather<K>& g, const TopK<K>& c)
{
    int x[K], y[K];
    ids_load<K>(arg, bases[0], L, x);
    ids_load<K>(arg, bases[1], L, y);
    ids_kept<K, 0>(x, g, c);
    __builtin_amdgcn_sched_barrier(0);
    ids_load<K>(arg, bases[2], L, x);
    ids_kept<K, 1>(y, g, c);
    __builtin_amdgcn_sched_barrier(0);
    ids_load<K>(arg, bases[3], L, y);
    ids_kept<K, 2>(x, g, c);
    __builtin_amdgcn_sched_barrier(0);
    ids_load<K>(arg, bases[4], L, x);
    ids_kept<K, 3>(y, g, c);
    __builtin_amdgcn_sched_barrier(0);
    ids_load<K>(arg, bases[5], L, y);
    ids_kept<K, 4>(x, g, c);
    __builtin_amdgcn_sched_barrier(0);
    ids_load<K>(arg, bases[6], L, x);
    ids_kept<K, 5>(y, g, c);
    __builtin_amdgcn_sched_barrier(0);
    ids_kept<K, 6>(x, g, c);
    __builtin_amdgcn_sched_barrier(0);
}

// Rolled runs of z-major positions [from, to) through pass A restricted to
// the kept ids (ids only).
template <int K>
__device__ __forceinline__ void kept_run(const StepArgs& a, const Nbr& nb, int from, int to,
                                         Gather<K>& g, const TopK<K>& c)
{
#pragma unroll 1
    for (int q = from; q < to; ++q) scan_point_kept<K, -1>(a.arg, nbase_q(a, nb, q), a.L, g, c);
}

template <int K, bool G27>
__device__ __forceinline__ void gather_overflow(const StepArgs& a, const Nbr& nb, Gather<K>& g)
{
    TopK<K> c;
    c.n = 0;
#pragma unroll
    for (int q = 0; q < K; ++q) {
        c.id[q] = -1;
        c.mx[q] = 0.0f;
        c.pos[q] = 0;
    }
    gather_init(g);
    if (!G27) {
        const int nbs[7] = {nbase(a, nb, 1, 1, 0), nbase(a, nb, 1, 0, 1), nbase(a, nb, 0, 1, 1),
                            nbase(a, nb, 1, 1, 1), nbase(a, nb, 2, 1, 1), nbase(a, nb, 1, 2, 1),
                            nbase(a, nb, 1, 1, 2)};
        NbrEntries<K> e[2];
        entries_load<K>(a.arg, a.ph, nbs[0], a.L, e[0]);
#pragma unroll
        for (int u = 0; u < 7; ++u) {
            if (u + 1 < 7) entries_load<K>(a.arg, a.ph, nbs[u + 1], a.L, e[(u + 1) & 1]);
            entries_push<K>(e[u & 1], u, c);
            __builtin_amdgcn_sched_barrier(0);
        }
        kept_pass7<K>(a.arg, nbs, a.L, g, c);
    } else {
#pragma unroll 1
        for (int q = 0; q < 27; ++q) topk_point<K>(a.arg, a.ph, nbase_q(a, nb, q), a.L, q, c);
        kept_run<K>(a, nb, 0, 4, g, c);
        scan_point_kept<K, 0>(a.arg, nbase_q(a, nb, 4), a.L, g, c);
        kept_run<K>(a, nb, 5, 10, g, c);
        scan_point_kept<K, 1>(a.arg, nbase_q(a, nb, 10), a.L, g, c);
        kept_run<K>(a, nb, 11, 12, g, c);
        scan_point_kept<K, 2>(a.arg, nbase_q(a, nb, 12), a.L, g, c);
        scan_point_kept<K, 3>(a.arg, nbase_q(a, nb, 13), a.L, g, c);
        scan_point_kept<K, 4>(a.arg, nbase_q(a, nb, 14), a.L, g, c);
        kept_run<K>(a, nb, 15, 16, g, c);
        scan_point_kept<K, 5>(a.arg, nbase_q(a, nb, 16), a.L, g, c);
        kept_run<K>(a, nb, 17, 22, g, c);
        scan_point_kept<K, 6>(a.arg, nbase_q(a, nb, 22), a.L, g, c);
        kept_run<K>(a, nb, 23, 27, g, c);
    }
}

// The record buffer named by the fix-up list's header (kFixHeader): entry
// f = s*K + p of record q has its id at id[f * cap + q] and its phi at
// ph[f * cap + q]; ph follows the 7 * K * cap ids in the same allocation.
struct RecView {
    int* id;
    float* ph;
    int cap;
};

template <int K>
__device__ __forceinline__ RecView rec_view(const StepArgs& a)
{
    RecView r;
    r.cap = (int)a.fix_list[-4];
    r.id = *reinterpret_cast<int* const*>(a.fix_list - 2);
    r.ph = reinterpret_cast<float*>(r.id + (size_t)7 * K * (size_t)r.cap);
    return r;
}

// The same policy on a fix-up record: the point's 7 x K entries in scan order
// (D,B,L,C,R,T,U, then slot), entry f = s*K + p at rid/rph[f * cap + q]
// (written by the tile from its LDS ring, see rec_store).
template <int K>
__device__ __forceinline__ void gather_overflow_rec(const int* rid, const float* rph, int cap, int q,
                                                    Gather<K>& g)
{
    TopK<K> c;
    c.n = 0;
#pragma unroll
    for (int u = 0; u < K; ++u) {
        c.id[u] = -1;
        c.mx[u] = 0.0f;
        c.pos[u] = 0;
    }
    gather_init(g);
    NbrEntries<K> cur, nxt;
    entries_load<K>(rid, rph, q, cap, cur);
#pragma unroll 1
    for (int s = 0; s < 7; ++s) {
        if (s + 1 < 7) entries_load<K>(rid + (s + 1) * K * cap, rph + (s + 1) * K * cap, q, cap, nxt);
        entries_push<K>(cur, s, c);
        cur = nxt;
    }
    const int bases[7] = {q, q + K * cap, q + 2 * K * cap, q + 3 * K * cap, q + 4 * K * cap,
                          q + 5 * K * cap, q + 6 * K * cap};
    kept_pass7<K>(rid, bases, cap, g, c);
}

// One interior point: pass A, then pass B.  Returns 0 if no neighbour holds a
// grain id (every slot was written as the empty pair), 1 if slots were
// computed, 2 if the neighbourhood overflows K slots: the point is then left
// untouched here and the fix-up kernel redoes it under the overflow policy.
template <int K, bool G27>
__device__ __forceinline__ int apt_point(const StepArgs& a, int i, int j, int k, float a_s,
                                         float epsv)
{
    const Nbr nb = make_nbr(a, i, j, k);
    Gather<K> g;
    gather_init(g);
    scan_all<K, G27>(a, nb, g);
    if (g.ovf) return 2;
    int b[7];
    stencil_bases(a, nb, b);
    rhs_point<K>(a, i, j, k, b, g, a_s, epsv);
    return g.n > 0 ? 1 : 0;
}

// Overflowed fix-up list entry q (fix-up kernels only): point index, and
// whether the tile left the point's entries in the record buffer.
__device__ __forceinline__ void fixup_point(const StepArgs& a, unsigned q, int& i, int& j, int& k,
                                            bool& rec)
{
    const unsigned e = a.fix_list[q];
    rec = (e & kRecFlag) != 0u;
    const int C = (int)(e & ~kRecFlag);
    const int A = a.fnx * a.fny;
    k = C / A;
    const int r = C - k * A;
    j = r / a.fnx;
    i = r - j * a.fnx;
}

// The overflow policy and pass B for a point whose entries are in record q.
template <int K>
__device__ __forceinline__ void apt_point_overflow_rec(const StepArgs& a, int i, int j, int k, int q,
                                                       float a_s, float epsv)
{
    const RecView r = rec_view<K>(a);
    Gather<K> g;
    gather_overflow_rec<K>(r.id, r.ph, r.cap, q, g);
    const int C = i + j * a.fnx + k * a.fnx * a.fny;
    rhs_point<K>(a, i, j, k, C, g, a_s, epsv, [&](int s, uint32_t p) {
        return r.ph[(s * K + (int)p) * r.cap + q];
    });
}

// The same point under the overflow policy (fix-up kernel only).
template <int K, bool G27>
__device__ __forceinline__ void apt_point_overflow(const StepArgs& a, int i, int j, int k,
                                                   float a_s, float epsv)
{
    const Nbr nb = make_nbr(a, i, j, k);
    Gather<K> g;
    gather_overflow<K, G27>(a, nb, g);
    int b[7];
    stencil_bases(a, nb, b);
    rhs_point<K>(a, i, j, k, b, g, a_s, epsv);
}

// ---------------------------------------------------------------- tiled ---
// LDS-tiled z-march.  A workgroup owns a TX x TY tile of interior columns
// and a chunk of z planes.  The ids and phi of the tile plus a one-point halo
// (boundary maps applied while loading) live in a 3-plane LDS ring; the next
// plane is prefetched into registers while the current plane is computed, so
// HBM latency overlaps the gather arithmetic.
//
// Planes are skipped with the activity map.  If every tile plane that the
// stencil of plane k can reach (the 3 x 3 neighbouring tiles at k-1, k, k+1,
// through the boundary maps) holds only empty pairs in the source, every
// point of plane k gathers no id and its result is K empty pairs; if the
// destination tile plane is also all empty, writing them changes nothing.
// Such planes are neither loaded nor written, and a source plane known to be
// all empty is staged as empty pairs without reading memory.  The results are
// the same bits as processing everything.
constexpr int kTX = 64, kTY = 4, kHX = kTX + 2, kHY = kTY + 2, kNPT = kHX * kHY;

__device__ __forceinline__ int act_tile_x(int i) { return (i - 1) / kTX; }
__device__ __forceinline__ int act_tile_y(int j) { return (j - 1) / kTY; }

template <int K>
struct PlaneLoader {
    static constexpr int NE = K * kNPT;                   // slots per staged plane
    static constexpr int PER = (NE + kBlock - 1) / kBlock;
    int off[PER];                                         // global offset minus the plane term
    int vid[PER];
    float vph[PER];
    bool live;                                            // staged plane holds non-empty slots
};

// Staging offsets of a tile.  A partial tile (at the +x or +y end) clamps the
// positions past the halo to the halo column/row; only its inactive lanes
// read those entries.
template <int K>
__device__ __forceinline__ void loader_init(PlaneLoader<K>& ld, const StepArgs& a, int i0, int j0)
{
#pragma unroll
    for (int q = 0; q < PlaneLoader<K>::PER; ++q) {
        const int e = threadIdx.x + q * kBlock;
        const int s = e / kNPT;
        const int pt = e - s * kNPT;
        const int ly = pt / kHX;
        const int lx = pt - ly * kHX;
        ld.off[q] = (e < PlaneLoader<K>::NE)
                        ? bc_map(min(i0 - 1 + lx, a.fnx - 1), a.fnx, a.bcx)
                              + bc_map(min(j0 - 1 + ly, a.fny - 1), a.fny, a.bcy) * a.fnx + s * a.L
                        : -1;
    }
}

// Elements 0..PER-2 of every lane are always inside the plane; only the last
// one can fall past NE, and its offset is clamped so the load stays in bounds
// (the value is then simply not stored).  The loads are issued whatever the
// plane: `fetch` false (nothing will be staged) and `active` false (all empty
// pairs in the source, staged as (-1, -1.0f)) read element 0 instead, so the
// number of vector-memory operations in flight is the same on every path.
template <int K>
__device__ __forceinline__ void loader_fetch(PlaneLoader<K>& ld, const StepArgs& a, int plane,
                                             bool active, bool fetch = true)
{
    const bool real = active && fetch;
    ld.live = active;
#pragma unroll
    for (int q = 0; q < PlaneLoader<K>::PER; ++q) {
        const int o = (!real || (q == PlaneLoader<K>::PER - 1 && ld.off[q] < 0)) ? 0 : ld.off[q] + plane;
        ld.vid[q] = a.arg[o];
        ld.vph[q] = a.ph[o];
    }
}

template <int K>
__device__ __forceinline__ void loader_store(const PlaneLoader<K>& ld, int* ids, float* phs)
{
#pragma unroll
    for (int q = 0; q < PlaneLoader<K>::PER; ++q) {
        if (q < PlaneLoader<K>::PER - 1 || ld.off[q] >= 0) {
            ids[threadIdx.x + q * kBlock] = ld.live ? ld.vid[q] : -1;
            phs[threadIdx.x + q * kBlock] = ld.live ? ld.vph[q] : -1.0f;
        }
    }
}

// Stage a plane straight from global memory into an LDS ring slot (no
// registers in between), for the start of a run when every slot of the ring
// is free.  Element q * kBlock + threadIdx.x goes to the same LDS word as in
// loader_store: the LDS address of a global_load_lds is the wave-uniform base
// plus 4 * lane.  The caller waits for the loads (vmcnt) and then for the
// other waves (barrier) before reading the slot.
template <int K>
__device__ __forceinline__ void loader_direct(const PlaneLoader<K>& ld, const StepArgs& a, int plane,
                                              bool active, int* ids, float* phs)
{
    const int wbase = threadIdx.x & ~63;
    if (active) {
#pragma unroll
        for (int q = 0; q < PlaneLoader<K>::PER; ++q) {
            if (q < PlaneLoader<K>::PER - 1 || ld.off[q] >= 0) {
                const int o = ld.off[q] + plane;
                __builtin_amdgcn_global_load_lds(
                    (const __attribute__((address_space(1))) void*)(a.arg + o),
                    (__attribute__((address_space(3))) void*)(ids + q * kBlock + wbase), 4, 0, 0);
                __builtin_amdgcn_global_load_lds(
                    (const __attribute__((address_space(1))) void*)(a.ph + o),
                    (__attribute__((address_space(3))) void*)(phs + q * kBlock + wbase), 4, 0, 0);
            }
        }
    } else {
#pragma unroll
        for (int q = 0; q < PlaneLoader<K>::PER; ++q) {
            if (q < PlaneLoader<K>::PER - 1 || ld.off[q] >= 0) {
                ids[threadIdx.x + q * kBlock] = -1;
                phs[threadIdx.x + q * kBlock] = -1.0f;
            }
        }
    }
}

// Pass A from LDS for one neighbour point (same probe as scan_point), in two
// parts: the K id reads, then the probes.  Reads issued ahead of the probes
// keep the LDS latency off the probe chain (a read inside the probe loop sits
// behind the previous probe's branches and waits once per slot).
// NPT: points per slot in the LDS layout (slot stride); HX (below): row stride.
template <int K, int NPT = kNPT>
__device__ __forceinline__ void lds_ids(const int* __restrict__ plane, int pt, int (&ids)[K])
{
#pragma unroll
    for (int p = 0; p < K; ++p) ids[p] = plane[p * NPT + pt];
}

template <int K, int S>
__device__ __forceinline__ void probe_ids(const int (&ids)[K], Gather<K>& g)
{
#pragma unroll
    for (int p = 0; p < K; ++p) {
        const int id = ids[p];
        if (id == -1) continue;
        int pos = g.n;
#pragma unroll
        for (int q = 0; q < K; ++q) pos = (g.la[q] == id) ? q : pos;
        if (pos >= K) {
            g.ovf = true;
            continue;
        }
        if (pos == g.n) {  // a new id takes the next free slot
#pragma unroll
            for (int q = 0; q < K; ++q) g.la[q] = (q == pos) ? id : g.la[q];
            ++g.n;
        }
        if (S >= 0) {
            const uint32_t sh = 4u * (uint32_t)pos;
            g.pm[S] = (g.pm[S] & ~(15u << sh)) | ((uint32_t)p << sh);
        }
    }
}

template <int K, int S, int NPT = kNPT>
__device__ __forceinline__ void scan_lds(const int* __restrict__ plane, int pt, Gather<K>& g)
{
    int ids[K];
    lds_ids<K, NPT>(plane, pt, ids);
    probe_ids<K, S>(ids, g);
    __builtin_amdgcn_sched_barrier(0);
}

// Rolled runs of z-major positions [from, to) of one plane (ids only).
template <int K, int NPT = kNPT, int HX = kHX>
__device__ __forceinline__ void scan_lds_run(const int* pl, int pt, int from, int to, Gather<K>& g)
{
#pragma unroll 1
    for (int q = from; q < to; ++q)
        scan_lds<K, -1, NPT>(pl, pt + ((q / 3) % 3 - 1) * HX + (q % 3 - 1), g);
}

template <int K, bool G27, int NPT = kNPT, int HX = kHX>
__device__ __forceinline__ void scan_all_lds(const int* pD, const int* pC, const int* pU, int pt,
                                             Gather<K>& g)
{
    if (!G27) {
        // two id buffers: the next neighbour's reads are in flight while the
        // current one is probed
        int b0[K], b1[K];
        lds_ids<K, NPT>(pD, pt, b0);
        lds_ids<K, NPT>(pC, pt - HX, b1);
        __builtin_amdgcn_sched_barrier(0);
        probe_ids<K, 0>(b0, g);
        lds_ids<K, NPT>(pC, pt - 1, b0);
        __builtin_amdgcn_sched_barrier(0);
        probe_ids<K, 1>(b1, g);
        lds_ids<K, NPT>(pC, pt, b1);
        __builtin_amdgcn_sched_barrier(0);
        probe_ids<K, 2>(b0, g);
        lds_ids<K, NPT>(pC, pt + 1, b0);
        __builtin_amdgcn_sched_barrier(0);
        probe_ids<K, 3>(b1, g);
        lds_ids<K, NPT>(pC, pt + HX, b1);
        __builtin_amdgcn_sched_barrier(0);
        probe_ids<K, 4>(b0, g);
        lds_ids<K, NPT>(pU, pt, b0);
        __builtin_amdgcn_sched_barrier(0);
        probe_ids<K, 5>(b1, g);
        probe_ids<K, 6>(b0, g);
        __builtin_amdgcn_sched_barrier(0);
    } else {
        // z-major 27-point order as in scan_all; non-stencil positions rolled,
        // each run within one plane (a run-time pick among the plane pointers
        // compiles to an indexed load from a private copy of them)
        scan_lds_run<K, NPT, HX>(pD, pt, 0, 4, g);
        scan_lds<K, 0, NPT>(pD, pt, g);
        scan_lds_run<K, NPT, HX>(pD, pt, 5, 9, g);
        scan_lds_run<K, NPT, HX>(pC, pt, 9, 10, g);
        scan_lds<K, 1, NPT>(pC, pt - HX, g);
        scan_lds_run<K, NPT, HX>(pC, pt, 11, 12, g);
        scan_lds<K, 2, NPT>(pC, pt - 1, g);
        scan_lds<K, 3, NPT>(pC, pt, g);
        scan_lds<K, 4, NPT>(pC, pt + 1, g);
        scan_lds_run<K, NPT, HX>(pC, pt, 15, 16, g);
        scan_lds<K, 5, NPT>(pC, pt + HX, g);
        scan_lds_run<K, NPT, HX>(pC, pt, 17, 18, g);
        scan_lds_run<K, NPT, HX>(pU, pt, 18, 22, g);
        scan_lds<K, 6, NPT>(pU, pt, g);
        scan_lds_run<K, NPT, HX>(pU, pt, 23, 27, g);
    }
}

// Overflow policy (above) for a 7-point point inside the tile, from the LDS
// ring, given the normal gather g of the point (g.n == K, g.ovf set): its slot
// list holds the first K distinct ids in first-appearance order.  If every
// entry outside that list has a smaller phi than the smallest of the list
// ids' largest phi, every other id ranks below all K list ids, so the list is
// the kept set, in the policy's (first-appearance) order, and g stands as it
// is.  A list id's largest phi is taken over the entries its row nibbles name;
// an entry a nibble does not name (another id, or a second entry of a list id
// in one row) counts as outside, which can only make the test fail, never pass
// wrongly.  Returns false when the test fails: the fix-up kernel then applies
// the policy in full.  (At the 256x256x512 start-up, 83-96 % of the overflowed
// points pass.)
template <int K, int NPT = kNPT, int HX = kHX>
__device__ __forceinline__ bool overflow_keeps_list(const int* iD, const int* iC, const int* iU,
                                                    const float* fD, const float* fC,
                                                    const float* fU, int pt, const Gather<K>& g)
{
    auto off = [&](int s) {
        return s == 1 ? pt - HX : s == 2 ? pt - 1 : s == 4 ? pt + 1 : s == 5 ? pt + HX : pt;
    };
    float best[K];
#pragma unroll
    for (int q = 0; q < K; ++q) best[q] = -__builtin_huge_valf();
    float outside = -__builtin_huge_valf();
#pragma unroll
    for (int s = 0; s < 7; ++s) {
        const int* ib = s == 0 ? iD : (s == 6 ? iU : iC);
        const float* fb = s == 0 ? fD : (s == 6 ? fU : fC);
        const int o = off(s);
        uint32_t named = 0u;  // neighbour slots a nibble of this row names
#pragma unroll
        for (int q = 0; q < K; ++q) {
            const uint32_t p = (g.pm[s] >> (4 * q)) & 15u;
            if (p == 15u) continue;
            named |= 1u << p;
            const float v = fb[(int)p * NPT + o];
            best[q] = v > best[q] ? v : best[q];
        }
#pragma unroll
        for (int p = 0; p < K; ++p) {
            if ((named >> p) & 1u) continue;
            if (ib[p * NPT + o] == -1) continue;
            const float v = fb[p * NPT + o];
            outside = v > outside ? v : outside;
        }
    }
    float low = __builtin_huge_valf();
#pragma unroll
    for (int q = 0; q < K; ++q) low = best[q] < low ? best[q] : low;
    return outside < low;
}

constexpr unsigned kNoRec = 0xFFFFFFFFu;

// Append the overflowed lanes of a wave to the fix-up list (one atomic per wave).
// With `rec`, an entry whose position has a record slot (below the record
// capacity in the list's header) is flagged and its position returned, for
// the caller to write the record there; kNoRec otherwise.
__device__ __forceinline__ unsigned fixup_append(const StepArgs& a, int ovf, unsigned point,
                                                 bool rec = false)
{
    const unsigned long long m = __ballot(ovf);
    if (m == 0ull) return kNoRec;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((long long)m) - 1;
    unsigned base = 0;
    if (lane == leader) base = atomicAdd(a.fix_count, (unsigned)__popcll(m));
    base = __shfl(base, leader);
    const unsigned pos = base + (unsigned)__popcll(m & ((1ull << lane) - 1ull));
    const unsigned cap = rec ? a.fix_list[-4] : 0u;
    const bool has = ovf && pos < cap;
    if (ovf) a.fix_list[pos] = point | (has ? kRecFlag : 0u);
    return has ? pos : kNoRec;
}

// Record of an overflowed point for the fix-up: its 7 x K (id, phi) entries
// from the LDS ring, in scan order, at record pos (structure of arrays, so the
// wave's stores and the fix-up's loads are coalesced over consecutive entries).
template <int K, int NPT = kNPT, int HX = kHX>
__device__ __forceinline__ void rec_store(const StepArgs& a, const int* iD, const int* iC,
                                          const int* iU, const float* fD, const float* fC,
                                          const float* fU, int pt, unsigned pos)
{
    const RecView r = rec_view<K>(a);
    // one row at a time: only 2K values in flight, not 14K registers' worth
#pragma unroll 1
    for (int s = 0; s < 7; ++s) {
        const int o = s == 1 ? pt - HX : s == 2 ? pt - 1 : s == 4 ? pt + 1 : s == 5 ? pt + HX : pt;
        const int* ib = s == 0 ? iD : (s == 6 ? iU : iC);
        const float* fb = s == 0 ? fD : (s == 6 ? fU : fC);
#pragma unroll
        for (int p = 0; p < K; ++p) {
            r.id[(s * K + p) * r.cap + (int)pos] = ib[p * NPT + o];
            r.ph[(s * K + p) * r.cap + (int)pos] = fb[p * NPT + o];
        }
    }
}

// One z plane of a tile.  On entry the ring holds planes k-1, k, k+1.
// Order of the vector-memory work: gather from LDS, then the point's own
// reads (orientation rows, paraboloid field), then the prefetch of plane k+2
// (when plane k+1 is processed next), then the slot arithmetic, then one
// straight-line block of 2K stores, then plane k+2 is staged into the ring
// slot plane k-1 leaves free.  Vector-memory operations retire in issue
// order, so the arithmetic only waits for the point's reads, never for the
// prefetch, and staging only for the prefetch, never for the stores; every
// path into an iteration has the same operations in flight.
// REC: hand overflowed points to the fix-up with a record of their entries
// (the step kernel variant used while neighbourhoods overflow).
template <int K, bool G27, bool REC = true>
__device__ __forceinline__ void tile_plane(const StepArgs& a, PlaneLoader<K>& ld, int (*rid)[K * kNPT],
                                           float (*rph)[K * kNPT], const float* zs, int i, int j,
                                           int k, int k0, int tile, int pt, bool lane_on, float xi,
                                           float yj, float a_s, float epsv, bool stage_next,
                                           bool act_next2, unsigned& nres)
{
    const int A = a.fnx * a.fny;
    __syncthreads();
    const int sD = (k + 2) % 3, sC = k % 3, sU = (k + 1) % 3;
    Gather<K> g;
    gather_init(g);
    scan_all_lds<K, G27>(rid[sD], rid[sC], rid[sU], pt, g);
    if (!lane_on) {  // lane past the end of a partial tile: nothing to compute or store
        g.n = 0;
        g.ovf = false;
    }
    // 7-point overflow settled here from the ring when the gather's list is
    // the kept set (the wave skips this when no lane overflowed); the rest go
    // to the fix-up kernel.  nres counts the wave's settled points for the
    // overflow diagnostic.
    if (!G27 && REC) {
        const bool was = g.ovf;
        if (__ballot(was) != 0ull) {
            if (was && overflow_keeps_list<K>(rid[sD], rid[sC], rid[sU], rph[sD], rph[sC], rph[sU], pt, g))
                g.ovf = false;
            nres += (unsigned)__popcll(__ballot(was && !g.ovf));
        }
    }
    const int ovf = g.ovf ? 1 : 0;
    const int C = i + j * a.fnx + k * A;
    const float* pD = rph[sD];
    const float* pC = rph[sC];
    const float* pU = rph[sU];
    auto phi_at = [&](int s, uint32_t p) {
        const int o = (int)p * kNPT;
        switch (s) {
        case 0: return pD[o + pt];
        case 1: return pC[o + pt - kHX];
        case 2: return pC[o + pt - 1];
        case 3: return pC[o + pt];
        case 4: return pC[o + pt + 1];
        case 5: return pC[o + pt + kHX];
        default: return pU[o + pt];
        }
    };
    SlotPrep<K> sp;
    if (ovf) g.n = 0;  // an overflowed point computes nothing here (fix-up kernel)
    slot_prep<K>(a, g, xi, yj, zs[k - k0], phi_at, sp);
    loader_fetch(ld, a, bc_map(k + 2, a.fnz, a.bcz) * A, act_next2, stage_next);
    int oid[K];
    float oph[K];
    bool wr;
    slot_finish<K>(a, g, sp, a_s, epsv, phi_at, oid, oph, wr);
    // Overflowed points are written by the fix-up kernel; skipped slots keep the
    // destination's old value.  Both send their stores to the per-lane sink.
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int q = 0; q < K; ++q) {
        const bool w = lane_on && !ovf && (wr || q >= g.n);
        int* pa = w ? a.arg_new + C + q * a.L : a.sink_i + lane;
        float* pp = w ? a.ph_new + C + q * a.L : a.sink_f + lane;
        *pa = oid[q];
        *pp = oph[q];
    }
    // A point that gathered an id may leave non-empty slots (an overflowed one
    // will, through the fix-up): mark the destination tile plane.
    if (ovf || g.n > 0) a.act_dst[k * a.ntiles + tile] = a.ep_new;
    // Overflowed lanes are redone after the step by the per-point fix-up, from
    // a record of their entries where one fits (7-point gather).
    const unsigned pos = fixup_append(a, ovf, (unsigned)C, !G27 && REC);
    if (!G27 && REC && pos != kNoRec)
        rec_store<K>(a, rid[sD], rid[sC], rid[sU], rph[sD], rph[sC], rph[sU], pt, pos);
    if (stage_next) {
        __syncthreads();
        loader_store(ld, rid[sD], rph[sD]);
    }
}

// The LDS ring of a step workgroup: 3 planes of a 66x6 tile region (tile plus
// halo), every slot's id and phi.  One allocation per K, shared by the tile
// march and the edge strip (edge_tile), which never run at the same time in a
// workgroup.
template <int K>
struct LdsRing {
    int id[3][K * kNPT];
    float ph[3][K * kNPT];
};

template <int K>
__device__ __forceinline__ LdsRing<K>& lds_ring()
{
    __shared__ LdsRing<K> r;
    return r;
}

// The plan kernels' unit of work: wave gw of a simulation covers z-chunk
// group cg = gw / (waves per group) and plan_tiles consecutive full tiles.
// Lane l reads the activity of tile slot l / plan_rows at plane kb - 1 +
// l % plan_rows (kb: first plane of the group): the 3 x 3 source words and
// the tile's own destination word.  Two ballots then hold every chunk's bits,
// and lane l = c * plan_tiles + slot assembles the masks of chunk c of the
// group for that tile slot, chunk = [k0, k1): need bit m for plane k0+m to
// process (some source word of the 3 x 3 tiles around the tile is active at a
// plane within one of it, through the z map, or its own destination word is),
// act bit r for the source activity of plane k0-1+r.  Items leave a wave
// chunk-major, so the list runs through z-chunk groups in order and the items
// in flight together lie in few planes.  put: the chunk has planes to
// process; item = its work-list code.
__device__ __forceinline__ void plan_wave(const StepArgs& a, int gw, int lane, bool& put, int& item,
                                          uint64_t& need, uint64_t& act)
{
    const int R = plan_rows(a.kchunk);
    const int TW = plan_tiles(a.kchunk);
    const int C = plan_group(a.kchunk);
    const int wpg = (a.items_chunk + TW - 1) / TW;  // waves per chunk group
    const int cg = gw / wpg;
    const int t0 = (gw - cg * wpg) * TW;
    const int zc0 = cg * C;
    const int kb = 1 + zc0 * a.kchunk;
    const int klast = min(kb + C * a.kchunk, a.fnz - 1);  // upper halo plane of the last chunk
    bool on = false, dst = false;
    {
        const int t = t0 + lane / R;
        const int kk = kb - 1 + lane % R;
        if (a.use_act && t < a.items_chunk && kk <= klast) {
            const int ty = t / a.tiles_x;
            const int tx = t - ty * a.tiles_x;
            const int i0 = 1 + tx * kTX, j0 = 1 + ty * kTY;
            const int xl = act_tile_x(bc_map(i0 - 1, a.fnx, a.bcx));
            const int xr = act_tile_x(bc_map(min(i0 + kTX, a.fnx - 1), a.fnx, a.bcx));
            const int yl = act_tile_y(bc_map(j0 - 1, a.fny, a.bcy));
            const int yr = act_tile_y(bc_map(min(j0 + kTY, a.fny - 1), a.fny, a.bcy));
            const int rl = yl * a.ntx, rc = ty * a.ntx, rr = yr * a.ntx;
            const unsigned* w = a.act_src + bc_map(kk, a.fnz, a.bcz) * a.ntiles;
            const unsigned e = a.ep_src;
            // bitwise ors: the nine reads go out together
            on = (w[xl + rl] >= e) | (w[tx + rl] >= e) | (w[xr + rl] >= e) | (w[xl + rc] >= e) |
                 (w[tx + rc] >= e) | (w[xr + rc] >= e) | (w[xl + rr] >= e) | (w[tx + rr] >= e) |
                 (w[xr + rr] >= e);
            dst = kk >= kb && kk < klast && a.act_dst[kk * a.ntiles + tx + rc] >= a.ep_dst;
        }
    }
    const uint64_t A = __ballot(on);
    const uint64_t D = __ballot(dst);
    const int c = lane / TW;
    const int slot = lane - c * TW;
    const int zc = zc0 + c;
    const int t = t0 + slot;
    put = false;
    item = 0;
    need = 0ull;
    act = 0ull;
    if (c < C && zc < a.nzchunk && t < a.items_chunk) {
        const int k0 = 1 + zc * a.kchunk;
        const int n = min(k0 + a.kchunk, a.fnz - 1) - k0;
        const uint64_t full = (1ull << n) - 1ull;  // n <= 61
        if (!a.use_act) {
            need = full;
            act = ~0ull;
        } else {
            const int sh = slot * R + c * a.kchunk;  // lane of plane k0 - 1 for this tile
            const uint64_t ac = (A >> sh) & ((1ull << (n + 2)) - 1ull);
            act = ac;
            need = ((ac | (ac >> 1) | (ac >> 2)) | ((D >> (sh + 1)) & full)) & full;
        }
        put = need != 0ull;
        item = zc * a.items_chunk + t;
    }
}

// List position of this lane's chunk if `put`: a heavy chunk (more than half
// of its planes to process) from the front of the list, a light one from the
// back (cap - 1 downward).  Every thread of the workgroup (kPlanWaves waves)
// calls it; waves take consecutive runs of positions in wave order.
template <int kPlanWaves>
__device__ __forceinline__ unsigned plan_slot(bool put, bool heavy, unsigned* front, unsigned* back,
                                              unsigned cap, unsigned* s_cnt)
{
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const unsigned long long mh = __ballot(put && heavy);
    const unsigned long long ml = __ballot(put && !heavy);
    if (lane == 0) {
        s_cnt[w] = (unsigned)__popcll(mh);
        s_cnt[kPlanWaves + 1 + w] = (unsigned)__popcll(ml);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned th = 0u, tl = 0u;
        for (int u = 0; u < kPlanWaves; ++u) {
            const unsigned ch = s_cnt[u], cl = s_cnt[kPlanWaves + 1 + u];
            s_cnt[u] = th;
            s_cnt[kPlanWaves + 1 + u] = tl;
            th += ch;
            tl += cl;
        }
        s_cnt[kPlanWaves] = th ? atomicAdd(front, th) : 0u;
        s_cnt[2 * kPlanWaves + 1] = tl ? atomicAdd(back, tl) : 0u;
    }
    __syncthreads();
    const unsigned long long below = (1ull << lane) - 1ull;
    if (heavy) return s_cnt[kPlanWaves] + s_cnt[w] + (unsigned)__popcll(mh & below);
    return cap - 1u - (s_cnt[2 * kPlanWaves + 1] + s_cnt[kPlanWaves + 1 + w] +
                       (unsigned)__popcll(ml & below));
}

__device__ __forceinline__ bool plan_heavy(uint64_t need, int kchunk)
{
    return 2 * __popcll(need) > kchunk;
}

// One workgroup of a step's plan (apt_plan, and the plan half of
// apt_fixplan): plan block pb covers waves 4*pb .. 4*pb+3; block 0 also
// clears the counter set of the step after this one.
__device__ __forceinline__ void plan_block(const StepArgs& a, int pb)
{
    constexpr int kPlanWaves = kBlock / 64;
    __shared__ unsigned s_cnt[2 * kPlanWaves + 2];
    const int g = pb * kBlock + (int)threadIdx.x;
    // the counter set of the step after this one (last used two steps
    // before, whose fix-up has finished) starts from zero
    if (g < 4) a.cnt_next[g] = 0u;
    const int gw = g >> 6;  // wave-uniform
    const int lane = threadIdx.x & 63;
    bool put = false;
    int item = 0;
    uint64_t need = 0ull, act = 0ull;
    if (gw < plan_waves(a.kchunk, a.nzchunk, a.items_chunk))
        plan_wave(a, gw, lane, put, item, need, act);
    const unsigned q = plan_slot<kPlanWaves>(put, plan_heavy(need, a.kchunk), a.wl_count, a.wl_light,
                                             (unsigned)(a.nzchunk * a.items_chunk), s_cnt);
    if (put) {
        a.wl_item[q] = item;
        a.wl_need[q] = need;
        a.wl_act[q] = act;
    }
}

// nres (wave-uniform) accumulates the overflowed points the tile resolved.
template <int K, bool G27, bool REC = true>
__device__ __forceinline__ void tile_march(const StepArgs& a, int tx, int ty, int k0, int k1,
                                           uint64_t need, uint64_t act, float a_s, float epsv,
                                           unsigned& nres)
{
    int (*rid)[K * kNPT] = lds_ring<K>().id;
    float (*rph)[K * kNPT] = lds_ring<K>().ph;
    __shared__ float zs[64];  // z of the chunk's planes (a chunk has at most 61)
    const int i0 = 1 + tx * kTX, j0 = 1 + ty * kTY;
    const int lx = threadIdx.x & (kTX - 1);
    const int ly = threadIdx.x / kTX;
    const int i = i0 + lx, j = j0 + ly;
    const int pt = (ly + 1) * kHX + (lx + 1);
    const int A = a.fnx * a.fny;
    const int tile = tx + ty * a.ntx;
    auto act_at = [&](int kk) { return kk <= k1 && ((act >> (kk - k0 + 1)) & 1ull) != 0ull; };
    auto need_at = [&](int kk) { return kk < k1 && ((need >> (kk - k0)) & 1ull) != 0ull; };
    // coordinates for the thermal field: per lane once, z through LDS (made
    // visible by the barrier that opens every run)
    const bool lane_on = i <= a.fnx - 2 && j <= a.fny - 2;
    const float xi = a.x[min(i, a.fnx - 2)], yj = a.y[min(j, a.fny - 2)];
    if ((int)threadIdx.x < k1 - k0) zs[threadIdx.x] = a.z[k0 + threadIdx.x];
    PlaneLoader<K> ld;
    loader_init(ld, a, i0, j0);
    int k = k0;
    while (k < k1) {
        if (!need_at(k)) {
            ++k;
            continue;
        }
        // start of a run of processed planes: k-1, k, k+1 into the ring once
        // every wave is done with the ring
        __syncthreads();
        loader_direct(ld, a, bc_map(k - 1, a.fnz, a.bcz) * A, act_at(k - 1), rid[(k + 2) % 3],
                      rph[(k + 2) % 3]);
        loader_direct(ld, a, k * A, act_at(k), rid[k % 3], rph[k % 3]);
        loader_direct(ld, a, bc_map(k + 1, a.fnz, a.bcz) * A, act_at(k + 1), rid[(k + 1) % 3],
                      rph[(k + 1) % 3]);
        __builtin_amdgcn_s_waitcnt(0);  // the three planes are in LDS (barrier in tile_plane)
        do {
            tile_plane<K, G27, REC>(a, ld, rid, rph, zs, i, j, k, k0, tile, pt, lane_on, xi, yj, a_s,
                               epsv, need_at(k + 1), act_at(k + 2), nres);
            ++k;
        } while (need_at(k));
    }
}

// A narrow strip of columns left at the +x end (fewer than kPartialMin; wider
// leftovers, and every leftover in y, are partial LDS tiles) takes the direct
// path, one point per lane: point p of z chunk [k0, k1) is strip column
// p % edge_cols (row-major over the strip) at plane k0 + p / edge_cols.  A
// point is skipped under the same rule as a tile plane: all of its stencil's
// tile planes empty in the source and its own tile plane empty in the
// destination.
// Whether interior point (i, j, k) of the +x strip must be processed: some
// tile plane of its 3x3x3 tile neighbourhood is active in the source, or its
// own tile plane is in the destination (the tile plane rule of the activity
// map, per point).
__device__ __forceinline__ bool edge_need(const StepArgs& a, int i, int j, int k, int tile)
{
    if (!a.use_act) return true;
    const int xl = act_tile_x(bc_map(i - 1, a.fnx, a.bcx)), xc = act_tile_x(i);
    const int xr = act_tile_x(bc_map(i + 1, a.fnx, a.bcx));
    const int rl = act_tile_y(bc_map(j - 1, a.fny, a.bcy)) * a.ntx, rc = act_tile_y(j) * a.ntx;
    const int rr = act_tile_y(bc_map(j + 1, a.fny, a.bcy)) * a.ntx;
    const unsigned e = a.ep_src;
    // bitwise ors: all 28 reads go out together (a short-circuit chain
    // would issue them one memory latency apart)
    bool any = a.act_dst[k * a.ntiles + tile] >= a.ep_dst;
#pragma unroll
    for (int dz = -1; dz <= 1; ++dz) {
        const unsigned* w = a.act_src + bc_map(k + dz, a.fnz, a.bcz) * a.ntiles;
        any = any | (w[xl + rl] >= e) | (w[xc + rl] >= e) | (w[xr + rl] >= e) |
              (w[xl + rc] >= e) | (w[xc + rc] >= e) | (w[xr + rc] >= e) |
              (w[xl + rr] >= e) | (w[xc + rr] >= e) | (w[xr + rr] >= e);
    }
    return any;
}

template <int K, bool G27>
__device__ __forceinline__ void edge_point(const StepArgs& a, int p, int k0, int k1, float a_s,
                                           float epsv)
{
    const int r = p / a.edge_cols;
    const int c = p - r * a.edge_cols;
    const int k = k0 + r;
    const int ex = (a.fnx - 2) - a.tiles_x * kTX;  // strip width, > 0 whenever edge items exist
    const int j = 1 + c / ex;
    const int i = 1 + a.tiles_x * kTX + (c - (j - 1) * ex);
    const bool live = k < k1;
    const int tile = act_tile_x(i) + act_tile_y(j) * a.ntx;
    const bool need = live && edge_need(a, i, j, k, tile);
    int res = 0;
    if (need) {
        res = apt_point<K, G27>(a, i, j, k, a_s, epsv);
        if (res != 0) a.act_dst[k * a.ntiles + tile] = a.ep_new;
    }
    fixup_append(a, res == 2, (unsigned)(i + j * a.fnx + k * a.fnx * a.fny));
}

// The strip is one column wide (the reference sizes its grids as 64 m + 1
// points, e.g. 257 = 4 x 64 + 1): it runs as edge tiles (7-point gather).
__device__ __forceinline__ bool edge_strip_is_column(const StepArgs& a)
{
    return a.edge_cols == a.fny - 2;
}

// One edge tile: rows 1 + 64 seg .. 64 seg + 64 of the one-column strip
// i = fnx - 2 over planes [k0, k1).  Lanes are (row, plane), 4 planes per
// pass.  A pass stages the columns i-1, i, i+1 of its 66 rows (halo included)
// and 6 planes, every slot's id and phi, straight from global memory into
// the step's LDS ring with global_load_lds (K x 1188 points, exactly the
// ring's size), then each lane runs the tile lane's gather, keep test,
// records and slot arithmetic with row stride 3 and plane stride 198.  A
// lane processes its point under edge_point's activity rule (the others
// neither store nor mark nor hand over), and a pass none of whose lanes
// does stages nothing.  Against the direct path (neighbours through L1/L2,
// one dependent round trip per neighbour) the staging is one round trip per
// pass, and overflowed points get the keep test and records (the direct path
// hands every one to the fix-up without a record).  Used by the batched step
// and by the single step's edge kernel of overflow steps (apt_edge).  Returns
// the overflowed points the keep test settled.  REC false: neither keep test
// nor records (every overflowed point goes to the fix-up whole).
template <int K, bool REC = true>
__device__ __forceinline__ unsigned edge_tile(const StepArgs& a, int seg, int k0, int k1,
                                              float a_s, float epsv)
{
    unsigned nres = 0u;
    constexpr int EW = 3, EH = kEdgeRows + 2, EPL = EW * EH, ENPT = EPL * 6;
    static_assert(ENPT == 3 * kNPT, "an edge tile fills the step's LDS ring exactly");
    constexpr int NE = K * ENPT;  // staged entries per pass
    constexpr int PER = (NE + kBlock - 1) / kBlock;
    int* eid = &lds_ring<K>().id[0][0];
    float* eph = &lds_ring<K>().ph[0][0];
    const int A = a.fnx * a.fny;
    const int i = a.fnx - 2;  // the strip column
    const int lane = threadIdx.x & 63;
    const int ly = (int)threadIdx.x % kEdgeRows, lz = (int)threadIdx.x / kEdgeRows;
    const int j = 1 + seg * kEdgeRows + ly;
    const bool row_on = j <= a.fny - 2;
    const int jc = min(j, a.fny - 2);
    const float xi = a.x[i], yj = a.y[jc];
    const int tile = act_tile_x(i) + act_tile_y(jc) * a.ntx;
    const int pt = lz * EPL + (ly + 1) * EW + 1;  // in plane lz + 1 through iC / fC
    const int* iD = eid;
    const int* iC = eid + EPL;
    const int* iU = eid + 2 * EPL;
    const float* fD = eph;
    const float* fC = eph + EPL;
    const float* fU = eph + 2 * EPL;
    const int wbase = threadIdx.x & ~63;
#pragma unroll 1
    for (int kb = k0; kb < k1; kb += kBlock / kEdgeRows) {
        const int k = kb + lz;
        const int kc = min(k, a.fnz - 2);
        const bool need = row_on && k < k1 && edge_need(a, i, jc, kc, tile);
        // the barrier also orders the previous pass's LDS reads before this
        // pass's staging
        if (!__syncthreads_or(need ? 1 : 0)) continue;
#pragma unroll 1
        for (int q = 0; q < PER; ++q) {
            const int e = q * kBlock + (int)threadIdx.x;
            if (e < NE) {
                const int p = e / ENPT;
                const int r = e - p * ENPT;
                const int z = r / EPL;
                const int r2 = r - z * EPL;
                const int y = r2 / EW;
                const int x = r2 - y * EW;
                const int gx = bc_map(i - 1 + x, a.fnx, a.bcx);
                const int gy = bc_map(min(seg * kEdgeRows + y, a.fny - 1), a.fny, a.bcy);
                const int gz = bc_map(min(kb - 1 + z, a.fnz - 1), a.fnz, a.bcz);
                const int o = gx + gy * a.fnx + gz * A + p * a.L;
                __builtin_amdgcn_global_load_lds(
                    (const __attribute__((address_space(1))) void*)(a.arg + o),
                    (__attribute__((address_space(3))) void*)(eid + q * kBlock + wbase), 4, 0, 0);
                __builtin_amdgcn_global_load_lds(
                    (const __attribute__((address_space(1))) void*)(a.ph + o),
                    (__attribute__((address_space(3))) void*)(eph + q * kBlock + wbase), 4, 0, 0);
            }
        }
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        Gather<K> g;
        gather_init(g);
        scan_all_lds<K, false, ENPT, EW>(iD, iC, iU, pt, g);
        if (!need) {
            g.n = 0;
            g.ovf = false;
        }
        if (REC) {  // the keep test of tile_plane
            const bool was = g.ovf;
            if (__ballot(was) != 0ull) {
                if (was && overflow_keeps_list<K, ENPT, EW>(iD, iC, iU, fD, fC, fU, pt, g))
                    g.ovf = false;
                nres += (unsigned)__popcll(__ballot(was && !g.ovf));
            }
        }
        const int ovf = g.ovf ? 1 : 0;
        const int C = i + jc * a.fnx + kc * A;
        auto phi_at = [&](int s, uint32_t p) {
            const int o = (int)p * ENPT;
            switch (s) {
            case 0: return fD[o + pt];
            case 1: return fC[o + pt - EW];
            case 2: return fC[o + pt - 1];
            case 3: return fC[o + pt];
            case 4: return fC[o + pt + 1];
            case 5: return fC[o + pt + EW];
            default: return fU[o + pt];
            }
        };
        SlotPrep<K> sp;
        if (ovf) g.n = 0;
        slot_prep<K>(a, g, xi, yj, a.z[kc], phi_at, sp);
        int oid[K];
        float oph[K];
        bool wr;
        slot_finish<K>(a, g, sp, a_s, epsv, phi_at, oid, oph, wr);
#pragma unroll
        for (int q = 0; q < K; ++q) {
            const bool w = need && !ovf && (wr || q >= g.n);
            int* pa = w ? a.arg_new + C + q * a.L : a.sink_i + lane;
            float* pp = w ? a.ph_new + C + q * a.L : a.sink_f + lane;
            *pa = oid[q];
            *pp = oph[q];
        }
        if (need && (ovf || g.n > 0)) a.act_dst[kc * a.ntiles + tile] = a.ep_new;
        const unsigned pos = fixup_append(a, ovf, (unsigned)C, REC);
        if (REC && pos != kNoRec) rec_store<K, ENPT, EW>(a, iD, iC, iU, fD, fC, fU, pt, pos);
    }
    return nres;
}

__host__ __device__ inline int cdiv(long long a, long long b) { return (int)((a + b - 1) / b); }

}  // namespace

}  // namespace cupf
