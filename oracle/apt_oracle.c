/*
 * apt_oracle.c -- CPU restatement of the cuPF 3-D APT step (see apt_oracle.h).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path, never the product.
 *
 * Every function below restates one reference kernel as plain scalar loops,
 * in the same arithmetic order, with the same implicit float->double
 * promotions (build the REAL=float variant with -ffp-contract=off).
 */
#include "apt_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef ORC_F64
typedef double REAL;
#define RSQRT(v) sqrt((double)(v))
#define REXP(v) exp((double)(v))
#else
typedef float REAL;
#define RSQRT(v) sqrtf((float)(v))
#define REXP(v) expf((float)(v))
#endif

/* 3DgrainMPI/APTPhaseField.cu:25 -- a double literal, so every `> LS` test
 * promotes its float operand to double. */
#define LS (-0.995)
/* 3DgrainMPI/APTPhaseField.cu:26 */
#define OMEGA 12

static inline long idx4(const orc_params* P, int i, int j, int k, int s)
{
    long L = (long)P->fnx * P->fny * P->fnz;
    return i + (long)j * P->fnx + (long)k * P->fnx * P->fny + (long)s * L;
}

/* ------------------------------------------------------------------ BC --- */
/* APTsetBC3D (3DgrainMPI/APTPhaseField.cu:93-174) with processor (0,0,0) of
 * (1,1,1).  bc=0: out[0]=in[2], out[n-1]=in[n-3]; bc=1: out[0]=in[n-2],
 * out[n-1]=in[1].  Faces are filled z, then y, then x, each pass over the
 * full face including halo rows, which composes the per-axis maps at edges
 * and corners (the reference races there). */
void orc_set_bc(const orc_params* P, void* vph, int* arg)
{
    REAL* ph = (REAL*)vph;
    int fnx = P->fnx, fny = P->fny, fnz = P->fnz;
    for (int s = 0; s < P->K; s++) {
        int zd = P->bcz == 0 ? 2 : fnz - 2, zu = P->bcz == 0 ? fnz - 3 : 1;
        for (int j = 0; j < fny; j++)
            for (int i = 0; i < fnx; i++) {
                long o = idx4(P, i, j, 0, s), n = idx4(P, i, j, zd, s);
                ph[o] = ph[n]; arg[o] = arg[n];
                o = idx4(P, i, j, fnz - 1, s); n = idx4(P, i, j, zu, s);
                ph[o] = ph[n]; arg[o] = arg[n];
            }
        int yb = P->bcy == 0 ? 2 : fny - 2, yt = P->bcy == 0 ? fny - 3 : 1;
        for (int k = 0; k < fnz; k++)
            for (int i = 0; i < fnx; i++) {
                long o = idx4(P, i, 0, k, s), n = idx4(P, i, yb, k, s);
                ph[o] = ph[n]; arg[o] = arg[n];
                o = idx4(P, i, fny - 1, k, s); n = idx4(P, i, yt, k, s);
                ph[o] = ph[n]; arg[o] = arg[n];
            }
        int xl = P->bcx == 0 ? 2 : fnx - 2, xr = P->bcx == 0 ? fnx - 3 : 1;
        for (int k = 0; k < fnz; k++)
            for (int j = 0; j < fny; j++) {
                long o = idx4(P, 0, j, k, s), n = idx4(P, xl, j, k, s);
                ph[o] = ph[n]; arg[o] = arg[n];
                o = idx4(P, fnx - 1, j, k, s); n = idx4(P, xr, j, k, s);
                ph[o] = ph[n]; arg[o] = arg[n];
            }
    }
}

/* ------------------------------------------------------------ kine_ani --- */
/* 3DgrainMPI/APTPhaseField.cu:36-52.  The reference rotation
 * (cosa*cosb, sina*cosb, -sinb; -sina, cosa, 0; cosa*sinb, sina*sinb, cosb)
 * is one row-major R9; each row is evaluated as (m0*ux + m1*uy) + m2*uz. */
static REAL kine_ani(const orc_params* P, REAL ux, REAL uy, REAL uz, const float* m)
{
    REAL kd = P->kin_delta;
    REAL a_s = (REAL)1.0f + (REAL)3.0f * kd;
    REAL epsilon = (REAL)-4.0f * kd / a_s;
    REAL m0 = m[0], m1 = m[1], m2 = m[2], m3 = m[3], m4 = m[4], m5 = m[5],
         m6 = m[6], m7 = m[7], m8 = m[8];
    REAL ux2 = m0 * ux + m1 * uy + m2 * uz;
    ux2 = ux2 * ux2;
    REAL uy2 = m3 * ux + m4 * uy + m5 * uz;
    uy2 = uy2 * uy2;
    REAL uz2 = m6 * ux + m7 * uy + m8 * uz;
    uz2 = uz2 * uz2;
    REAL MAG_sq = (ux2 + uy2 + uz2);
    REAL MAG_sq2 = MAG_sq * MAG_sq;
    if (MAG_sq > (REAL)P->eps)
        return a_s * ((REAL)1.0f + epsilon * (ux2 * ux2 + uy2 * uy2 + uz2 * uz2) / MAG_sq2);
    return (REAL)1.0f;
}

double orc_kine_ani(const orc_params* P, double ux, double uy, double uz, const float* R9)
{
    return (double)kine_ani(P, (REAL)ux, (REAL)uy, (REAL)uz, R9);
}

/* ------------------------------------------------------ interpolation --- */
/* interp3Dtemperature (3DgrainMPI/devicefunc.cu_inl:85-125).  Indices are
 * clamped into the macro grid; the reference reads out of bounds there. */
static REAL interp3(const float* u, REAL x, REAL y, REAL z, int Nx, int Ny, int Nz, REAL Dx)
{
    int kx = (int)(x / Dx), ky = (int)(y / Dx), kz = (int)(z / Dx);
    REAL dx = x / Dx - kx, dy = y / Dx - ky, dz = z / Dx - kz;
    if (kx == Nx - 1) { kx = Nx - 2; dx = (REAL)1.0f; }
    if (ky == Ny - 1) { ky = Ny - 2; dy = (REAL)1.0f; }
    if (kz == Nz - 1) { kz = Nz - 2; dz = (REAL)1.0f; }
    if (kx < 0) kx = 0; if (kx > Nx - 2) kx = Nx - 2;
    if (ky < 0) ky = 0; if (ky > Ny - 2) ky = Ny - 2;
    if (kz < 0) kz = 0; if (kz > Nz - 2) kz = Nz - 2;
    long o = kx + (long)ky * Nx + (long)kz * Nx * Ny;
    long on = kx + (long)ky * Nx + (long)(kz + 1) * Nx * Ny;
    REAL one = (REAL)1.0f;
    return ((one - dx) * (one - dy) * (REAL)u[o] + (one - dx) * dy * (REAL)u[o + Nx]
            + dx * (one - dy) * (REAL)u[o + 1] + dx * dy * (REAL)u[o + Nx + 1]) * (one - dz)
         + ((one - dx) * (one - dy) * (REAL)u[on] + (one - dx) * dy * (REAL)u[on + Nx]
            + dx * (one - dy) * (REAL)u[on + 1] + dx * dy * (REAL)u[on + Nx + 1]) * dz;
}

/* interp4Dtemperature (3DgrainMPI/devicefunc.cu_inl:32-83), clamped likewise. */
static REAL interp4(const float* u, REAL x, REAL y, REAL z, REAL t, int Nx, int Ny, int Nz,
                    int Nt, REAL Dx, REAL Dt)
{
    int kx = (int)(x / Dx), ky = (int)(y / Dx), kz = (int)(z / Dx), kt = (int)(t / Dt);
    REAL dx = x / Dx - kx, dy = y / Dx - ky, dz = z / Dx - kz, dt = t / Dt - kt;
    if (kx == Nx - 1) { kx = Nx - 2; dx = (REAL)1.0f; }
    if (ky == Ny - 1) { ky = Ny - 2; dy = (REAL)1.0f; }
    if (kz == Nz - 1) { kz = Nz - 2; dz = (REAL)1.0f; }
    if (kt == Nt - 1) { kt = Nt - 2; dt = (REAL)1.0f; }
    if (kx < 0) kx = 0; if (kx > Nx - 2) kx = Nx - 2;
    if (ky < 0) ky = 0; if (ky > Ny - 2) ky = Ny - 2;
    if (kz < 0) kz = 0; if (kz > Nz - 2) kz = Nz - 2;
    if (kt < 0) kt = 0; if (kt > Nt - 2) kt = Nt - 2;
    long NN = (long)Nx * Ny * Nz;
    long o = kx + (long)ky * Nx + (long)kz * Nx * Ny + (long)kt * NN;
    long on = kx + (long)ky * Nx + (long)(kz + 1) * Nx * Ny + (long)kt * NN;
    REAL one = (REAL)1.0f;
    REAL up = ((one - dx) * (one - dy) * (REAL)u[o] + (one - dx) * dy * (REAL)u[o + Nx]
               + dx * (one - dy) * (REAL)u[o + 1] + dx * dy * (REAL)u[o + Nx + 1]) * (one - dz)
            + ((one - dx) * (one - dy) * (REAL)u[on] + (one - dx) * dy * (REAL)u[on + Nx]
               + dx * (one - dy) * (REAL)u[on + 1] + dx * dy * (REAL)u[on + Nx + 1]) * dz;
    o += NN; on += NN;
    REAL un = ((one - dx) * (one - dy) * (REAL)u[o] + (one - dx) * dy * (REAL)u[o + Nx]
               + dx * (one - dy) * (REAL)u[o + 1] + dx * dy * (REAL)u[o + Nx + 1]) * (one - dz)
            + ((one - dx) * (one - dy) * (REAL)u[on] + (one - dx) * dy * (REAL)u[on + Nx]
               + dx * (one - dy) * (REAL)u[on + 1] + dx * dy * (REAL)u[on + Nx + 1]) * dz;
    return (one - dt) * up + dt * un;
}

/* Exported for the oracle's own pins (exactness on multilinear fields). */
double orc_interp3(const float* u, double x, double y, double z, int Nx, int Ny, int Nz, double Dx)
{
    return (double)interp3(u, (REAL)x, (REAL)y, (REAL)z, Nx, Ny, Nz, (REAL)Dx);
}

double orc_interp4(const float* u, double x, double y, double z, double t, int Nx, int Ny, int Nz,
                   int Nt, double Dx, double Dt)
{
    return (double)interp4(u, (REAL)x, (REAL)y, (REAL)z, (REAL)t, Nx, Ny, Nz, Nt, (REAL)Dx, (REAL)Dt);
}

/* ------------------------------------------------------------- thermal --- */
/* Thermal branch of APTrhs_psi (3DgrainMPI/APTPhaseField.cu:278-327) and the
 * 3Dgrain line law (3Dgrain/APTPhaseField.cu:303).  Literal double constants
 * are kept, so the float build promotes exactly where the reference does. */
static int thermal(const orc_params* P, REAL xi, REAL yj, REAL zk, const float* psi_mac,
                   REAL t, int use_iuc, REAL* Tout)
{
    REAL T = (REAL)0.0f;
    REAL G = P->G, z0 = P->z0, uc = P->uc, uc0 = P->uc0, lyd = P->lyd, lzd = P->lzd,
         r0 = P->r0, V = P->V;
    if (use_iuc) {
        T = -uc0 * 1e6 * t;
    } else if (P->thermal_type == 1) {
        T = G * (zk - z0) - uc * 1e6 * t;
    } else if (P->thermal_type == 2) {
        REAL a = yj - (REAL)0.5f * lyd, b = zk - z0 - lzd;
        REAL dist = RSQRT(a * a + b * b) - r0;
        T = -G * dist - uc * 1e6 * t;
    } else if (P->thermal_type == 3) {
        REAL lm = (r0 - z0) / (REAL)P->tan_angle;
        REAL x_start = 0;
        REAL xl = xi - x_start;
        REAL r0_x = z0 + (r0 - z0) * xl / lm;
        REAL z_tilt = z0 + lzd - zk;
        REAL a = yj - (REAL)0.5f * lyd;
        REAL dist = (REAL)P->cos_angle * (RSQRT(a * a + z_tilt * z_tilt) - r0_x);
        if (xl > lm + V * t * 1e6) return 1;
        T = -G * dist - uc * 1e6 * t;
    } else if (P->thermal_type == 4) {
        REAL xq = xi - V * t * 1e6;
        if (xq > (REAL)P->mp_len + (REAL)P->x0 || xq < (REAL)0.0f) return 1;
        REAL v = interp3(psi_mac, xq - (REAL)P->X0_mac, yj - (REAL)P->Y0_mac, zk - (REAL)P->Z0_mac,
                         P->Nx, P->Ny, P->Nz, (REAL)P->Dx_mac);
        if (P->order == 100) T = -v;
        else T = -G * v;
    } else if (P->thermal_type == 5) {
        REAL R = P->R;
        T = G * (zk - R * 1e6 * t - 2);
    }
    *Tout = T;
    return 0;
}

int orc_thermal(const orc_params* P, const float* x, const float* y, const float* z,
                const float* psi_mac, int i, int j, int k, double t, int use_iuc, double* T_out)
{
    REAL T = 0;
    int skip = thermal(P, x[i], y[j], z[k], psi_mac, (REAL)t, use_iuc, &T);
    *T_out = (double)T;
    return skip;
}

/* ---------------------------------------------------------------- gather --- */
/* One neighbourhood entry in scan order: id and the phi stored with it
 * (stencil row 0..6 for D,B,L,C,R,T,U or -1 when the point is outside the
 * 7-point stencil, 27-point gather only). */
typedef struct { int id; int row; REAL ph; } entry_t;

static int collect_entries(const orc_params* P, const REAL* ph, const int* arg, long C,
                           entry_t* e)
{
    long L = (long)P->fnx * P->fny * P->fnz;
    long sx = 1, sy = P->fnx, sz = (long)P->fnx * P->fny;
    int n = 0;
    if (!P->gather27) {
        /* stencil = -3..3 -> D,B,L,C,R,T,U (3DgrainMPI/APTPhaseField.cu:211-214) */
        const long off[7] = {-sz, -sy, -sx, 0, sx, sy, sz};
        for (int s = 0; s < 7; s++)
            for (int p = 0; p < P->K; p++) {
                long g = C + off[s] + p * L;
                e[n].id = arg[g]; e[n].row = s; e[n].ph = ph[g]; n++;
            }
    } else {
        /* z-major 27-point scan; phi kept for |xi|+|yi|+|zi|<=1
         * (3Dgrain/APTPhaseField.cu:241-268). */
        int row = 0;
        for (int zi = -1; zi <= 1; zi++)
            for (int yi = -1; yi <= 1; yi++)
                for (int xi = -1; xi <= 1; xi++) {
                    int in7 = abs(xi) + abs(yi) + abs(zi) <= 1;
                    for (int p = 0; p < P->K; p++) {
                        long g = C + xi * sx + yi * sy + zi * sz + p * L;
                        e[n].id = arg[g]; e[n].row = in7 ? row : -1; e[n].ph = ph[g]; n++;
                    }
                    if (in7) row++;
                }
    }
    return n;
}

/* First-appearance probe insert (3DgrainMPI/APTPhaseField.cu:215-232).  With
 * more than K distinct ids (undefined behaviour in the reference) the kept set
 * is chosen by the overflow policy:
 *   repeat K times: among entries whose id is not yet kept, take the first
 *   entry (scan order) with the largest phi; keep its id.
 * Kept ids then take slots in first-appearance order. */
static int insert_all(const orc_params* P, const entry_t* e, int n, const int* kept,
                      int* la, REAL lp[7][ORC_KMAX])
{
    int K = P->K;
    for (int a = 0; a < K; a++) {
        la[a] = -1;
        for (int s = 0; s < 7; s++) lp[s][a] = (REAL)-1.0f;
    }
    for (int q = 0; q < n; q++) {
        int id = e[q].id;
        if (id == -1) continue;
        if (kept) {
            int ok = 0;
            for (int u = 0; u < K; u++) if (kept[u] == id) { ok = 1; break; }
            if (!ok) continue;
        }
        int a;
        for (a = 0; a < K; a++) {
            if (la[a] == id) break;
            if (la[a] == -1) { la[a] = id; break; }
        }
        if (a == K) return 1;  /* a (K+1)-th distinct id */
        if (e[q].row >= 0) lp[e[q].row][a] = e[q].ph;
    }
    return 0;
}

static int gather(const orc_params* P, const entry_t* e, int n, int* la, REAL lp[7][ORC_KMAX])
{
    if (!insert_all(P, e, n, NULL, la, lp)) return 0;
    int K = P->K, kept[ORC_KMAX];
    for (int r = 0; r < K; r++) {
        int best = -1;
        for (int q = 0; q < n; q++) {
            if (e[q].id == -1) continue;
            int taken = 0;
            for (int u = 0; u < r; u++) if (kept[u] == e[q].id) { taken = 1; break; }
            if (taken) continue;
            if (best < 0 || e[q].ph > e[best].ph) best = q;
        }
        kept[r] = e[best].id;
    }
    insert_all(P, e, n, kept, la, lp);
    return 1;
}

/* ------------------------------------------------------------------ rhs --- */
long orc_rhs(const orc_params* P, const float* x, const float* y, const float* z,
             const float* R9, const float* psi_mac, const void* vph, const int* arg,
             void* vph_new, int* arg_new, double tt, int use_iuc)
{
    const REAL* ph = (const REAL*)vph;
    REAL* ph_new = (REAL*)vph_new;
    int fnx = P->fnx, fny = P->fny, fnz = P->fnz, K = P->K;
    long L = (long)fnx * fny * fnz;
    REAL t = (REAL)tt;
    const float ident[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    long overflow = 0;
#pragma omp parallel for schedule(static) reduction(+ : overflow)
    for (int k = 1; k < fnz - 1; k++) {
        entry_t e[27 * ORC_KMAX];
        int la[ORC_KMAX];
        REAL lp[7][ORC_KMAX];
        for (int j = 1; j < fny - 1; j++)
            for (int i = 1; i < fnx - 1; i++) {
                long C = i + (long)j * fnx + (long)k * fnx * fny;
                int n = collect_entries(P, ph, arg, C, e);
                overflow += gather(P, e, n, la, lp);
                for (int a = 0; a < K; a++) {
                    long o = C + a * L;
                    if (la[a] == -1) { arg_new[o] = -1; ph_new[o] = (REAL)-1.0f; continue; }
                    int g = la[a];
                    REAL phD = lp[0][a], phB = lp[1][a], phL = lp[2][a], phC = lp[3][a],
                         phR = lp[4][a], phT = lp[5][a], phU = lp[6][a];
                    REAL phxn = (phR - phL) * (REAL)0.5f;
                    REAL phyn = (phT - phB) * (REAL)0.5f;
                    REAL phzn = (phU - phD) * (REAL)0.5f;
                    const float* m = ident;
                    if (phC > LS) {
                        int tid = g <= P->num_theta ? g : (g % P->num_theta) + 1;
                        m = R9 + 9 * (long)tid;
                    }
                    REAL A2 = kine_ani(P, phxn, phyn, phzn, m);
                    REAL diff = phR + phL + phT + phB + phU + phD - 6 * phC;
                    REAL Tinterp = 0;
                    if (thermal(P, x[i], y[j], z[k], psi_mac, t, use_iuc, &Tinterp)) continue;
                    REAL Up = Tinterp / (REAL)P->L_cp;
                    REAL repul = (REAL)0.0f;
                    for (int p = 0; p < K; p++)
                        if (p != a)
                            repul += (REAL)0.25f * (lp[3][p] + (REAL)1.0f) * (lp[3][p] + (REAL)1.0f);
                    REAL hi = P->hi, lamd = P->lamd, dt = P->dt;
                    REAL rhs_psi = diff * hi * hi + ((REAL)1.0f - phC * phC) * phC
                        - lamd * Up * (((REAL)1.0f - phC * phC) * ((REAL)1.0f - phC * phC)
                                       - (REAL)0.5f * OMEGA * (phC + (REAL)1.0f) * repul);
                    REAL dphi = rhs_psi / A2;
                    REAL nv = phC + dt * dphi;
                    ph_new[o] = nv;
                    arg_new[o] = (nv < -0.9999) ? -1 : g;
                }
            }
    }
    return overflow;
}

/* -------------------------------------------------------- collect / ini --- */
void orc_collect(const orc_params* P, const void* vph, const int* arg, void* vphi, int* alpha)
{
    const REAL* ph = (const REAL*)vph;
    REAL* phi = (REAL*)vphi;
    long L = (long)P->fnx * P->fny * P->fnz;
    for (long C = 0; C < L; C++) {
        int am = 0;
        for (int s = 0; s < P->K; s++)
            if (ph[C + s * L] > ph[C + am * L]) am = s;
        if (ph[C + am * L] > LS) {
            phi[C] = ph[C + am * L];
            alpha[C] = arg[C + am * L];
        }
    }
}

void orc_ini(const orc_params* P, void* vph, int* arg, const float* phi0, const int* alpha0)
{
    REAL* ph = (REAL*)vph;
    long L = (long)P->fnx * P->fny * P->fnz;
    for (long q = 0; q < L * P->K; q++) { ph[q] = (REAL)-1.0f; arg[q] = -1; }
    for (long C = 0; C < L; C++)
        if (phi0[C] > LS) { ph[C] = (REAL)phi0[C]; arg[C] = alpha0[C]; }
}

/* ------------------------------------------------------ moving domain --- */
void orc_plane_sums(const orc_params* P, const void* vph, int64_t* sums)
{
    const REAL* ph = (const REAL*)vph;
    long A = (long)P->fnx * P->fny, L = A * P->fnz;
    for (int k = 0; k < P->fnz; k++) {
        int64_t acc = 0;
        for (int s = 0; s < P->K; s++)
            for (long q = 0; q < A; q++)
                acc += (int64_t)llrint(ldexp((double)ph[s * L + k * A + q], 40));
        sums[k] = acc;
    }
}

int orc_tip_scan(const orc_params* P, const int64_t* sums, int cur_tip)
{
    double denom = (double)P->K * P->fnx * P->fny;
    while (ldexp((double)sums[cur_tip], -40) / denom > LS && cur_tip < P->fnz - 1) cur_tip++;
    return cur_tip;
}

void orc_move_frame(const orc_params* P, void* vo, int* ao, void* vn, int* an, float* z,
                    const int* alpha, int* alpha_full, int* loss_area, int move_count)
{
    REAL* po = (REAL*)vo;
    REAL* pn = (REAL*)vn;
    int fnx = P->fnx, fny = P->fny, fnz = P->fnz;
    long A = (long)fnx * fny;
    float* zb = (float*)malloc(sizeof(float) * fnz);
    for (int k = 1; k < fnz - 2; k++) zb[k] = z[k + 1];
    zb[fnz - 2] = 2 * z[fnz - 2] - z[fnz - 3];
    for (int s = 0; s < P->K; s++)
        for (int k = 1; k < fnz - 1; k++)
            for (int j = 1; j < fny - 1; j++)
                for (int i = 1; i < fnx - 1; i++) {
                    long C = idx4(P, i, j, k, s);
                    if (k < fnz - 2) { pn[C] = po[C + A]; an[C] = ao[C + A]; }
                    else { pn[C] = (REAL)-1.0f; an[C] = -1; }
                }
    for (long q = A; q < 2 * A; q++) {
        alpha_full[(long)move_count * A + q] = alpha[q];
        if (alpha[q] >= 1 && alpha[q] <= P->num_theta) loss_area[alpha[q] - 1] += 1;
    }
    for (int k = 1; k < fnz - 1; k++) z[k] = zb[k];
    for (int s = 0; s < P->K; s++)
        for (int k = 1; k < fnz - 1; k++)
            for (int j = 1; j < fny - 1; j++)
                for (int i = 1; i < fnx - 1; i++) {
                    long C = idx4(P, i, j, k, s);
                    po[C] = pn[C]; ao[C] = an[C];
                }
    free(zb);
}

/* --------------------------------------------------------- nucleation --- */
/* splitmix64 finaliser (Steele, Lea & Flood, OOPSLA 2014). */
static uint64_t mix64(uint64_t v)
{
    v = (v ^ (v >> 30)) * 0xbf58476d1ce4e5b9ULL;
    v = (v ^ (v >> 27)) * 0x94d049bb133111ebULL;
    return v ^ (v >> 31);
}

/* Counter-based draw per (seed, coarse cell, call): u in (0,1] with 24-bit
 * resolution and a positive grain id in [1, 2^31-2]. */
void orc_draw(uint64_t seed, uint64_t cell, uint64_t call, float* u, int* id)
{
    uint64_t h = mix64(seed + 0x9e3779b97f4a7c15ULL);
    h = mix64(h ^ (cell * 0xd1b54a32d192ed03ULL));
    h = mix64(h ^ (call * 0x8cb92ba72f3d8dd7ULL + 0x632be59bd9b4e019ULL));
    *u = (float)((h >> 40) + 1) * (1.0f / 16777216.0f);
    *id = 1 + (int)(((h & 0xffffffffULL) * 0x7ffffffeULL) >> 32);
}

/* tanh ball seeded into slot 0 (3DgrainMPI/APTPhaseField.cu:515-529). */
void orc_nucl_ball(const orc_params* P, float* ball)
{
    int p = P->pts_cell, w = 2 * p + 1, q = 0;
    for (int lk = -p; lk <= p; lk++)
        for (int lj = -p; lj <= p; lj++)
            for (int li = -p; li <= p; li++) {
                float pc = (float)p;
                float d = P->dx * ((1.0f + pc) / 2.0f - sqrtf((float)(li * li + lj * lj + lk * lk)));
                ball[q++] = tanhf(d / P->sqrt2);
            }
    (void)w;
}

/* nuncl_possibility (3DgrainMPI/APTPhaseField.cu:446-455). */
static REAL nucl_prob(const orc_params* P, REAL delT, REAL d_delT, REAL nmax)
{
    REAL mu = P->undcool_mean, sd = P->undcool_std, r = P->nuc_rad;
    REAL slope = (REAL)-0.5f * (delT - mu) * (delT - mu) / sd / sd;
    slope = REXP(slope);
    REAL s2p = RSQRT(2.0f * M_PI);
    REAL density = nmax / (s2p * sd) * slope * d_delT;
    return (REAL)8.0f * r * r * r * density;
}

double orc_nucl_prob(const orc_params* P, double delT, double d_delT, double nmax)
{
    return (double)nucl_prob(P, (REAL)delT, (REAL)d_delT, (REAL)nmax);
}

/* init_nucl_status (3DgrainMPI/APTPhaseField.cu:420-444): note the block
 * centre omits the halo offset that add_nucl adds (kept). */
void orc_init_nucl_status(const orc_params* P, const void* vph, int* status, int cnx, int cny, int cnz)
{
    const REAL* ph = (const REAL*)vph;
    int p = P->pts_cell, w = 2 * p + 1;
    for (int k = 0; k < cnz; k++)
        for (int j = 0; j < cny; j++)
            for (int i = 0; i < cnx; i++) {
                long g = (long)(w * k + p) * P->fnx * P->fny + (long)(w * j + p) * P->fnx + (w * i + p);
                status[i + (long)j * cnx + (long)k * cnx * cny] = ph[g] > LS ? 1 : 0;
            }
}

/* add_nucl (3DgrainMPI/APTPhaseField.cu:459-536) with the counter-based draw. */
int orc_add_nucl(const orc_params* P, const float* x, const float* y, const float* z,
                 const float* T3d, const float* ball, void* vph, int* arg, int* status,
                 int cnx, int cny, int cnz, double dt_nuc_d, double t_d, int use_iuc,
                 uint64_t seed, uint64_t call_index)
{
    REAL* ph = (REAL*)vph;
    int p = P->pts_cell, w = 2 * p + 1, events = 0;
    long A = (long)P->fnx * P->fny, L = A * P->fnz;
    REAL dt = (REAL)dt_nuc_d, t = (REAL)t_d;
    for (int k = 0; k < cnz; k++)
        for (int j = 0; j < cny; j++)
            for (int i = 0; i < cnx; i++) {
                long C = i + (long)j * cnx + (long)k * cnx * cny;
                if (status[C] != 0) continue;
                int gi = w * i + p + 1, gj = w * j + p + 1, gk = w * k + p + 1;
                long gC = (long)gk * A + (long)gj * P->fnx + gi;
                for (int s = 0; s < P->K; s++)
                    if (ph[gC + s * L] > LS) status[C] = 1;
                if (status[C] != 0) continue;
                REAL posb;
                if (use_iuc) {
                    REAL uc0 = P->uc0;
                    REAL delT = uc0 * 1e6 * (t + dt);
                    REAL d_delT = uc0 * 1e6 * dt;
                    posb = nucl_prob(P, delT, d_delT, (REAL)P->nuc_Nmax0);
                } else {
                    REAL xx = (REAL)x[gi] - (REAL)P->X0_mac, yy = (REAL)y[gj] - (REAL)P->Y0_mac,
                         zz = (REAL)z[gk] - (REAL)P->Z0_mac;
                    REAL Tc = interp4(T3d, xx, yy, zz, t - (REAL)P->t0_mac, P->Nx, P->Ny, P->Nz,
                                      P->Nt, (REAL)P->Dx_mac, (REAL)P->Dt_mac);
                    REAL Tcd = interp4(T3d, xx, yy, zz, t + dt - (REAL)P->t0_mac, P->Nx, P->Ny,
                                       P->Nz, P->Nt, (REAL)P->Dx_mac, (REAL)P->Dt_mac);
                    posb = nucl_prob(P, -Tcd, Tc - Tcd, (REAL)P->nuc_Nmax);
                }
                float u; int id;
                orc_draw(seed, (uint64_t)C, call_index, &u, &id);
                if (u < posb) {
                    int q = 0;
                    for (int lk = -p; lk <= p; lk++)
                        for (int lj = -p; lj <= p; lj++)
                            for (int li = -p; li <= p; li++) {
                                long o = gC + lk * A + (long)lj * P->fnx + li;
                                ph[o] = (REAL)ball[q++];
                                if (ph[o] > (REAL)-0.9999f) arg[o] = id;
                            }
                    status[C] = 1;
                    events++;
                }
            }
    return events;
}
