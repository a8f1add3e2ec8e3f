/*
 * apt_oracle.h -- CPU restatement of the cuPF 3-D APT phase-field step.
 *
 * TEST INFRASTRUCTURE ONLY.  This oracle is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product (cupf_amd) never links, imports or calls it.
 *
 * One source, two builds:
 *   liboracle_f32.so : REAL=float, -ffp-contract=off, no fast-math.  Keeps the
 *                      reference's implicit double promotions (the 1e6 thermal
 *                      factors, the -0.9999 kill threshold, the LS=-0.995 tests).
 *   liboracle_f64.so : REAL=double; the same float parameters, double arithmetic.
 *
 * Layout (3DgrainMPI/devicefunc.cu_inl:24-29): element (i,j,k,slot) sits at
 *   i + j*fnx + k*fnx*fny + slot*fnx*fny*fnz.
 */
#ifndef APT_ORACLE_H
#define APT_ORACLE_H

#include <stdint.h>

#define ORC_KMAX 16

/* Flat parameter block.  Field meaning follows GlobalConstants
 * (3DgrainMPI/params.h:10-56); cone trig is evaluated once on the host. */
typedef struct orc_params {
    int fnx, fny, fnz;      /* full grid incl. the 1-cell halo             */
    int K;                  /* active slots per point (APT_NUM_PF)         */
    int num_theta;          /* NG: orientation rows 0..NG                  */
    int bcx, bcy, bcz;      /* 0 = no-flux mirror, 1 = periodic            */
    int thermal_type;       /* 0 none,1 line,2 cylinder,3 cone,4 paraboloid,
                               5 = 3Dgrain line formula G*(z-R*1e6*t-2)    */
    int gather27;           /* 0: 7-point id gather, 1: 27-point (3Dgrain) */
    float dt, hi, lamd, L_cp, kin_delta, eps;
    float G, R, uc, uc0, z0, r0, lyd, lzd, V, mp_len, x0, order;
    float tan_angle, cos_angle;
    /* macro grid (type-4 thermal field, nucleation temperature) */
    int Nx, Ny, Nz, Nt;
    float X0_mac, Y0_mac, Z0_mac, t0_mac, Dx_mac, Dt_mac;
    /* nucleation */
    int pts_cell;           /* p: blocks are (2p+1)^3                       */
    float nuc_rad, undcool_mean, undcool_std, nuc_Nmax, nuc_Nmax0, dx, sqrt2;
} orc_params;

#ifdef __cplusplus
extern "C" {
#endif

/* Sequential halo fill: z faces, then y faces, then x faces
 * (3DgrainMPI/APTPhaseField.cu:93-174 applied in a fixed order). */
void orc_set_bc(const orc_params* P, void* ph, int* arg);

/* One APTrhs_psi application src -> dst (3DgrainMPI/APTPhaseField.cu:178-354).
 * Reads the halo of src as stored (call orc_set_bc first).  Returns the
 * number of points whose neighbourhood held more than K distinct ids. */
long orc_rhs(const orc_params* P, const float* x, const float* y, const float* z,
             const float* R9, const float* psi_mac,
             const void* ph, const int* arg, void* ph_new, int* arg_new,
             double t, int use_initial_uc);

/* APTcollect_PF (3DgrainMPI/APTPhaseField.cu:55-75). */
void orc_collect(const orc_params* P, const void* ph, const int* arg, void* phi, int* alpha);

/* set_minus1 + memset(-1) + APTini_PF (3DgrainMPI/APTPhaseField.cu:78-89). */
void orc_ini(const orc_params* P, void* ph, int* arg, const float* phi0, const int* alpha0);

/* kine_ani with a row-major 3x3 rotation (3DgrainMPI/APTPhaseField.cu:36-52). */
double orc_kine_ani(const orc_params* P, double ux, double uy, double uz, const float* R9);

/* Thermal value at (i,j,k,t); returns 1 when the reference `continue`s. */
int orc_thermal(const orc_params* P, const float* x, const float* y, const float* z,
                const float* psi_mac, int i, int j, int k, double t, int use_initial_uc,
                double* T_out);

/* interp3Dtemperature / interp4Dtemperature (3DgrainMPI/devicefunc.cu_inl:32-125),
 * indices clamped into the macro grid. */
double orc_interp3(const float* u, double x, double y, double z, int Nx, int Ny, int Nz, double Dx);
double orc_interp4(const float* u, double x, double y, double z, double t, int Nx, int Ny, int Nz,
                   int Nt, double Dx, double Dt);

/* Per-plane sum of phi over all slots and (i,j) incl. halo as 2^-40 fixed point
 * (ave_x, 3DgrainMPI/PhaseField.cu:366-378, made order independent). */
void orc_plane_sums(const orc_params* P, const void* ph, int64_t* sums);

/* tip_mvf scan (3DgrainMPI/PhaseField.cu:382-399) on fixed-point sums. */
int orc_tip_scan(const orc_params* P, const int64_t* sums, int cur_tip);

/* APTmove_frame + APTcopy_frame (3DgrainMPI/APTPhaseField.cu:357-410):
 * shift every slot down one plane in both buffers, archive alpha plane 1. */
void orc_move_frame(const orc_params* P, void* ph_old, int* arg_old, void* ph_new, int* arg_new,
                    float* z, const int* alpha, int* alpha_full, int* loss_area, int move_count);

/* Nucleation (3DgrainMPI/APTPhaseField.cu:420-536) with the counter-based draw
 * orc_draw(); ball[] is the (2p+1)^3 tanh template. Returns #nucleation events. */
void orc_init_nucl_status(const orc_params* P, const void* ph, int* status, int cnx, int cny, int cnz);
int orc_add_nucl(const orc_params* P, const float* x, const float* y, const float* z,
                 const float* T3d, const float* ball, void* ph, int* arg, int* status,
                 int cnx, int cny, int cnz, double dt_nuc, double t, int use_initial_uc,
                 uint64_t seed, uint64_t call_index);
void orc_nucl_ball(const orc_params* P, float* ball);
double orc_nucl_prob(const orc_params* P, double delT, double d_delT, double nmax);
void orc_draw(uint64_t seed, uint64_t cell, uint64_t call, float* u, int* id);

#ifdef __cplusplus
}
#endif
#endif
