/*
 * pf_capi.h -- C ABI of the MI355X-native APT phase-field step (cupf_amd).
 *
 * The reference has no FFI: its step sits behind the C++ class surface
 * PDE -> PhaseField -> APTPhaseField (3DgrainMPI/cuPDE.h:4-21,
 * 3DgrainMPI/PhaseField.h:11-59, 3DgrainMPI/APTPhaseField.h:11-30).  Every
 * device-side member of that surface maps onto one entry point below; the C++
 * drop-in class in cupf_amd/host/ calls only these.
 *
 * Conventions
 *   - Plain pointers and sizes, no exceptions across the ABI.
 *   - Every call returns int: PF_OK (0) or a negative PF_E* code; the message
 *     of the last failure on the calling thread is pf_last_error().
 *   - Host memory is caller-owned; device memory is context-owned.
 *   - A context is bound to one device and one HIP stream and must be driven
 *     from one host thread at a time.
 *   - Field layout (host and device, 3DgrainMPI/devicefunc.cu_inl:24-29):
 *     element (i,j,k,slot) at i + j*fnx + k*fnx*fny + slot*fnx*fny*fnz, with a
 *     1-cell halo on every face.  phi is float32, arg (grain id) is int32 with
 *     -1 meaning an empty slot.
 */
#ifndef PF_CAPI_H
#define PF_CAPI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PF_OK 0
#define PF_E_INVALID (-1)   /* bad argument / size / unsupported K           */
#define PF_E_HIP (-2)       /* HIP runtime error                              */
#define PF_E_NOMEM (-3)     /* device allocation failed                       */
#define PF_E_STATE (-4)     /* call out of order (e.g. step before init)      */

/* pf_step flags */
#define PF_STEP_INITIAL_UC 1  /* useInitialUnderCooling=true (preMt phase)    */
#define PF_STEP_NUCLEATE 2    /* add_nucl after every odd step                */

/* Thermal field types: 0..4 as selected by input-file name in
 * 3DgrainMPI/PhaseFieldCpu.cpp:201-242; 5 is the 3Dgrain line law
 * T = G*(z - R*1e6*t - 2) (3Dgrain/APTPhaseField.cu:303). */
#define PF_THERMAL_NONE 0
#define PF_THERMAL_LINE 1
#define PF_THERMAL_CYLINDER 2
#define PF_THERMAL_CONE 3
#define PF_THERMAL_PARABOLOID 4
#define PF_THERMAL_LINE_3DGRAIN 5

/* Device-side parameter block: the subset of GlobalConstants
 * (3DgrainMPI/params.h:10-56) that the step kernels read from `cP`
 * (3DgrainMPI/APTPhaseField.cu:34), plus tau0 and a draw seed. */
typedef struct pf_params {
    int fnx, fny, fnz;      /* full grid incl. 1-cell halo                    */
    int K;                  /* active slots per point: 5 or 8                 */
    int num_theta;          /* NG; orientation rows 0..NG                     */
    int bcx, bcy, bcz;      /* 0 = no-flux mirror, 1 = periodic               */
    int thermal_type;       /* PF_THERMAL_*                                   */
    int gather27;           /* 0: 7-point id gather (3DgrainMPI);
                               1: 27-point id gather (3Dgrain)               */
    float dt, hi, lamd, L_cp, kin_delta, eps;
    float G, R, uc, uc0, z0, r0, lyd, lzd, V, mp_len, x0, order;
    float tan_angle, cos_angle;   /* tanf/cosf(angle), evaluated on the host  */
    int Nx, Ny, Nz, Nt;           /* macro thermal grid                       */
    float X0_mac, Y0_mac, Z0_mac, t0_mac, Dx_mac, Dt_mac;
    int pts_cell;                 /* nucleation block half-width p           */
    float nuc_rad, undcool_mean, undcool_std, nuc_Nmax, nuc_Nmax0, dx, sqrt2;
    float tau0;                   /* time of step s is (float)s*dt*tau0       */
    uint64_t seed;                /* nucleation draw seed                     */
} pf_params;

typedef struct pf_ctx pf_ctx;

/* Message of the last failing call on this thread ("" if none). */
const char* pf_last_error(void);

/* Lifecycle.  Replaces APTPhaseField::cudaSetup allocation
 * (3DgrainMPI/APTPhaseField.cu:569-613) and the destructor (:540-566).
 * Allocates both ping-pong K-slot buffers for phi and arg on `device`. */
int pf_create(const pf_params* params, int device, pf_ctx** out);
int pf_destroy(pf_ctx* ctx);

/* Coordinate vectors x[fnx], y[fny], z[fnz] (3DgrainMPI/APTPhaseField.cu:603-606). */
int pf_set_coords(pf_ctx* ctx, const float* x, const float* y, const float* z);

/* Orientation table: nrows row-major 3x3 rotations, row g for grain id g
 * (0..NG).  The reference's (theta_x, theta_z) tables
 * (3DgrainMPI/APTPhaseField.cu:623-634, used at :258-273) are the special case
 * built by pf_orientations_from_angles. */
int pf_set_orientations(pf_ctx* ctx, const float* R9, int nrows);

/* Host helper, no device work: R9 rows from the reference's angle array
 * theta[2*NG+1] = [0, thx(1..NG), thz(1..NG)] with sinf/cosf as in
 * 3DgrainMPI/PhaseFieldCpu.cpp:379-385.  R9 must hold 9*(NG+1) floats. */
int pf_orientations_from_angles(const float* theta, int num_theta, float* R9);

/* Macro thermal inputs: psi_mac[Nx*Ny*Nz] (type-4 field) and T3d[Nx*Ny*Nz*Nt]
 * (nucleation temperature), 3DgrainMPI/APTPhaseField.cu:617-631.  Either may be NULL. */
int pf_set_thermal_macro(pf_ctx* ctx, const float* psi_mac, const float* T3d);

/* Initial state from the single-field phi0/alpha0 of initField: every slot
 * -1/-1, then slot 0 = (phi0, alpha0) where phi0 > -0.995, halo included, in
 * both buffers (set_minus1 + cudaMemset + APTini_PF,
 * 3DgrainMPI/APTPhaseField.cu:727-728,817-818).  Also stores phi0/alpha0 as
 * the context's single-field phi/alpha (phi_old/alpha_m). */
int pf_init_from_phi_alpha(pf_ctx* ctx, const float* phi0, const int* alpha0);

/* Raw K-slot state transfer; which = 0 current buffer, 1 the other one.
 * For parity tests and checkpoint/restart. */
int pf_upload_state(pf_ctx* ctx, int which, const float* ph, const int* arg);
int pf_download_state(pf_ctx* ctx, int which, float* ph, int* arg);

/* Fill the halo of one buffer from the interior (APTsetBC3D,
 * 3DgrainMPI/APTPhaseField.cu:93-174,700-707).  Every halo cell, edges and
 * corners included, receives the value at its per-axis mapped interior point. */
int pf_apply_bc(pf_ctx* ctx, int which);

/* Advance nsteps time steps.  Step s = first_step .. first_step+nsteps-1 runs
 * APTrhs_psi (3DgrainMPI/APTPhaseField.cu:178-354) from the current buffer
 * into the other one at t = (float)s*dt*tau0, with the boundary condition
 * fused into the neighbour loads, then swaps the buffers.  With
 * PF_STEP_NUCLEATE, add_nucl (:459-536) runs after every odd step s with
 * dt_nuc = 2*dt*tau0 (:866-870).  This is the loop body of
 * APTPhaseField::evolve (:852-922) without its host-side QoI work. */
int pf_step(pf_ctx* ctx, int nsteps, int first_step, int flags);

/* pf_step bracketed by HIP events on the context's stream; *ms = elapsed. */
int pf_step_timed(pf_ctx* ctx, int nsteps, int first_step, int flags, float* ms);

/* Diagnostic: pf_step with HIP events on the context's stream around every
 * launch of the work-list plan kernel, of the main step kernel (with, in
 * steps whose neighbourhoods overflow, the edge kernel that runs beside it on
 * a second stream) and of the overflow fix-up kernel, each launched on its own
 * (pf_step sends a fix-up and the next plan as one launch); the summed
 * elapsed times are returned in ms (any pointer may be NULL).  No nucleation.  No counterpart in the reference (its
 * only timer is the wall clock around the loop, 3DgrainMPI/APTPhaseField.cu:850,940). */
int pf_step_profiled(pf_ctx* ctx, int nsteps, int first_step, int flags, float* ms_main,
                     float* ms_fixup, float* ms_plan);

/* pf_collect flags */
#define PF_COLLECT_CLEAR_ALPHA 1  /* zero alpha first (the cudaMemset at
                                     3DgrainMPI/APTPhaseField.cu:663,809,820,907);
                                     without it stale alpha persists (:683,694,899) */
#define PF_COLLECT_NO_BC 2        /* collect the buffer's halo as stored instead
                                     of filling it first */

/* APTcollect_PF (3DgrainMPI/APTPhaseField.cu:55-75) on buffer `which`
 * (0 current, 1 other) into the context's single-field phi/alpha, after a
 * halo fill unless PF_COLLECT_NO_BC.  phi_out/alpha_out (fnx*fny*fnz) may be
 * NULL to keep the result on the device. */
int pf_collect(pf_ctx* ctx, int which, int flags, float* phi_out, int* alpha_out);

/* Device-side QoIs on the context's alpha field as the last pf_collect left
 * it; only the results are copied to the host.
 *
 * QOI::calculateQoIs (3DgrainMPI/QOI.cpp:198-231): for grain ids
 * 1..max_id, counts[id] = interior points with that alpha and first[id] =
 * scan index (k-major, then j, then i, from 0) of the first such point, or
 * INT_MAX if none; ordering the ids with a count by first[] gives the
 * reference's first-seen order.  counts and first hold max_id + 1 ints;
 * *n_above = interior points whose alpha exceeds max_id (nucleated ids). */
int pf_qoi_volume(pf_ctx* ctx, int max_id, int* counts, int* first, long long* n_above);

/* No counterpart in the reference.  pf_qoi_volume for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): the same results, with the per-grain
 * buffers in a fixed device array of this library instead of allocations
 * (PF_NOALLOC_MAX_ID + 1 counts and first indices plus one 64-bit count:
 * 8 MiB per GPU, held until the process ends; its address is resolved once
 * per GPU by pf_create).  max_id above PF_NOALLOC_MAX_ID is rejected with
 * PF_E_INVALID.  Every call on one GPU uses that GPU's array, so calls that
 * overlap on one GPU (from several threads or streams) race. */
#define PF_NOALLOC_MAX_ID 1048575
int pf_qoi_volume_noalloc(pf_ctx* ctx, int max_id, int* counts, int* first, long long* n_above);

/* QOI3D::searchJunctionsOnImage (3DgrainMPI/QOI.cpp:233-311): solid interior
 * points whose 19-point neighbourhood holds no liquid and exactly 4 distinct
 * ids, none more than 10 times.  Up to cap rows of 8 ints (i, j, k, largest
 * count, the 4 ids ascending) in the reference's scan order (k, then j, then
 * i); *count = junctions found (rows past cap are not written).  The ids are
 * sorted because the reference's order comes from unordered_map iteration. */
int pf_qoi_junctions3d(pf_ctx* ctx, int* rows, int cap, int* count);

/* QOI::calculateLineQoIs (3DgrainMPI/QOI.cpp:38-95) without its loss-area
 * term and printout: *new_tip = the tip plane after cur_tip (the first plane
 * above cur_tip whose interior holds a liquid point, minus one; fnz - 2 if no
 * plane up to fnz - 2 holds liquid, where the reference would read past the
 * field), total[g] / extra[g] = interior points with alpha = g + 1 in all
 * planes / in planes above *new_tip, for g < num_theta (the reference's
 * out-of-range writes for larger ids are skipped).  The caller adds the
 * moving domain's loss area to total. */
int pf_qoi_line(pf_ctx* ctx, int cur_tip, int* new_tip, int* total, int* extra);

/* QOI::sampleHeights (3DgrainMPI/QOI.cpp:162-185), called by moveDomain
 * (3DgrainMPI/APTPhaseField.cu:696): *new_tip = the first plane above
 * cur_tip whose interior holds a liquid point (alpha 0) on the collected
 * alpha, minus one (fnz - 2 if none up to fnz - 2).  Only fnz plane flags
 * cross PCIe instead of the whole alpha field. */
int pf_qoi_tip(pf_ctx* ctx, int cur_tip, int* new_tip);

/* Liquid-start re-initialisation (3DgrainMPI/APTPhaseField.cu:809-818): both
 * buffers set to -1, then slot 0 = (initial phi0, collected alpha) where
 * phi0 > -0.995.  Follows a pf_collect with PF_COLLECT_CLEAR_ALPHA. */
int pf_reinit_from_collect(pf_ctx* ctx);

/* Moving domain (3DgrainMPI/APTPhaseField.cu:639-697, PhaseField.cu:366-399).
 * pf_md_enable allocates alpha_full[fnx*fny*fnz_full] and loss_area[NG].
 * pf_md_tip: per-plane phi sums of the current buffer (order-independent
 * fixed point), then the tip scan from *tip.
 * pf_md_shift: one APTmove_frame + APTcopy_frame with archive index move_count. */
int pf_md_enable(pf_ctx* ctx, int fnz_full);
int pf_md_tip(pf_ctx* ctx, int* tip);
int pf_md_shift(pf_ctx* ctx, int move_count);
int pf_md_download(pf_ctx* ctx, int* alpha_full, int* loss_area, float* z);

/* Nucleation set-up (3DgrainMPI/APTPhaseField.cu:733-745,781): mode 0 zeroes
 * the status array, mode 1 sets it from the stored single-field phi
 * (init_nucl_status).  Computes cnx/cny/cnz and the tanh ball template. */
int pf_nucl_reset(pf_ctx* ctx, int mode);
/* Number of nucleation events so far. */
int pf_nucl_events(pf_ctx* ctx, long long* events);

/* Points whose neighbourhood held more than K distinct ids so far. */
int pf_overflow_count(pf_ctx* ctx, long long* count);

/* Wait for all work on the context's stream. */
int pf_sync(pf_ctx* ctx);

/* Diagnostic: activity map of buffer `which` (0 current, 1 other).  The step
 * skips z planes of 64x4 column tiles whose stencil reaches only empty slot
 * pairs (-1, -1.0) and whose destination is empty too.  Tiles are numbered
 * t = tx + ty*ceil((fnx-2)/64), tx = (i-1)/64, ty = (j-1)/4 over interior
 * columns.  map (nullable, fnz*ntiles bytes, index k*ntiles + t) receives 1
 * for a tile plane that may hold a non-empty slot; *active their number,
 * *total all interior tile planes.  No counterpart in the reference, which
 * updates every point; results are identical either way.  Environment
 * variable CUPF_NO_ACTIVITY=1 at pf_create turns the skipping off. */
int pf_activity(pf_ctx* ctx, int which, unsigned char* map, long long* active, long long* total);

/* ---- Batched ensembles --------------------------------------------------
 * No counterpart in the reference, which runs one simulation per process
 * (3Dgrain/createcmd:1 launches them one after another).  A batch steps
 * several contexts of one device (same K and gather variant; grids, thermal
 * parameters and seeds may differ) with one plan, one persistent step, one
 * fix-up and one advance kernel per time step over all of them, so small
 * simulations share the GPU instead of each being launch-bound.  Every member
 * gets exactly the bits pf_step would give it: batching changes the schedule,
 * not the arithmetic.  Members stay ordinary contexts: set them up, upload,
 * download, collect and step them singly as before; a batch only borrows
 * them during pf_batch_step. */
typedef struct pf_batch pf_batch;

/* pf_batch_step flags (beside PF_STEP_INITIAL_UC) */
#define PF_BATCH_GRAPH 4      /* replay groups of 8 steps from one captured hipGraph */
#define PF_BATCH_AUTOTUNE 8   /* every 400 steps, time 4 real steps with each z-chunk
                                 depth 2, 4, 8, 16, twice (forward, then reverse
                                 order), and keep the fastest for the next 400 steps
                                 (a tuning round needs >= 56 steps left) */

/* n <= 1024 distinct contexts. */
int pf_batch_create(pf_ctx* const* ctxs, int n, pf_batch** out);
int pf_batch_destroy(pf_batch* b);
/* nsteps steps of every member; member i runs steps first_step[i] ..
 * first_step[i]+nsteps-1 (as pf_step's first_step).  No nucleation. */
int pf_batch_step(pf_batch* b, int nsteps, const int* first_step, int flags);
/* z planes per work item (2..61); default 4, or CUPF_KCHUNK at creation. */
int pf_batch_set_kchunk(pf_batch* b, int kchunk);
int pf_batch_get_kchunk(pf_batch* b, int* kchunk);

/* Diagnostic: bandwidth of a plain device-to-device copy kernel over `bytes`
 * (read + write counted), iters launches; the reference point for the HBM
 * roofline and for scaling HBM byte counters.  No counterpart in the reference. */
int pf_copy_bandwidth(int device, unsigned long long bytes, int iters, float* gbs);

/* Library build information, e.g. "cupf_amd gfx950 K={5,8}". */
const char* pf_build_info(void);

#ifdef __cplusplus
}
#endif
#endif
