// apt_internal.h -- device-side argument blocks shared by the kernels and the
// C-ABI layer of cupf_amd.  Not part of the public interface.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cupf {

// LS = -0.995 (3DgrainMPI/APTPhaseField.cu:25) is a double literal in the
// reference: every `phi > LS` test compares in double.
constexpr double kLS = -0.995;
// Slot kill threshold (3DgrainMPI/APTPhaseField.cu:341), also a double compare.
constexpr double kKill = -0.9999;
// The same two comparisons in float, exact for every float x:
// (double)x > -0.995 <=> x > -0.995f, because -0.995f is the largest float not
// above -0.995; (double)x < -0.9999 <=> x < -0.9999f, because -0.9999f is the
// smallest float not below -0.9999 (checked exhaustively on [-1, -0.99] by
// tests/test_oracle.py).
constexpr float kLSf = -0.995f;
constexpr float kKillf = -0.9999f;

// Fix-up list entry flag: the point's entries are in the record buffer at the
// entry's position (point indices stay below 2^31 since K*L does).
constexpr unsigned kRecFlag = 0x80000000u;

// The fix-up list is allocated with kFixHeader words in front of entry 0.
// Words -4 .. -1 describe the record buffer: record capacity, 0, and the
// buffer's address (low, high word).  The step kernel reads them from memory
// only in waves that hand a point to the fix-up, so they cost it no uniform
// registers (it is at its scalar register limit).
constexpr int kFixHeader = 64;

// Fix-up list length from which a step counts as overflowing for the host's
// choice of step kernel variant: below it the plain variant's fix-up (every
// overflowed point, gathered again through L1/L2) costs less than the
// overflow variant's records, keep test and edge kernel.
constexpr unsigned kOvfHeavy = 32768u;

// Leftover columns at the +x end narrower than this take the direct path;
// wider ones form a partial LDS tile (its idle lanes cost less than the direct
// path's per-point neighbour loads).
constexpr int kPartialMin = 16;

// Everything one APT step launch reads.  Passed by value (kernel argument
// segment -> SGPRs), so no constant-memory upload per step.
struct StepArgs {
    const float* ph;       // source buffer, K slot planes
    const int* arg;
    float* ph_new;         // destination buffer
    int* arg_new;
    const float* x;
    const float* y;
    const float* z;
    const float* R12;      // orientation rows, 12 floats per grain (9 used)
    const float* psi_mac;  // type-4 thermal field
    unsigned long long* overflow;  // count of points with > K distinct ids
    unsigned* fix_list;    // interior points left to the fix-up pass (capacity: interior points);
                           // kRecFlag set: the entries of the point are in record (list position);
                           // preceded by the record header (kFixHeader)
    // Counters of this step: fix_count[0] entries in fix_list; wl_count (below)
    // points into the same 4-word set.  Steps take three sets in turn;
    // cnt_next is the next step's, cleared by this step's plan kernel (its
    // last user, two steps back, has finished; this step's fix-up, which may
    // run beside the next step's plan, keeps its own).
    unsigned* fix_count;
    unsigned* cnt_next;
    int* sink_i;           // 64-lane write sinks for lanes that must not store
    float* sink_f;
    // Activity map: one word per (column tile, z plane) and buffer.  A tile
    // plane of a buffer may hold a slot other than the empty pair (-1, -1.0f)
    // only if its word is >= that buffer's epoch.  Epochs come from one
    // increasing counter, so stale words never need clearing.
    const unsigned* act_src;
    unsigned* act_dst;
    unsigned ep_src;       // epoch of the source buffer
    unsigned ep_dst;       // epoch the destination buffer had before this step
    unsigned ep_new;       // epoch the destination buffer gets from this step
    int use_act;           // 0: process every plane (the map is still maintained)
    int ntx, nty, ntiles;  // activity tiles (64x4, partial ones included)
    // Work of one step: nzchunk * edge_items edge items (the leftover columns:
    // kEdgePoints points or one edge tile each; the first work indices, or a
    // kernel of their own in a single overflow step), and the tile chunks
    // listed by the plan kernel:
    // wl_item = z chunk * items_chunk + full tile, with the chunk's plane masks.
    // Chunks with more than half their planes to process are listed from the
    // front (wl_count[0] of them), the others from the back of the list
    // (wl_light[0] of them, down from nzchunk * items_chunk - 1), so the long
    // items are taken first and the short ones fill the end.  wl_count[1] is
    // the next work index to take (words 2 and 3 of this step's counter set).
    int* wl_item;
    unsigned long long* wl_need;
    unsigned long long* wl_act;
    unsigned* wl_count;
    unsigned* wl_light;    // chunks listed from the end of the list (word 1 of the counter set)
    int items_chunk;       // full tiles per z chunk
    int edge_cols;         // interior columns outside the full tiles
    int edge_items;        // edge items per z chunk (edge_items_per_chunk)
    int step_blocks;       // persistent workgroups of the step kernel
    int fix_blocks;        // workgroups (one wave each) of the overflow fix-up kernel
    int fnx, fny, fnz;
    int L;                 // slot stride fnx*fny*fnz
    int ncolx, ncols;      // interior columns: (fnx-2), (fnx-2)*(fny-2)
    int nzchunk, kchunk;   // z-chunks per column, planes per chunk
    int tiles_x, tiles_y;  // 64x4 column tiles of the LDS path (partial ones at the +x/+y
                           // end included; see kPartialMin)
    int bcx, bcy, bcz;
    int num_theta;
    int thermal_type;
    int use_iuc;
    float t;
    float dt, hi, lamd, L_cp, kin_delta, eps;
    float G, R, uc, uc0, z0, r0, lyd, lzd, V, mp_len, x0, order;
    float tan_angle, cos_angle;
    int Nx, Ny, Nz;
    float X0_mac, Y0_mac, Z0_mac, Dx_mac;
    // The host's step kernel variant (not read by the kernels): nonzero for
    // the one with fix-up records, keep test and the edge kernel, used while
    // neighbourhoods overflow.  The fix-up reports for that choice: it writes
    // launch_index to seen[0] and, when its list holds kOvfHeavy points or
    // more, to seen[1] (pinned host memory; null: no report).
    unsigned long long* seen;
    unsigned long long launch_index;
    int ovf_variant;
};

// One simulation of a batch, in device memory.  `a` holds the arguments of
// its next step; after every step the advance kernel swaps the buffers, moves
// the epochs and the step index and recomputes t, exactly as the host does
// between two single-simulation launches.
struct BatchSim {
    StepArgs a;
    float* ph[2];
    int* arg[2];
    unsigned* act[2];
    unsigned ep[2];        // epochs of the two buffers
    unsigned ep_ctr;       // epoch counter of the simulation
    int cur;               // source buffer of the next step
    int step;              // index s of the next step (t = (float)s*dt*tau0)
    int parity;            // counter set of the next step (0..2, in turn)
    unsigned* cnt;         // the simulation's two 4-word counter sets
    float tau0;
    int plan_base;         // first plan wave of this simulation
    int edge_base;         // first edge work index of this simulation
};

// Plan waves: a wave's 64 lanes cover plan_tiles tiles, plan_rows lanes (z
// planes) each: a group of plan_group z chunks plus one halo plane at each
// end (kchunk <= 61).
__host__ __device__ inline int plan_rows(int kchunk) { return kchunk <= 14 ? 16 : (kchunk <= 30 ? 32 : 64); }
__host__ __device__ inline int plan_tiles(int kchunk) { return 64 / plan_rows(kchunk); }
__host__ __device__ inline int plan_group(int kchunk) { return (plan_rows(kchunk) - 2) / kchunk; }
// Plan waves of a simulation with nzchunk chunks of items_chunk tiles.
__host__ __device__ inline int plan_waves(int kchunk, int nzchunk, int items_chunk)
{
    const int g = plan_group(kchunk), tw = plan_tiles(kchunk);
    return ((nzchunk + g - 1) / g) * ((items_chunk + tw - 1) / tw);
}

// Points of one edge item: kEdgePerLane per lane of a 256-lane workgroup
// (4 per lane was measured 8 % slower on the 256x256x512 step: edge items are
// taken first, and fewer, longer ones hold workgroups the tile items need).
constexpr int kEdgePerLane = 1;
constexpr int kEdgePoints = 256 * kEdgePerLane;
// Rows of one edge tile (a one-column strip, 7-point gather): 64 rows x 4
// planes per pass of a 256-lane workgroup.
constexpr int kEdgeRows = 64;

// Edge items per z chunk for a strip of edge_cols = width x (fny - 2) points:
// kEdgePoints points each, or (tiles: batched or overflow steps, 7-point
// gather, one-column strip) edge tiles of kEdgeRows rows.
inline int edge_items_per_chunk(int edge_cols, int fny, int kchunk, bool gather27, bool tiles)
{
    if (edge_cols == 0) return 0;
    if (tiles && !gather27 && edge_cols == fny - 2) return (fny - 2 + kEdgeRows - 1) / kEdgeRows;
    return (int)(((long long)kchunk * edge_cols + kEdgePoints - 1) / kEdgePoints);
}

// Most simulations one batch may hold.
constexpr int kMaxBatch = 1024;

// A batched step over n simulations sharing K and the gather variant.  The
// work list holds tile chunks of every simulation; cnt holds two 4-word
// counter sets ([0] chunks listed from the front, [1] next work index, [2]
// chunks listed from the back) used by alternate steps, selected by cnt[8];
// cnt[9] counts the batched fix-up's launches.
struct BatchArgs {
    BatchSim* sims;
    int n;
    int plan_total;        // waves of the plan kernel (all simulations)
    int edge_total;        // edge work items of all simulations
    int* wl_sim;
    int* wl_item;
    unsigned long long* wl_need;
    unsigned long long* wl_act;
    unsigned* cnt;
    int wl_cap;            // list capacity: light chunks are listed down from wl_cap - 1
    int step_blocks;
    int fix_blocks;        // workgroups (one wave each) of the batched fix-up kernel
    // The host's batch_step variant: nonzero for the one with fix-up records
    // and keep test (7-point gather).  The batched fix-up reports for that
    // choice (pinned host memory; null: no report): it counts its launches in
    // cnt[9] and writes the count to seen[0] and, when the lists of all
    // members hold kOvfHeavy points or more, to seen[1].
    unsigned long long* seen;
    int ovf_variant;
};

// Nucleation launch block (3DgrainMPI/APTPhaseField.cu:459-536).
struct NuclArgs {
    float* ph;
    int* arg;
    int* status;
    const float* x;
    const float* y;
    const float* z;
    const float* T3d;
    const float* ball;     // (2p+1)^3 tanh template
    unsigned long long* events;
    unsigned* act;         // activity map of the buffer nucleated into, and its epoch
    unsigned ep;
    int ntx, ntiles;
    int fnx, fny, fnz, L, K;
    int cnx, cny, cnz, p;
    int use_iuc;
    float t, dt;
    float uc0, undcool_mean, undcool_std, nuc_Nmax, nuc_Nmax0, nuc_rad;
    int Nx, Ny, Nz, Nt;
    float X0_mac, Y0_mac, Z0_mac, t0_mac, Dx_mac, Dt_mac;
    uint64_t seed;
    uint64_t call;
};

// Launchers (apt_misc.hip, step_k*_g*.hip).  All enqueue on
// `stream` and return the hipError_t of the launch.
// Workgroups of the persistent step kernel that fit the device at once.
int step_kernel_slots(int K, int gather27, int device);
int batch_kernel_slots(int K, int gather27, int device);
// Workgroups of the (single or batched) overflow fix-up kernel that fit at once.
int fixup_kernel_slots(int K, int gather27, int device, bool batch);
// Batched step: plan, persistent step, fix-up and advance kernels.
hipError_t launch_batch_step(const BatchArgs& b, int K, int gather27, hipStream_t stream);
// The second stream of an overflow step's edge kernel and the events that
// order it: `ready` (recorded on the step's stream after the plan kernel),
// `done` (on the edge stream after the edge kernel; the fix-up waits for it).
struct EdgeLaunch {
    hipStream_t stream;
    hipEvent_t ready, done;
};

// Pieces shared by the per-K launchers.
hipError_t launch_plan(const StepArgs& a, hipStream_t stream);
hipError_t launch_batch_plan(const BatchArgs& b, hipStream_t stream);
hipError_t launch_batch_advance(const BatchArgs& b, hipStream_t stream);
hipError_t launch_step_k5_g7(const StepArgs& a, hipStream_t stream, const EdgeLaunch* edge,
                             hipEvent_t e0, hipEvent_t e1, hipEvent_t e2, hipEvent_t e3,
                             bool planned, const StepArgs* next);
hipError_t launch_batch_k5_g7(const BatchArgs& b, hipStream_t stream);
int step_slots_k5_g7(int device, int which);
hipError_t launch_step_k5_g27(const StepArgs& a, hipStream_t stream, const EdgeLaunch* edge,
                             hipEvent_t e0, hipEvent_t e1, hipEvent_t e2, hipEvent_t e3,
                             bool planned, const StepArgs* next);
hipError_t launch_batch_k5_g27(const BatchArgs& b, hipStream_t stream);
int step_slots_k5_g27(int device, int which);
hipError_t launch_step_k8_g7(const StepArgs& a, hipStream_t stream, const EdgeLaunch* edge,
                             hipEvent_t e0, hipEvent_t e1, hipEvent_t e2, hipEvent_t e3,
                             bool planned, const StepArgs* next);
hipError_t launch_batch_k8_g7(const BatchArgs& b, hipStream_t stream);
int step_slots_k8_g7(int device, int which);
hipError_t launch_step_k8_g27(const StepArgs& a, hipStream_t stream, const EdgeLaunch* edge,
                             hipEvent_t e0, hipEvent_t e1, hipEvent_t e2, hipEvent_t e3,
                             bool planned, const StepArgs* next);
hipError_t launch_batch_k8_g27(const BatchArgs& b, hipStream_t stream);
int step_slots_k8_g27(int device, int which);
// Work-list plan (unless `planned`: the previous launch's apt_fixplan made
// it), main step kernel, then the overflow fix-up, or with `next` the fix-up
// together with the plan of the step `next` describes (apt_fixplan).
// Optional events are recorded before the plan, after it, after the main
// kernel and after the fix-up.
hipError_t launch_step(const StepArgs& a, int K, int gather27, hipStream_t stream,
                       const EdgeLaunch* edge, hipEvent_t ev_begin = nullptr,
                       hipEvent_t ev_planned = nullptr, hipEvent_t ev_mid = nullptr,
                       hipEvent_t ev_end = nullptr, bool planned = false,
                       const StepArgs* next = nullptr);
// Device-side QoIs on a collected alpha field (qoi.hip).
hipError_t launch_qoi_volume(const int* alpha, int fnx, int fny, int fnz, int max_id, int* counts,
                             int* first, unsigned long long* above, hipStream_t stream);
// Junction rows in scan order: count pass into blk_cnt, scan into blk_off and
// *count, then (cap > 0) the rows; blk_cnt/blk_off hold qoi_junction_blocks words.
int qoi_junction_blocks(int fnx, int fny, int fnz);
hipError_t launch_qoi_junctions(const int* alpha, int fnx, int fny, int fnz, unsigned* blk_cnt,
                                unsigned* blk_off, int* rows, int cap, unsigned* count,
                                hipStream_t stream);
hipError_t launch_qoi_liquid_planes(const int* alpha, int fnx, int fny, int fnz, unsigned* liq,
                                   hipStream_t stream);
hipError_t launch_qoi_area(const int* alpha, int fnx, int fny, int fnz, int ng, int tip, int* total,
                           int* extra, hipStream_t stream);
hipError_t launch_bc_fill(float* ph, int* arg, int fnx, int fny, int fnz, int K, int bcx,
                          int bcy, int bcz, hipStream_t stream);
hipError_t launch_collect(const float* ph, const int* arg, float* phi, int* alpha, int L, int K,
                          int clear_alpha, hipStream_t stream);
hipError_t launch_ini(float* ph0, int* arg0, float* ph1, int* arg1, const float* phi0,
                      const int* alpha0, int L, int K, hipStream_t stream);
// Exact sums of planes k_from .. k_to-1 into sums[k_from .. k_to) (the rest untouched).
hipError_t launch_plane_sums(const float* ph, long long* sums, int fnx, int fny, int fnz, int K,
                             int k_from, int k_to, hipStream_t stream);
hipError_t launch_move_frame(float* ph_old, int* arg_old, float* ph_new, int* arg_new, float* z,
                             float* zbuf, const int* alpha, int* alpha_full, int* loss_area,
                             int move_count, int fnx, int fny, int fnz, int K, int num_theta,
                             hipStream_t stream);
hipError_t launch_nucl_status(const float* phi0, int* status, int cnx, int cny, int cnz, int p,
                              int fnx, int fny, hipStream_t stream);
hipError_t launch_nucl(const NuclArgs& a, hipStream_t stream);
// Activity map of one buffer from its contents: every tile plane holding a
// non-empty slot gets `ep`.
hipError_t launch_act_rebuild(const float* ph, const int* arg, unsigned* act, unsigned ep, int fnx,
                              int fny, int fnz, int K, int ntx, int ntiles, hipStream_t stream);
hipError_t launch_copy(const void* src, void* dst, size_t bytes, hipStream_t stream);

}  // namespace cupf
