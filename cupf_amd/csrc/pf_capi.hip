// pf_capi.hip -- C ABI of cupf_amd (declared in include/pf_capi.h).
//
// Owns the device state of one simulation: the two ping-pong K-slot buffers
// (phi float32 + grain id int32), coordinates, orientation and thermal tables,
// single-field phi/alpha, moving-domain archive and nucleation state.  All
// work goes to the context's own HIP stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <array>
#include <tuple>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pf_capi.h"
#include "apt_internal.h"
#include "capacity.h"

using namespace cupf;

struct pf_ctx {
    pf_params p;
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // the stream of the overflow steps' edge kernel and the events that fork
    // it off `stream` and join it back before the fix-up
    EdgeLaunch edge = {nullptr, nullptr, nullptr};
    // Marks every kLeadEvery launches (kLeadSlots of them in turn): before a
    // slot is reused, the host waits for the launch it marked, so it runs at
    // most kLeadEvery * kLeadSlots launches ahead of the device and the fix-up's
    // overflow report it chooses the step variant by is at most that old.
    static constexpr int kLeadEvery = 8, kLeadSlots = 4;
    hipEvent_t lead[kLeadSlots] = {};
    int L = 0;
    int cur = 0;                    // index of the current buffer
    float* ph[2] = {nullptr, nullptr};
    int* arg[2] = {nullptr, nullptr};
    float *x = nullptr, *y = nullptr, *z = nullptr, *zbuf = nullptr;
    float* R12 = nullptr;
    float* psi_mac = nullptr;
    float* T3d = nullptr;
    float* phi0 = nullptr;          // initial single field (phi_old of the reference)
    float* phi = nullptr;           // collect target
    int* alpha = nullptr;           // collect target / initial alpha (alpha_m)
    unsigned long long* counters = nullptr;  // [0] overflow points, [1] nucleation events
    unsigned* fix_buf = nullptr;    // record header (kFixHeader words), then the fix-up list
    unsigned* fix_list = nullptr;   // overflow fix-up list, one entry per interior point
    int* rec = nullptr;             // fix-up records: 7 x K ids, then 7 x K phis (structure of arrays)
    unsigned rec_cap = 0;           // records: min(interior points / 8, 2^20), none for 27-point
    unsigned* fix_count = nullptr;  // three 4-word counter sets, used by steps in turn
    int* sink = nullptr;            // 128 words written by lanes that must not store
    long long* sums = nullptr;
    std::vector<long long> h_sums;
    int* alpha_full = nullptr;
    int* loss = nullptr;
    int fnz_full = 0;
    int* nstatus = nullptr;
    float* ball = nullptr;
    int cnx = 0, cny = 0, cnz = 0;
    unsigned long long ncall = 0;
    long long wl_alloc = 0;         // entries allocated for wl_item / wl_need / wl_act
    long long fix_alloc = 0;        // entries allocated for fix_list
    int kchunk = 4;                 // z planes per work item
    // Activity map per buffer (one word per 64x4 column tile and z plane) and
    // the epoch each buffer was last written at; see StepArgs.
    unsigned* act[2] = {nullptr, nullptr};
    unsigned ep[2] = {1u, 1u};
    unsigned ep_ctr = 1u;
    int use_act = 1;
    int ntx = 0, nty = 0, ntiles = 0;
    // Work list of the step (filled by the plan kernel, cleared by the fix-up).
    int* wl_item = nullptr;
    unsigned long long* wl_need = nullptr;
    unsigned long long* wl_act = nullptr;
    unsigned long long nlaunch = 0;  // step launches so far (selects the counter set)
    // Step kernel variant: the fix-up's report (pinned host memory: launch of
    // the last fix-up, launch of the last one with a non-empty list), and the
    // launch up to which the overflow variant is used regardless (a new state).
    unsigned long long* seen = nullptr;
    unsigned long long ovf_until = 0;
    int variant_force = -1;         // CUPF_STEP_VARIANT: 0 / 1 fixes the variant (tests, timing)
    int nzchunk = 0, edge_cols = 0, edge_items = 0, items_chunk = 0;
    int tiles_x = 0, tiles_y = 0;   // LDS tiles (partial ones included)
    int step_blocks = 0;            // persistent step workgroups (occupancy x CUs)
    int fix_blocks = 0;             // fix-up workgroups (occupancy x CUs)
    // overflow on the LDS path by tile plane (see StepArgs)
    bool have_coords = false, have_orient = false, have_state = false;
};

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(PF_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));      \
    } while (0)

template <class T>
int dalloc(T** p, size_t n)
{
    *p = nullptr;
    if (n == 0) return PF_OK;
    hipError_t e = hipMalloc((void**)p, n * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(PF_E_NOMEM, "hipMalloc of " + std::to_string(n * sizeof(T)) +
                                    " bytes failed: " + hipGetErrorString(e));
    }
    return PF_OK;
}

// Device memory of pf_qoi_volume_noalloc: counts[PF_NOALLOC_MAX_ID + 1],
// first[PF_NOALLOC_MAX_ID + 1], then the 64-bit count of larger ids.  Host code
// never uses the name `scratch` as an address (it is a host-side stand-in):
// pf_create resolves the device address once per GPU with
// hipGetSymbolAddress and keeps it in g_scratch.
constexpr size_t kScratchIds = (size_t)PF_NOALLOC_MAX_ID + 1;
static __device__ __align__(256) unsigned char scratch[2 * kScratchIds * sizeof(int) + 256];

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch[kMaxScratchDevices] = {};

// Resolve this GPU's copy of `scratch` (the device must be current).
int scratch_setup(int device)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0 || device >= kMaxScratchDevices)
        return fail(PF_E_INVALID, "device index beyond the scratch table");
    if (g_scratch[device]) return PF_OK;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return fail(PF_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[device] = ptr;
    return PF_OK;
}

void* scratch_of(int device)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return (device >= 0 && device < kMaxScratchDevices) ? g_scratch[device] : nullptr;
}

int ctx_check(pf_ctx* c)
{
    if (!c) return fail(PF_E_INVALID, "null context");
    if (hipSetDevice(c->device) != hipSuccess) return fail(PF_E_HIP, "hipSetDevice failed");
    return PF_OK;
}

int buf_index(pf_ctx* c, int which) { return which == 0 ? c->cur : 1 - c->cur; }

// Launches without a reported heavy overflow (a fix-up list of kOvfHeavy
// points or more) after which the step kernel without fix-up records and keep
// test, its edge items inside it, is used (7-point gather).  The report lags
// the launches by the queue's depth (at most kLeadEvery * kLeadSlots); either
// variant gives the same bits.  The lag and the quiet period are kept short:
// the overflow variant's edge kernel costs 15-35 us per launch once little
// overflows.
constexpr unsigned long long kOvfQuiet = 16;
// Launches after a new state that take the overflow variant whatever the
// reports say (those in flight are still the old state's).
constexpr unsigned long long kOvfNew =
    (unsigned long long)(pf_ctx::kLeadEvery * pf_ctx::kLeadSlots) + kOvfQuiet;

// Before launch c->nlaunch: every kLeadEvery launches, wait for the launch
// marked kLeadEvery * kLeadSlots launches ago, then mark this one.
hipError_t bound_lead(pf_ctx* c)
{
    if (c->nlaunch % pf_ctx::kLeadEvery != 0) return hipSuccess;
    hipEvent_t& e = c->lead[(c->nlaunch / pf_ctx::kLeadEvery) % pf_ctx::kLeadSlots];
    hipError_t r = hipEventSynchronize(e);  // an event never recorded completes at once
    if (r != hipSuccess) return r;
    return hipEventRecord(e, c->stream);
}

int ovf_variant(const pf_ctx* c)
{
    if (c->p.gather27) return 0;
    if (c->variant_force >= 0) return c->variant_force;
    if (c->nlaunch < c->ovf_until || !c->seen) return 1;
    const volatile unsigned long long* r = c->seen;
    const unsigned long long last = r[1], now = r[0];
    if (now == 0) return 1;  // nothing reported yet
    return now < last + kOvfQuiet ? 1 : 0;
}

StepArgs make_step_args(pf_ctx* c, int src, float t, int use_iuc)
{
    const pf_params& p = c->p;
    StepArgs a;
    a.ph = c->ph[src];
    a.arg = c->arg[src];
    a.ph_new = c->ph[1 - src];
    a.arg_new = c->arg[1 - src];
    a.x = c->x;
    a.y = c->y;
    a.z = c->z;
    a.R12 = c->R12;
    a.psi_mac = c->psi_mac;
    a.overflow = c->counters;
    a.fix_list = c->fix_list;
    {
        // three sets in turn (advanced once the launch succeeded): this step's
        // plan clears the next step's set, whose last user (two steps back) has
        // finished; the next step's plan may run beside this step's fix-up
        // (apt_fixplan), so it must not clear this step's set
        const int set = (int)(c->nlaunch % 3ull);
        a.fix_count = c->fix_count + 4 * set;
        a.cnt_next = c->fix_count + 4 * ((set + 1) % 3);
        a.wl_count = a.fix_count + 2;
        a.wl_light = a.fix_count + 1;
    }
    a.sink_i = c->sink;
    a.sink_f = reinterpret_cast<float*>(c->sink + 64);
    a.fnx = p.fnx;
    a.fny = p.fny;
    a.fnz = p.fnz;
    a.L = c->L;
    a.ncolx = p.fnx - 2;
    a.ncols = (p.fnx - 2) * (p.fny - 2);
    a.kchunk = c->kchunk;
    a.nzchunk = c->nzchunk;
    a.tiles_x = c->tiles_x;
    a.tiles_y = c->tiles_y;
    a.act_src = c->act[src];
    a.act_dst = c->act[1 - src];
    a.ep_src = c->ep[src];
    a.ep_dst = c->ep[1 - src];
    a.ep_new = ++c->ep_ctr;
    a.use_act = c->use_act;
    a.ntx = c->ntx;
    a.nty = c->nty;
    a.ntiles = c->ntiles;
    a.edge_cols = c->edge_cols;
    a.ovf_variant = ovf_variant(c);
    a.seen = c->seen;
    a.launch_index = c->nlaunch + 1;
    a.edge_items = edge_items_per_chunk(c->edge_cols, p.fny, c->kchunk, p.gather27 != 0,
                                        a.ovf_variant != 0);
    a.items_chunk = c->items_chunk;
    a.step_blocks = c->step_blocks;
    a.fix_blocks = c->fix_blocks;
    a.wl_item = c->wl_item;
    a.wl_need = c->wl_need;
    a.wl_act = c->wl_act;

    a.bcx = p.bcx;
    a.bcy = p.bcy;
    a.bcz = p.bcz;
    a.num_theta = p.num_theta;
    a.thermal_type = p.thermal_type;
    a.use_iuc = use_iuc;
    a.t = t;
    a.dt = p.dt;
    a.hi = p.hi;
    a.lamd = p.lamd;
    a.L_cp = p.L_cp;
    a.kin_delta = p.kin_delta;
    a.eps = p.eps;
    a.G = p.G;
    a.R = p.R;
    a.uc = p.uc;
    a.uc0 = p.uc0;
    a.z0 = p.z0;
    a.r0 = p.r0;
    a.lyd = p.lyd;
    a.lzd = p.lzd;
    a.V = p.V;
    a.mp_len = p.mp_len;
    a.x0 = p.x0;
    a.order = p.order;
    a.tan_angle = p.tan_angle;
    a.cos_angle = p.cos_angle;
    a.Nx = p.Nx;
    a.Ny = p.Ny;
    a.Nz = p.Nz;
    a.X0_mac = p.X0_mac;
    a.Y0_mac = p.Y0_mac;
    a.Z0_mac = p.Z0_mac;
    a.Dx_mac = p.Dx_mac;
    return a;
}

int nucl_launch(pf_ctx* c, float t, float dt_nuc, int use_iuc)
{
    if (!c->nstatus) return fail(PF_E_STATE, "nucleation requested before pf_nucl_reset");
    const pf_params& p = c->p;
    NuclArgs a;
    a.ph = c->ph[c->cur];
    a.arg = c->arg[c->cur];
    a.status = c->nstatus;
    a.x = c->x;
    a.y = c->y;
    a.z = c->z;
    a.T3d = c->T3d;
    a.ball = c->ball;
    a.events = c->counters + 1;
    a.act = c->act[c->cur];
    a.ep = c->ep[c->cur];
    a.ntx = c->ntx;
    a.ntiles = c->ntiles;
    a.fnx = p.fnx;
    a.fny = p.fny;
    a.fnz = p.fnz;
    a.L = c->L;
    a.K = p.K;
    a.cnx = c->cnx;
    a.cny = c->cny;
    a.cnz = c->cnz;
    a.p = p.pts_cell;
    a.use_iuc = use_iuc;
    a.t = t;
    a.dt = dt_nuc;
    a.uc0 = p.uc0;
    a.undcool_mean = p.undcool_mean;
    a.undcool_std = p.undcool_std;
    a.nuc_Nmax = p.nuc_Nmax;
    a.nuc_Nmax0 = p.nuc_Nmax0;
    a.nuc_rad = p.nuc_rad;
    a.Nx = p.Nx;
    a.Ny = p.Ny;
    a.Nz = p.Nz;
    a.Nt = p.Nt;
    a.X0_mac = p.X0_mac;
    a.Y0_mac = p.Y0_mac;
    a.Z0_mac = p.Z0_mac;
    a.t0_mac = p.t0_mac;
    a.Dx_mac = p.Dx_mac;
    a.Dt_mac = p.Dt_mac;
    a.seed = p.seed;
    a.call = c->ncall++;
    if (!use_iuc && !c->T3d)
        return fail(PF_E_STATE, "nucleation needs the T3d macro field (pf_set_thermal_macro)");
    HIPCHK(launch_nucl(a, c->stream));
    return PF_OK;
}

// Activity map of buffer b from its contents, under a fresh epoch.
hipError_t act_rebuild(pf_ctx* c, int b)
{
    const pf_params& p = c->p;
    c->ep[b] = ++c->ep_ctr;
    return launch_act_rebuild(c->ph[b], c->arg[b], c->act[b], c->ep[b], p.fnx, p.fny, p.fnz, p.K,
                              c->ntx, c->ntiles, c->stream);
}

// Before the epoch counter could wrap within the next `ahead` steps: clear
// both maps and rebuild them.
hipError_t act_renew_if_needed(pf_ctx* c, unsigned ahead = 1u)
{
    if (c->ep_ctr < 0x7fffffffu - ahead) return hipSuccess;
    const size_t n = (size_t)c->ntiles * c->p.fnz;
    hipError_t e;
    for (int b = 0; b < 2; ++b)
        if ((e = hipMemsetAsync(c->act[b], 0, n * sizeof(unsigned), c->stream)) != hipSuccess) return e;
    c->ep_ctr = 1u;
    if ((e = act_rebuild(c, 0)) != hipSuccess) return e;
    return act_rebuild(c, 1);
}

void free_ctx(pf_ctx* c)
{
    (void)hipSetDevice(c->device);
    // a step that failed between its fork and join leaves the edge stream unjoined
    if (c->edge.stream) (void)hipStreamSynchronize(c->edge.stream);
    void* bufs[] = {c->ph[0], c->ph[1], c->arg[0], c->arg[1], c->x, c->y, c->z, c->zbuf,
                    c->R12, c->psi_mac, c->T3d, c->phi0, c->phi, c->alpha, c->counters,
                    c->sums, c->alpha_full, c->loss, c->nstatus, c->ball,
                    c->fix_buf, c->fix_count, c->sink, c->act[0], c->act[1],
                    c->wl_item, c->wl_need, c->wl_act, c->rec};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    if (c->seen) (void)hipHostFree(c->seen);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    for (hipEvent_t e : c->lead)
        if (e) (void)hipEventDestroy(e);
    if (c->edge.ready) (void)hipEventDestroy(c->edge.ready);
    if (c->edge.done) (void)hipEventDestroy(c->edge.done);
    if (c->edge.stream) (void)hipStreamDestroy(c->edge.stream);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

}  // namespace

// A batch: several contexts of one device stepped together (see pf_batch_step).
struct pf_batch {
    std::vector<pf_ctx*> ctxs;
    int device = 0, K = 0, g27 = 0;
    hipStream_t stream = nullptr;
    int kchunk = 4;
    int step_blocks = 0;
    int fix_blocks = 0;
    BatchSim* sims = nullptr;       // device array, n entries
    int* wl_sim = nullptr;
    int* wl_item = nullptr;
    unsigned long long* wl_need = nullptr;
    unsigned long long* wl_act = nullptr;
    unsigned* cnt = nullptr;        // 2 x 4 list counters, parity word, fix-up count
    size_t wl_cap = 0;
    // hipGraphs of kGraphSteps batched steps, one per step kernel variant,
    // each captured for one kchunk
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    int graph_kchunk[2] = {0, 0};
    static constexpr int kGraphSteps = 8;
    static constexpr int kMinChunk = 2;
    // Step kernel variant choice, as for a single context: batch_fixup
    // reports (pinned) the index within the current run of its step and of
    // the last step whose lists held kOvfHeavy points or more in all.
    unsigned long long* seen = nullptr;
    unsigned long long nstep = 0;     // batched steps enqueued over the batch's life
    unsigned long long seen_base = 0; // nstep when the current run began
    unsigned long long heavy_at = 0;  // last step known to overflow heavily (1-based)
    int variant_force = -1;           // CUPF_STEP_VARIANT
    // lead bound: a mark every kGraphSteps steps, kLeadSlots of them in turn
    static constexpr int kLeadSlots = 4;
    hipEvent_t lead[kLeadSlots] = {};
    unsigned long long lead_n = 0, next_mark = 0;
};

namespace {

void batch_free(pf_batch* b)
{
    (void)hipSetDevice(b->device);
    for (int v = 0; v < 2; ++v) {
        if (b->gexec[v]) (void)hipGraphExecDestroy(b->gexec[v]);
        if (b->graph[v]) (void)hipGraphDestroy(b->graph[v]);
    }
    void* bufs[] = {b->sims, b->wl_sim, b->wl_item, b->wl_need, b->wl_act, b->cnt};
    for (void* p : bufs)
        if (p) (void)hipFree(p);
    if (b->seen) (void)hipHostFree(b->seen);
    for (hipEvent_t e : b->lead)
        if (e) (void)hipEventDestroy(e);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

// Step arguments of every member for steps first_step[i].., with the batch's
// chunk depth, as device BatchSim records (and the work-index bases).
int batch_upload(pf_batch* b, const std::vector<int>& first_step, int use_iuc, BatchArgs& ba)
{
    const int n = (int)b->ctxs.size();
    std::vector<BatchSim> h(n);
    int plan_total = 0, edge_total = 0;
    for (int i = 0; i < n; ++i) {
        pf_ctx* c = b->ctxs[i];
        const pf_params& p = c->p;
        const float t = first_step[i] * p.dt * p.tau0;
        BatchSim& m = h[i];
        m.a = make_step_args(c, c->cur, t, use_iuc);  // takes this step's epoch from the context
        m.a.kchunk = b->kchunk;
        m.a.nzchunk = (p.fnz - 2 + b->kchunk - 1) / b->kchunk;
        m.a.edge_items = edge_items_per_chunk(c->edge_cols, p.fny, b->kchunk, p.gather27 != 0, true);
        m.a.step_blocks = b->step_blocks;
        m.a.fix_blocks = b->fix_blocks;
        for (int q = 0; q < 2; ++q) {
            m.ph[q] = c->ph[q];
            m.arg[q] = c->arg[q];
            m.act[q] = c->act[q];
            m.ep[q] = c->ep[q];
        }
        m.ep_ctr = c->ep_ctr;
        m.cur = c->cur;
        m.step = first_step[i];
        m.parity = (int)(c->nlaunch % 3ull);
        m.cnt = c->fix_count;
        m.tau0 = p.tau0;
        m.plan_base = plan_total;
        m.edge_base = edge_total;
        plan_total += plan_waves(b->kchunk, m.a.nzchunk, m.a.items_chunk);
        edge_total += m.a.nzchunk * m.a.edge_items;
    }
    HIPCHK(hipMemcpyAsync(b->sims, h.data(), sizeof(BatchSim) * n, hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipMemsetAsync(b->cnt, 0, 10 * sizeof(unsigned), b->stream));
    ba.sims = b->sims;
    ba.n = n;
    ba.plan_total = plan_total;
    ba.edge_total = edge_total;
    ba.wl_sim = b->wl_sim;
    ba.wl_item = b->wl_item;
    ba.wl_need = b->wl_need;
    ba.wl_act = b->wl_act;
    ba.wl_cap = (int)b->wl_cap;
    ba.cnt = b->cnt;
    ba.step_blocks = b->step_blocks;
    ba.fix_blocks = b->fix_blocks;
    ba.seen = b->seen;
    ba.ovf_variant = 0;
    return PF_OK;
}

// The step kernel variant of the next batched step (1: fix-up records and
// keep test), from the fix-up's reports as for a single context.
int batch_variant(pf_batch* b)
{
    if (b->g27) return 0;
    if (b->variant_force >= 0) return b->variant_force;
    const volatile unsigned long long* r = b->seen;
    const unsigned long long now = r[0], last = r[1];
    if (last) b->heavy_at = std::max(b->heavy_at, b->seen_base + last);
    return b->seen_base + now < b->heavy_at + kOvfQuiet ? 1 : 0;
}

// Before the steps from b->nstep on: every kGraphSteps steps, wait for the
// mark set kLeadSlots marks before, then mark the stream, so the host runs at
// most ~kLeadSlots * kGraphSteps steps ahead of the reports it chooses by.
hipError_t batch_bound_lead(pf_batch* b)
{
    if (b->nstep < b->next_mark) return hipSuccess;
    hipEvent_t& e = b->lead[b->lead_n % pf_batch::kLeadSlots];
    hipError_t r = hipEventSynchronize(e);  // an event never recorded completes at once
    if (r != hipSuccess) return r;
    ++b->lead_n;
    b->next_mark = b->nstep + pf_batch::kGraphSteps;
    return hipEventRecord(e, b->stream);
}

// The host side of nsteps batched steps (what step_impl does per launch),
// after make_step_args already took the first step's epoch.
void batch_mirror(pf_batch* b, int nsteps)
{
    for (pf_ctx* c : b->ctxs) {
        for (int s = 0; s < nsteps; ++s) {
            if (s > 0) ++c->ep_ctr;
            c->ep[1 - c->cur] = c->ep_ctr;
            c->cur = 1 - c->cur;
            ++c->nlaunch;
        }
    }
}

// nsteps batched steps of every member from first_step (advanced on return).
int batch_run(pf_batch* b, int nsteps, std::vector<int>& first_step, int use_iuc, bool use_graph)
{
    if (nsteps <= 0) return PF_OK;
    for (pf_ctx* c : b->ctxs) {
        HIPCHK(hipSetDevice(c->device));
        HIPCHK(act_renew_if_needed(c, (unsigned)nsteps + 2u));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    BatchArgs ba;
    int rc = batch_upload(b, first_step, use_iuc, ba);
    if (rc) return rc;
    // the previous run ended with a synchronisation: fold its last reports in,
    // then count this run's steps from 1 (batch_upload cleared the device count)
    (void)batch_variant(b);
    b->seen[0] = b->seen[1] = 0ull;
    b->seen_base = b->nstep;
    int left = nsteps;
    while (left > 0) {
        HIPCHK(batch_bound_lead(b));
        const int v = batch_variant(b);
        ba.ovf_variant = v;
        if (use_graph && left >= pf_batch::kGraphSteps) {
            if (!b->gexec[v] || b->graph_kchunk[v] != b->kchunk) {
                if (b->gexec[v]) (void)hipGraphExecDestroy(b->gexec[v]);
                if (b->graph[v]) (void)hipGraphDestroy(b->graph[v]);
                b->gexec[v] = nullptr;
                b->graph[v] = nullptr;
                HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
                hipError_t e = hipSuccess;
                for (int q = 0; q < pf_batch::kGraphSteps && e == hipSuccess; ++q)
                    e = launch_batch_step(ba, b->K, b->g27, b->stream);
                hipGraph_t g = nullptr;
                const hipError_t e2 = hipStreamEndCapture(b->stream, &g);
                if (e != hipSuccess || e2 != hipSuccess) {
                    if (g) (void)hipGraphDestroy(g);
                    return fail(PF_E_HIP, std::string("graph capture: ") +
                                              hipGetErrorString(e != hipSuccess ? e : e2));
                }
                b->graph[v] = g;
                HIPCHK(hipGraphInstantiate(&b->gexec[v], b->graph[v], nullptr, nullptr, 0));
                b->graph_kchunk[v] = b->kchunk;
            }
            HIPCHK(hipGraphLaunch(b->gexec[v], b->stream));
            left -= pf_batch::kGraphSteps;
            b->nstep += pf_batch::kGraphSteps;
        } else {
            HIPCHK(launch_batch_step(ba, b->K, b->g27, b->stream));
            --left;
            ++b->nstep;
        }
    }
    HIPCHK(hipStreamSynchronize(b->stream));
    batch_mirror(b, nsteps);
    for (int& s : first_step) s += nsteps;
    return PF_OK;
}

}  // namespace

extern "C" {

const char* pf_last_error(void) { return g_err.c_str(); }

const char* pf_build_info(void) { return "cupf_amd gfx950 K={5,8} gather={7,27}"; }

int pf_create(const pf_params* params, int device, pf_ctx** out)
{
    if (!params || !out) return fail(PF_E_INVALID, "null argument");
    *out = nullptr;
    const pf_params& p = *params;
    if (p.K != 5 && p.K != 8) return fail(PF_E_INVALID, "K must be 5 or 8");
    if (p.fnx < 4 || p.fny < 4 || p.fnz < 4)
        return fail(PF_E_INVALID, "grid too small: fnx, fny, fnz must be >= 4");
    if (p.num_theta < 1) return fail(PF_E_INVALID, "num_theta must be >= 1");
    if (p.bcx < 0 || p.bcx > 1 || p.bcy < 0 || p.bcy > 1 || p.bcz < 0 || p.bcz > 1)
        return fail(PF_E_INVALID, "bc digits must be 0 or 1");
    if (p.thermal_type < 0 || p.thermal_type > 5)
        return fail(PF_E_INVALID, "thermal_type must be 0..5");
    const long long L = (long long)p.fnx * p.fny * p.fnz;
    if (L * p.K >= (1LL << 31))
        return fail(PF_E_INVALID, "K*fnx*fny*fnz must stay below 2^31 (32-bit slot indexing)");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(PF_E_HIP, "no HIP device available");
    if (device < 0 || device >= ndev) return fail(PF_E_INVALID, "device index out of range");
    HIPCHK(hipSetDevice(device));
    if (int rs = scratch_setup(device)) return rs;

    pf_ctx* c = new pf_ctx();
    c->p = p;
    if (const char* kc = getenv("CUPF_KCHUNK")) {
        const int v = atoi(kc);
        if (v >= 1 && v <= 61) c->kchunk = v;  // plane masks of a chunk fit 64 bits
    }
    if (const char* na = getenv("CUPF_NO_ACTIVITY")) c->use_act = atoi(na) ? 0 : 1;
    if (const char* sv = getenv("CUPF_STEP_VARIANT")) c->variant_force = atoi(sv) ? 1 : 0;
    c->ntx = (p.fnx - 2 + 63) / 64;
    c->nty = (p.fny - 2 + 3) / 4;
    c->ntiles = c->ntx * c->nty;
    {
        // 64x4 tiles take the LDS path, a partial tile where the +x leftover is
        // at least kPartialMin wide and always in y; a narrower +x strip goes to
        // edge items of 256 points each
        int tiles_x = (p.fnx - 2) / 64;
        int ex = (p.fnx - 2) - tiles_x * 64;
        if (ex >= kPartialMin) {
            ++tiles_x;
            ex = 0;
        }
        c->tiles_x = tiles_x;
        c->tiles_y = (p.fny - 2 + 3) / 4;
        c->edge_cols = ex * (p.fny - 2);
        c->edge_items = edge_items_per_chunk(c->edge_cols, p.fny, c->kchunk, p.gather27 != 0, false);
        c->items_chunk = c->tiles_x * c->tiles_y;
        c->nzchunk = (p.fnz - 2 + c->kchunk - 1) / c->kchunk;
    }
    c->step_blocks = step_kernel_slots(p.K, p.gather27, device);
    c->fix_blocks = fixup_kernel_slots(p.K, p.gather27, device, false);
    if (c->step_blocks < 1 || c->fix_blocks < 1) {
        delete c;
        return fail(PF_E_HIP, "occupancy query for the step or fix-up kernel failed");
    }
    // fewer fix-up workgroups (each lane then redoes more overflowed points in
    // turn); the results do not depend on it
    if (const char* fb = getenv("CUPF_FIX_BLOCKS")) {
        const int v = atoi(fb);
        if (v >= 1 && v < c->fix_blocks) c->fix_blocks = v;
    }
    c->device = device;
    c->L = (int)L;
    int rc = PF_OK;
    auto bail = [&](int code) {
        free_ctx(c);
        return code;
    };
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(PF_E_HIP, "hipStreamCreate failed"));
    if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess)
        return bail(fail(PF_E_HIP, "hipEventCreate failed"));
    for (hipEvent_t& e : c->lead)
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
            return bail(fail(PF_E_HIP, "hipEventCreate failed"));
    if (c->edge_items > 0 && !p.gather27 &&
        (hipStreamCreateWithFlags(&c->edge.stream, hipStreamNonBlocking) != hipSuccess ||
         hipEventCreateWithFlags(&c->edge.ready, hipEventDisableTiming) != hipSuccess ||
         hipEventCreateWithFlags(&c->edge.done, hipEventDisableTiming) != hipSuccess))
        return bail(fail(PF_E_HIP, "edge stream or event creation failed"));
    const size_t nst = (size_t)L * p.K;
    for (int b = 0; b < 2; ++b) {
        if ((rc = dalloc(&c->ph[b], nst)) || (rc = dalloc(&c->arg[b], nst))) return bail(rc);
    }
    if ((rc = dalloc(&c->x, p.fnx)) || (rc = dalloc(&c->y, p.fny)) || (rc = dalloc(&c->z, p.fnz)) ||
        (rc = dalloc(&c->zbuf, p.fnz)) || (rc = dalloc(&c->phi0, (size_t)L)) ||
        (rc = dalloc(&c->phi, (size_t)L)) || (rc = dalloc(&c->alpha, (size_t)L)) ||
        (rc = dalloc(&c->counters, 2)) || (rc = dalloc(&c->sums, p.fnz)) ||
        (rc = dalloc(&c->fix_buf, kFixHeader + (size_t)(c->fix_alloc = (long long)(p.fnx - 2) * (p.fny - 2) * (p.fnz - 2)))) ||
        (rc = dalloc(&c->fix_count, 12)) || (rc = dalloc(&c->sink, 128)) ||
        (rc = dalloc(&c->act[0], (size_t)c->ntiles * p.fnz)) ||
        (rc = dalloc(&c->act[1], (size_t)c->ntiles * p.fnz)) ||
        (rc = dalloc(&c->wl_item, (size_t)(c->wl_alloc = (long long)c->nzchunk * c->items_chunk + 1))) ||
        (rc = dalloc(&c->wl_need, (size_t)c->nzchunk * c->items_chunk + 1)) ||
        (rc = dalloc(&c->wl_act, (size_t)c->nzchunk * c->items_chunk + 1)))
        return bail(rc);
    if (!p.gather27) {
        // records of the overflowed points the tile hands to the fix-up (the
        // first positions of the list; later ones are gathered from the state)
        const long long interior = (long long)(p.fnx - 2) * (p.fny - 2) * (p.fnz - 2);
        c->rec_cap = (unsigned)std::max(1LL, std::min(interior / 8, 1LL << 20));
        if ((rc = dalloc(&c->rec, (size_t)c->rec_cap * 7 * p.K * 2))) return bail(rc);
    }
    c->fix_list = c->fix_buf + kFixHeader;
    unsigned hdr[4] = {c->rec_cap, 0u, 0u, 0u};
    std::memcpy(&hdr[2], &c->rec, sizeof(int*));
    static_assert(sizeof(int*) == 8, "record header holds a 64-bit address");
    c->h_sums.resize(p.fnz);
    for (int b = 0; b < 2; ++b) {
        if (hipMemsetD32Async((hipDeviceptr_t)c->ph[b], 0xBF800000u, nst, c->stream) != hipSuccess ||
            hipMemsetAsync(c->arg[b], 0xFF, nst * sizeof(int), c->stream) != hipSuccess)
            return bail(fail(PF_E_HIP, "initial fill failed"));
    }
    if (hipMemsetAsync(c->counters, 0, 2 * sizeof(unsigned long long), c->stream) != hipSuccess ||
        hipMemsetAsync(c->fix_count, 0, 12 * sizeof(unsigned), c->stream) != hipSuccess ||
        hipMemcpyAsync(c->fix_list - 4, hdr, sizeof(hdr), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipMemsetAsync(c->alpha, 0, L * sizeof(int), c->stream) != hipSuccess ||
        hipMemsetAsync(c->act[0], 0, (size_t)c->ntiles * p.fnz * sizeof(unsigned), c->stream) != hipSuccess ||
        hipMemsetAsync(c->act[1], 0, (size_t)c->ntiles * p.fnz * sizeof(unsigned), c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return bail(fail(PF_E_HIP, "initial fill failed"));
    if (hipHostMalloc(&c->seen, 2 * sizeof(unsigned long long)) != hipSuccess) {
        c->seen = nullptr;
        return bail(fail(PF_E_NOMEM, "hipHostMalloc of the step variant report failed"));
    }
    c->seen[0] = c->seen[1] = 0ull;
    c->ovf_until = kOvfNew;
    *out = c;
    return PF_OK;
}

int pf_destroy(pf_ctx* ctx)
{
    if (!ctx) return PF_OK;
    (void)hipStreamSynchronize(ctx->stream);
    free_ctx(ctx);
    return PF_OK;
}

int pf_set_coords(pf_ctx* c, const float* x, const float* y, const float* z)
{
    int rc = ctx_check(c);
    if (rc) return rc;
    if (!x || !y || !z) return fail(PF_E_INVALID, "null coordinate array");
    HIPCHK(hipMemcpyAsync(c->x, x, sizeof(float) * c->p.fnx, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->y, y, sizeof(float) * c->p.fny, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->z, z, sizeof(float) * c->p.fnz, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->have_coords = true;
    return PF_OK;
}

int pf_orientations_from_angles(const float* theta, int ng, float* R9)
{
    if (!theta || !R9 || ng < 1) return fail(PF_E_INVALID, "bad arguments");
    for (int g = 0; g <= ng; ++g) {
        // sint/cost tables of 3DgrainMPI/PhaseFieldCpu.cpp:379-385, indexed as
        // in 3DgrainMPI/APTPhaseField.cu:262-265 (row 0 pairs theta[0] with theta[NG]).
        const float sina = sinf(theta[g]), cosa = cosf(theta[g]);
        const float sinb = sinf(theta[g + ng]), cosb = cosf(theta[g + ng]);
        float* m = R9 + 9 * (size_t)g;
        m[0] = cosa * cosb;
        m[1] = sina * cosb;
        m[2] = -sinb;
        m[3] = -sina;
        m[4] = cosa;
        m[5] = 0.0f;
        m[6] = cosa * sinb;
        m[7] = sina * sinb;
        m[8] = cosb;
    }
    return PF_OK;
}

int pf_set_orientations(pf_ctx* c, const float* R9, int nrows)
{
    int rc = ctx_check(c);
    if (rc) return rc;
    if (!R9) return fail(PF_E_INVALID, "null orientation table");
    if (nrows < c->p.num_theta + 1)
        return fail(PF_E_INVALID, "orientation table needs num_theta+1 rows");
    std::vector<float> r12((size_t)nrows * 12, 0.0f);
    for (int g = 0; g < nrows; ++g)
        for (int q = 0; q < 9; ++q) r12[12 * (size_t)g + q] = R9[9 * (size_t)g + q];
    if (c->R12) (void)hipFree(c->R12);
    c->R12 = nullptr;
    if ((rc = dalloc(&c->R12, r12.size()))) return rc;
    HIPCHK(hipMemcpyAsync(c->R12, r12.data(), sizeof(float) * r12.size(), hipMemcpyHostToDevice,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->have_orient = true;
    return PF_OK;
}

int pf_set_thermal_macro(pf_ctx* c, const float* psi_mac, const float* T3d)
{
    int rc = ctx_check(c);
    if (rc) return rc;
    const pf_params& p = c->p;
    const size_t n3 = (size_t)p.Nx * p.Ny * p.Nz;
    if ((psi_mac || T3d) && (p.Nx < 2 || p.Ny < 2 || p.Nz < 2))
        return fail(PF_E_INVALID, "macro grid needs Nx, Ny, Nz >= 2");
    if (psi_mac) {
        if (c->psi_mac) (void)hipFree(c->psi_mac);
        c->psi_mac = nullptr;
        if ((rc = dalloc(&c->psi_mac, n3))) return rc;
        HIPCHK(hipMemcpyAsync(c->psi_mac, psi_mac, sizeof(float) * n3, hipMemcpyHostToDevice,
                              c->stream));
    }
    if (T3d) {
        if (p.Nt < 2) return fail(PF_E_INVALID, "T3d needs Nt >= 2");
        if (c->T3d) (void)hipFree(c->T3d);
        c->T3d = nullptr;
        if ((rc = dalloc(&c->T3d, n3 * p.Nt))) return rc;
        HIPCHK(hipMemcpyAsync(c->T3d, T3d, sizeof(float) * n3 * p.Nt, hipMemcpyHostToDevice,
                              c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return PF_OK;
}

int pf_init_from_phi_alpha(pf_ctx* c, const float* phi0, const int* alpha0)
{
    int rc = ctx_check(c);
    if (rc) return rc;
    if (!phi0 || !alpha0) return fail(PF_E_INVALID, "null initial field");
    const size_t L = c->L;
    HIPCHK(hipMemcpyAsync(c->phi0, phi0, sizeof(float) * L, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->phi, phi0, sizeof(float) * L, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->alpha, alpha0, sizeof(int) * L, hipMemcpyHostToDevice, c->stream));
    c->cur = 0;
    HIPCHK(launch_ini(c->ph[0], c->arg[0], c->ph[1], c->arg[1], c->phi0, c->alpha, c->L, c->p.K,
                      c->stream));
    HIPCHK(act_rebuild(c, 0));
    HIPCHK(act_rebuild(c, 1));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->have_state = true;
    c->ovf_until = c->nlaunch + kOvfNew;  // a new state may overflow
    return PF_OK;
}

int pf_reinit_from_collect(pf_ctx* c)
{
    int rc = ctx_check(c);
    if (rc) return rc;
    if (!c->have_state) return fail(PF_E_STATE, "no initial field");
    c->cur = 0;
    HIPCHK(launch_ini(c->ph[0], c->arg[0], c->ph[1], c->arg[1], c->phi0, c->alpha, c->L, c->p.K,
                      c->stream));
    HIPCHK(act_rebuild(c, 0));
    HIPCHK(act_rebuild(c, 1));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->ovf_until = c->nlaunch + kOvfNew;  // a new state may overflow
    return PF_OK;
}

int pf_upload_state(pf_ctx* c, int which, const float* ph, const int* arg)
{
    int rc = ctx_check(c);
    if (rc) return rc;
    if (!ph || !arg) return fail(PF_E_INVALID, "null state array");
    const int b = buf_index(c, which);
    const size_t n = (size_t)c->L * c->p.K;
    HIPCHK(hipMemcpyAsync(c->ph[b], ph, sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->arg[b], arg, sizeof(int) * n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(act_rebuild(c, b));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->have_state = true;
    c->ovf_until = c->nlaunch + kOvfNew;  // a new state may overflow
    return PF_OK;
}

int pf_download_state(pf_ctx* c, int which, float* ph, int* arg)
{
    int rc = ctx_check(c);
    if (rc) return rc;
    const int b = buf_index(c, which);
    const size_t n = (size_t)c->L * c->p.K;
    if (ph) HIPCHK(hipMemcpyAsync(ph, c->ph[b], sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
    if (arg) HIPCHK(hipMemcpyAsync(arg, c->arg[b], sizeof(int) * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return PF_OK;
}

int pf_apply_bc(pf_ctx* c, int which)
{
    int rc = ctx_check(c);
    if (rc) return rc;
    const int b = buf_index(c, which);
    const pf_params& p = c->p;
    HIPCHK(launch_bc_fill(c->ph[b], c->arg[b], p.fnx, p.fny, p.fnz, p.K, p.bcx, p.bcy, p.bcz,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return PF_OK;
}

// The step's lists hold their worst case, recomputed from the grid and the
// chunk depth (capacity.h), and the derived sizes the kernels index with
// agree with it; otherwise nothing is launched.
static int check_capacity(const pf_ctx* c)
{
    const pf_params& p = c->p;
    const long long items = worst_list_items(p.fnx, p.fny, p.fnz, c->kchunk, kPartialMin);
    if ((long long)c->nzchunk * c->items_chunk != items || items > c->wl_alloc ||
        worst_fix_entries(p.fnx, p.fny, p.fnz) > c->fix_alloc)
        return fail(PF_E_STATE, "work list capacity below its worst case (" + std::to_string(items) +
                                    " chunks, " + std::to_string(c->wl_alloc) + " allocated)");
    return PF_OK;
}

static int step_impl(pf_ctx* c, int nsteps, int first_step, int flags)
{
    if (!c->have_coords || !c->have_orient || !c->have_state)
        return fail(PF_E_STATE, "pf_step needs coordinates, orientations and an initial state");
    if (int rc = check_capacity(c)) return rc;
    const pf_params& p = c->p;
    if (p.thermal_type == PF_THERMAL_PARABOLOID && !c->psi_mac)
        return fail(PF_E_STATE, "paraboloid thermal field needs psi_mac");
    if (nsteps < 0) return fail(PF_E_INVALID, "nsteps must be >= 0");
    const int use_iuc = (flags & PF_STEP_INITIAL_UC) ? 1 : 0;
    const bool nucl = (flags & PF_STEP_NUCLEATE) != 0;
    const int end = first_step + nsteps;
    // Without nucleation (which runs between steps), each step's fix-up and
    // the next step's plan go out as one launch (apt_fixplan): the next
    // step's arguments are made before this step is launched, from the host
    // state as it stands after this launch.  The last step of the call ends
    // with its own fix-up, so no work is left pending between calls.
    StepArgs a{};
    bool have = false, planned = false;
    for (int s = first_step; s < end; ++s) {
        // t_cur_step = (2*kt+1)*dt*tau0 in float (3DgrainMPI/APTPhaseField.cu:863,919).
        const float t = s * p.dt * p.tau0;
        if (!have) {
            HIPCHK(act_renew_if_needed(c));
            HIPCHK(bound_lead(c));
            a = make_step_args(c, c->cur, t, use_iuc);
            planned = false;
        }
        const StepArgs cur = a;
        ++c->nlaunch;
        c->ep[1 - c->cur] = cur.ep_new;
        c->cur = 1 - c->cur;
        // an activity-map renewal (epoch counter near its end) must come
        // between this step and the next plan: no fused launch then
        have = !nucl && s + 1 < end && c->ep_ctr < 0x7fffffffu - 1u;
        if (have) {
            HIPCHK(bound_lead(c));
            a = make_step_args(c, c->cur, (s + 1) * p.dt * p.tau0, use_iuc);
        }
        HIPCHK(launch_step(cur, p.K, p.gather27, c->stream, &c->edge, nullptr, nullptr, nullptr,
                           nullptr, planned, have ? &a : nullptr));
        planned = have;
        if (nucl && (s & 1)) {
            int rc = nucl_launch(c, t, 2.0f * p.dt * p.tau0, use_iuc);
            if (rc) return rc;
        }
    }
    return PF_OK;
}

int pf_step(pf_ctx* c, int nsteps, int first_step, int flags)
{
    int rc = ctx_check(c);
    if (rc) return rc;
    return step_impl(c, nsteps, first_step, flags);
}

int pf_step_timed(pf_ctx* c, int nsteps, int first_step, int flags, float* ms)
{
    int rc = ctx_check(c);
    if (rc) return rc;
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    rc = step_impl(c, nsteps, first_step, flags);
    if (rc) return rc;
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    HIPCHK(hipEventSynchronize(c->ev1));
    float t = 0.0f;
    HIPCHK(hipEventElapsedTime(&t, c->ev0, c->ev1));
    if (ms) *ms = t;
    return PF_OK;
}

int pf_step_profiled(pf_ctx* c, int nsteps, int first_step, int flags, float* ms_main,
                     float* ms_fixup, float* ms_plan)
{
    int rc = ctx_check(c);
    if (rc) return rc;
    if (!c->have_coords || !c->have_orient || !c->have_state)
        return fail(PF_E_STATE, "pf_step needs coordinates, orientations and an initial state");
    if (flags & PF_STEP_NUCLEATE) return fail(PF_E_INVALID, "profiled stepping excludes nucleation");
    const pf_params& p = c->p;
    std::vector<hipEvent_t> ev(4 * (size_t)nsteps, nullptr);
    for (auto& e : ev) HIPCHK(hipEventCreate(&e));
    const int use_iuc = (flags & PF_STEP_INITIAL_UC) ? 1 : 0;
    hipError_t err = hipSuccess;
    for (int q = 0; q < nsteps && err == hipSuccess; ++q) {
        const int s = first_step + q;
        const float t = s * p.dt * p.tau0;
        err = act_renew_if_needed(c);
        if (err == hipSuccess) err = bound_lead(c);
        if (err != hipSuccess) break;
        StepArgs a = make_step_args(c, c->cur, t, use_iuc);
        err = launch_step(a, p.K, p.gather27, c->stream, &c->edge, ev[4 * q], ev[4 * q + 1],
                          ev[4 * q + 2], ev[4 * q + 3]);
        if (err != hipSuccess) break;
        ++c->nlaunch;
        c->ep[1 - c->cur] = a.ep_new;
        c->cur = 1 - c->cur;
    }
    if (err == hipSuccess) err = hipStreamSynchronize(c->stream);
    double tp = 0.0, tm = 0.0, tf = 0.0;
    for (int q = 0; q < nsteps && err == hipSuccess; ++q) {
        float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
        err = hipEventElapsedTime(&a0, ev[4 * q], ev[4 * q + 1]);
        if (err == hipSuccess) err = hipEventElapsedTime(&a1, ev[4 * q + 1], ev[4 * q + 2]);
        if (err == hipSuccess) err = hipEventElapsedTime(&a2, ev[4 * q + 2], ev[4 * q + 3]);
        tp += a0;
        tm += a1;
        tf += a2;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    if (err != hipSuccess) return fail(PF_E_HIP, std::string("profiled step: ") + hipGetErrorString(err));
    if (ms_main) *ms_main = (float)tm;
    if (ms_fixup) *ms_fixup = (float)tf;
    if (ms_plan) *ms_plan = (float)tp;
    return PF_OK;
}

int pf_collect(pf_ctx* c, int which, int flags, float* phi_out, int* alpha_out)
{
    int rc = ctx_check(c);
    if (rc) return rc;
    const int b = buf_index(c, which);
    const pf_params& p = c->p;
    if (!(flags & PF_COLLECT_NO_BC))
        HIPCHK(launch_bc_fill(c->ph[b], c->arg[b], p.fnx, p.fny, p.fnz, p.K, p.bcx, p.bcy, p.bcz,
                              c->stream));
    HIPCHK(launch_collect(c->ph[b], c->arg[b], c->phi, c->alpha, c->L, p.K,
                          (flags & PF_COLLECT_CLEAR_ALPHA) ? 1 : 0, c->stream));
    if (phi_out)
        HIPCHK(hipMemcpyAsync(phi_out, c->phi, sizeof(float) * c->L, hipMemcpyDeviceToHost, c->stream));
    if (alpha_out)
        HIPCHK(hipMemcpyAsync(alpha_out, c->alpha, sizeof(int) * c->L, hipMemcpyDeviceToHost,
                              c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return PF_OK;
}

int pf_qoi_volume(pf_ctx* c, int max_id, int* counts, int* first, long long* n_above)
{
    int rc = ctx_check(c);
    if (rc) return rc;
    if (max_id < 1 || !counts || !first) return fail(PF_E_INVALID, "max_id >= 1, counts and first required");
    const pf_params& p = c->p;
    int *dc = nullptr, *df = nullptr;
    unsigned long long* da = nullptr;
    if ((rc = dalloc(&dc, (size_t)max_id + 1)) || (rc = dalloc(&df, (size_t)max_id + 1)) ||
        (rc = dalloc(&da, 1))) {
        (void)hipFree(dc);
        (void)hipFree(df);
        return rc;
    }
    unsigned long long above = 0;
    hipError_t e = launch_qoi_volume(c->alpha, p.fnx, p.fny, p.fnz, max_id, dc, df, da, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(counts, dc, sizeof(int) * ((size_t)max_id + 1), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(first, df, sizeof(int) * ((size_t)max_id + 1), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(&above, da, sizeof(above), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(dc);
    (void)hipFree(df);
    (void)hipFree(da);
    if (e != hipSuccess) return fail(PF_E_HIP, std::string("qoi volume: ") + hipGetErrorString(e));
    if (n_above) *n_above = (long long)above;
    return PF_OK;
}

int pf_qoi_volume_noalloc(pf_ctx* c, int max_id, int* counts, int* first, long long* n_above)
{
    int rc = ctx_check(c);
    if (rc) return rc;
    if (max_id < 1 || !counts || !first) return fail(PF_E_INVALID, "max_id >= 1, counts and first required");
    if (max_id > PF_NOALLOC_MAX_ID)
        return fail(PF_E_INVALID, "max_id " + std::to_string(max_id) + " exceeds PF_NOALLOC_MAX_ID (" +
                                      std::to_string(PF_NOALLOC_MAX_ID) + ") of pf_qoi_volume_noalloc");
    void* base = scratch_of(c->device);
    if (!base) return fail(PF_E_HIP, "no scratch address for this device");
    const pf_params& p = c->p;
    int* dc = static_cast<int*>(base);
    int* df = dc + kScratchIds;
    auto* da = reinterpret_cast<unsigned long long*>(df + kScratchIds);
    // the launcher's fill kernels clear counts[0..max_id], first[0..max_id]
    // and the count of larger ids before the count kernel runs
    unsigned long long above = 0;
    hipError_t e = launch_qoi_volume(c->alpha, p.fnx, p.fny, p.fnz, max_id, dc, df, da, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(counts, dc, sizeof(int) * ((size_t)max_id + 1), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(first, df, sizeof(int) * ((size_t)max_id + 1), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(&above, da, sizeof(above), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(PF_E_HIP, std::string("qoi volume (no alloc): ") + hipGetErrorString(e));
    if (n_above) *n_above = (long long)above;
    return PF_OK;
}

int pf_qoi_junctions3d(pf_ctx* c, int* rows, int cap, int* count)
{
    int rc = ctx_check(c);
    if (rc) return rc;
    if (cap < 0 || (cap > 0 && !rows) || !count) return fail(PF_E_INVALID, "bad rows/cap/count");
    const pf_params& p = c->p;
    const int nblk = qoi_junction_blocks(p.fnx, p.fny, p.fnz);
    int* dr = nullptr;
    unsigned* dn = nullptr;
    unsigned* db = nullptr;
    if ((rc = dalloc(&dr, (size_t)8 * std::max(cap, 1))) || (rc = dalloc(&dn, 1)) ||
        (rc = dalloc(&db, (size_t)2 * nblk))) {
        (void)hipFree(dr);
        (void)hipFree(dn);
        return rc;
    }
    unsigned n = 0;
    // rows come back in the reference's scan order (k, j, i)
    hipError_t e = launch_qoi_junctions(c->alpha, p.fnx, p.fny, p.fnz, db, db + nblk, dr, cap, dn,
                                        c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(&n, dn, sizeof(n), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    const int m = (int)std::min<unsigned>(n, (unsigned)cap);
    if (e == hipSuccess && m > 0)
        e = hipMemcpy(rows, dr, sizeof(int) * 8 * (size_t)m, hipMemcpyDeviceToHost);
    (void)hipFree(dr);
    (void)hipFree(dn);
    (void)hipFree(db);
    if (e != hipSuccess) return fail(PF_E_HIP, std::string("qoi junctions: ") + hipGetErrorString(e));
    *count = (int)n;
    return PF_OK;
}

namespace {

// The tip plane after cur_tip on the collected alpha (3DgrainMPI/QOI.cpp:47-67,
// 162-185): per-plane liquid flags on the device, the scan over fnz flags on
// the host.  Leaves the flags' allocation to the caller's dl (fnz words).
hipError_t device_tip(pf_ctx* c, unsigned* dl, int cur_tip, int* tip)
{
    const pf_params& p = c->p;
    std::vector<unsigned> liq((size_t)p.fnz, 0u);
    hipError_t e = launch_qoi_liquid_planes(c->alpha, p.fnx, p.fny, p.fnz, dl, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(liq.data(), dl, sizeof(unsigned) * p.fnz, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return e;
    int t = cur_tip;
    for (;;) {  // step up to the first plane holding liquid
        ++t;
        if (t >= p.fnz - 1 || liq[t]) break;
    }
    *tip = t - 1;
    return hipSuccess;
}

}  // namespace

int pf_qoi_tip(pf_ctx* c, int cur_tip, int* new_tip)
{
    int rc = ctx_check(c);
    if (rc) return rc;
    const pf_params& p = c->p;
    if (!new_tip) return fail(PF_E_INVALID, "new_tip required");
    if (cur_tip < 0 || cur_tip > p.fnz - 2) return fail(PF_E_INVALID, "cur_tip out of range");
    unsigned* dl = nullptr;
    if ((rc = dalloc(&dl, (size_t)p.fnz))) return rc;
    int tip = cur_tip;
    const hipError_t e = device_tip(c, dl, cur_tip, &tip);
    (void)hipFree(dl);
    if (e != hipSuccess) return fail(PF_E_HIP, std::string("qoi tip: ") + hipGetErrorString(e));
    *new_tip = tip;
    return PF_OK;
}

int pf_qoi_line(pf_ctx* c, int cur_tip, int* new_tip, int* total, int* extra)
{
    int rc = ctx_check(c);
    if (rc) return rc;
    const pf_params& p = c->p;
    if (!new_tip || !total || !extra) return fail(PF_E_INVALID, "new_tip, total and extra required");
    if (cur_tip < 0 || cur_tip > p.fnz - 2) return fail(PF_E_INVALID, "cur_tip out of range");
    const int ng = p.num_theta;
    unsigned* dl = nullptr;
    int *dt = nullptr, *de = nullptr;
    if ((rc = dalloc(&dl, (size_t)p.fnz)) || (rc = dalloc(&dt, (size_t)ng)) ||
        (rc = dalloc(&de, (size_t)ng))) {
        (void)hipFree(dl);
        (void)hipFree(dt);
        return rc;
    }
    int tip = cur_tip;
    hipError_t e = device_tip(c, dl, cur_tip, &tip);
    if (e == hipSuccess) e = launch_qoi_area(c->alpha, p.fnx, p.fny, p.fnz, ng, tip, dt, de, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(total, dt, sizeof(int) * ng, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(extra, de, sizeof(int) * ng, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(dl);
    (void)hipFree(dt);
    (void)hipFree(de);
    if (e != hipSuccess) return fail(PF_E_HIP, std::string("qoi line: ") + hipGetErrorString(e));
    *new_tip = tip;
    return PF_OK;
}

int pf_md_enable(pf_ctx* c, int fnz_full)
{
    int rc = ctx_check(c);
    if (rc) return rc;
    if (fnz_full < c->p.fnz) return fail(PF_E_INVALID, "fnz_full must be >= fnz");
    if (c->alpha_full) (void)hipFree(c->alpha_full);
    if (c->loss) (void)hipFree(c->loss);
    c->alpha_full = nullptr;
    c->loss = nullptr;
    const size_t nf = (size_t)c->p.fnx * c->p.fny * fnz_full;
    if ((rc = dalloc(&c->alpha_full, nf)) || (rc = dalloc(&c->loss, c->p.num_theta))) return rc;
    HIPCHK(hipMemsetAsync(c->alpha_full, 0, nf * sizeof(int), c->stream));
    HIPCHK(hipMemsetAsync(c->loss, 0, c->p.num_theta * sizeof(int), c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->fnz_full = fnz_full;
    return PF_OK;
}

int pf_md_tip(pf_ctx* c, int* tip)
{
    int rc = ctx_check(c);
    if (rc) return rc;
    if (!tip) return fail(PF_E_INVALID, "null tip");
    const pf_params& p = c->p;
    const int b = c->cur;
    int t = *tip;
    if (t < 0 || t >= p.fnz) return fail(PF_E_INVALID, "tip out of range");
    HIPCHK(launch_bc_fill(c->ph[b], c->arg[b], p.fnx, p.fny, p.fnz, p.K, p.bcx, p.bcy, p.bcz,
                          c->stream));
    // tip_mvf scan (3DgrainMPI/PhaseField.cu:391-394) on the exact plane sums.
    // It reads planes from t up to the first liquid one, a few planes above
    // the tip: the sums come kTipWindow planes at a time, as the scan reaches
    // them (the same conditions in the same order as over all planes).
    constexpr int kTipWindow = 16;
    const double denom = (double)p.K * p.fnx * p.fny;
    int hi = t;  // sums known for planes [*tip, hi)
    for (;;) {
        if (t >= hi) {
            const int to = std::min(hi + kTipWindow, p.fnz);
            HIPCHK(launch_plane_sums(c->ph[b], c->sums, p.fnx, p.fny, p.fnz, p.K, hi, to, c->stream));
            HIPCHK(hipMemcpyAsync(c->h_sums.data() + hi, c->sums + hi, sizeof(long long) * (to - hi),
                                  hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            hi = to;
        }
        if (!(std::ldexp((double)c->h_sums[t], -40) / denom > kLS && t < p.fnz - 1)) break;
        ++t;
    }
    *tip = t;
    return PF_OK;
}

int pf_md_shift(pf_ctx* c, int move_count)
{
    int rc = ctx_check(c);
    if (rc) return rc;
    if (!c->alpha_full) return fail(PF_E_STATE, "pf_md_enable first");
    const pf_params& p = c->p;
    const size_t A = (size_t)p.fnx * p.fny;
    if ((size_t)(move_count + 2) * A > (size_t)c->fnz_full * A)
        return fail(PF_E_INVALID, "alpha_full archive is full");
    const int b = c->cur;
    // APTcollect_PF on the buffer as stored (3DgrainMPI/APTPhaseField.cu:683),
    // then APTmove_frame with the other buffer as scratch, then APTcopy_frame.
    HIPCHK(launch_collect(c->ph[b], c->arg[b], c->phi, c->alpha, c->L, p.K, 0, c->stream));
    HIPCHK(launch_move_frame(c->ph[b], c->arg[b], c->ph[1 - b], c->arg[1 - b], c->z, c->zbuf,
                             c->alpha, c->alpha_full, c->loss, move_count, p.fnx, p.fny, p.fnz,
                             p.K, p.num_theta, c->stream));
    // both buffers were rewritten (the other one served as scratch)
    HIPCHK(act_rebuild(c, b));
    HIPCHK(act_rebuild(c, 1 - b));
    HIPCHK(hipStreamSynchronize(c->stream));
    return PF_OK;
}

int pf_md_download(pf_ctx* c, int* alpha_full, int* loss_area, float* z)
{
    int rc = ctx_check(c);
    if (rc) return rc;
    const pf_params& p = c->p;
    if (alpha_full) {
        if (!c->alpha_full) return fail(PF_E_STATE, "pf_md_enable first");
        HIPCHK(hipMemcpyAsync(alpha_full, c->alpha_full,
                              sizeof(int) * (size_t)p.fnx * p.fny * c->fnz_full,
                              hipMemcpyDeviceToHost, c->stream));
    }
    if (loss_area) {
        if (!c->loss) return fail(PF_E_STATE, "pf_md_enable first");
        HIPCHK(hipMemcpyAsync(loss_area, c->loss, sizeof(int) * p.num_theta, hipMemcpyDeviceToHost,
                              c->stream));
    }
    if (z) HIPCHK(hipMemcpyAsync(z, c->z, sizeof(float) * p.fnz, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return PF_OK;
}

int pf_nucl_reset(pf_ctx* c, int mode)
{
    int rc = ctx_check(c);
    if (rc) return rc;
    const pf_params& p = c->p;
    if (p.pts_cell < 0) return fail(PF_E_INVALID, "pts_cell must be >= 0");
    const int w = 2 * p.pts_cell + 1;
    // cnx = (fnx - 1 - 2*haloWidth)/(2*pts_cell+1) (3DgrainMPI/APTPhaseField.cu:735-737)
    c->cnx = (p.fnx - 3) / w;
    c->cny = (p.fny - 3) / w;
    c->cnz = (p.fnz - 3) / w;
    const size_t n = (size_t)c->cnx * c->cny * c->cnz;
    if (c->nstatus) (void)hipFree(c->nstatus);
    if (c->ball) (void)hipFree(c->ball);
    c->nstatus = nullptr;
    c->ball = nullptr;
    if ((rc = dalloc(&c->nstatus, n > 0 ? n : 1)) || (rc = dalloc(&c->ball, (size_t)w * w * w)))
        return rc;
    // tanh ball of add_nucl (3DgrainMPI/APTPhaseField.cu:515-529), a function of
    // the offset only, evaluated once on the host.
    std::vector<float> ball((size_t)w * w * w);
    int q = 0;
    for (int lk = -p.pts_cell; lk <= p.pts_cell; ++lk)
        for (int lj = -p.pts_cell; lj <= p.pts_cell; ++lj)
            for (int li = -p.pts_cell; li <= p.pts_cell; ++li) {
                const float pc = (float)p.pts_cell;
                const float d = p.dx * ((1.0f + pc) / 2.0f
                                        - sqrtf((float)(li * li + lj * lj + lk * lk)));
                ball[q++] = tanhf(d / p.sqrt2);
            }
    HIPCHK(hipMemcpyAsync(c->ball, ball.data(), sizeof(float) * ball.size(),
                          hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemsetAsync(c->nstatus, 0, sizeof(int) * (n > 0 ? n : 1), c->stream));
    if (mode == 1)
        HIPCHK(launch_nucl_status(c->phi0, c->nstatus, c->cnx, c->cny, c->cnz, p.pts_cell, p.fnx,
                                  p.fny, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return PF_OK;
}

int pf_nucl_events(pf_ctx* c, long long* events)
{
    int rc = ctx_check(c);
    if (rc) return rc;
    unsigned long long v = 0;
    HIPCHK(hipMemcpyAsync(&v, c->counters + 1, sizeof(v), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (events) *events = (long long)v;
    return PF_OK;
}

int pf_overflow_count(pf_ctx* c, long long* count)
{
    int rc = ctx_check(c);
    if (rc) return rc;
    unsigned long long v = 0;
    HIPCHK(hipMemcpyAsync(&v, c->counters, sizeof(v), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (count) *count = (long long)v;
    return PF_OK;
}

int pf_copy_bandwidth(int device, unsigned long long bytes, int iters, float* gbs)
{
    if (bytes < 16 || iters < 1 || !gbs) return fail(PF_E_INVALID, "bad arguments");
    HIPCHK(hipSetDevice(device));
    bytes &= ~15ull;
    void *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = PF_OK;
    float ms = 0.0f;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
        rc = fail(PF_E_NOMEM, "copy buffers");
    } else if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
               hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess ||
               hipMemsetAsync(a, 0, bytes, st) != hipSuccess ||
               launch_copy(a, b, bytes, st) != hipSuccess ||
               hipEventRecord(e0, st) != hipSuccess) {
        rc = fail(PF_E_HIP, "copy set-up failed");
    } else {
        for (int q = 0; q < iters && rc == PF_OK; ++q)
            if (launch_copy(a, b, bytes, st) != hipSuccess) rc = fail(PF_E_HIP, "copy launch failed");
        if (rc == PF_OK && (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                            hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
            rc = fail(PF_E_HIP, "copy timing failed");
    }
    if (rc == PF_OK) *gbs = (float)(2.0 * bytes * iters / (ms * 1e-3) / 1e9);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}

int pf_activity(pf_ctx* c, int which, unsigned char* map, long long* active, long long* total)
{
    int rc = ctx_check(c);
    if (rc) return rc;
    const int b = buf_index(c, which);
    const pf_params& p = c->p;
    std::vector<unsigned> h((size_t)c->ntiles * p.fnz);
    HIPCHK(hipMemcpyAsync(h.data(), c->act[b], sizeof(unsigned) * h.size(), hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    long long n = 0;
    for (int k = 0; k < p.fnz; ++k)
        for (int t = 0; t < c->ntiles; ++t) {
            const size_t q = (size_t)k * c->ntiles + t;
            const bool on = k > 0 && k < p.fnz - 1 && h[q] >= c->ep[b];
            if (map) map[q] = on ? 1 : 0;
            n += on;
        }
    if (active) *active = n;
    if (total) *total = (long long)c->ntiles * (p.fnz - 2);
    return PF_OK;
}

int pf_batch_create(pf_ctx* const* ctxs, int n, pf_batch** out)
{
    if (!ctxs || !out || n < 1) return fail(PF_E_INVALID, "bad arguments");
    if (n > kMaxBatch) return fail(PF_E_INVALID, "a batch holds at most 1024 simulations");
    *out = nullptr;
    for (int i = 0; i < n; ++i) {
        if (!ctxs[i]) return fail(PF_E_INVALID, "null context in batch");
        for (int j = 0; j < i; ++j)
            if (ctxs[j] == ctxs[i]) return fail(PF_E_INVALID, "context listed twice in batch");
        if (ctxs[i]->device != ctxs[0]->device || ctxs[i]->p.K != ctxs[0]->p.K ||
            ctxs[i]->p.gather27 != ctxs[0]->p.gather27)
            return fail(PF_E_INVALID, "batch members must share device, K and gather variant");
    }
    HIPCHK(hipSetDevice(ctxs[0]->device));
    pf_batch* b = new pf_batch();
    b->ctxs.assign(ctxs, ctxs + n);
    b->device = ctxs[0]->device;
    b->K = ctxs[0]->p.K;
    b->g27 = ctxs[0]->p.gather27;
    if (const char* kc = getenv("CUPF_KCHUNK")) {
        const int v = atoi(kc);
        if (v >= pf_batch::kMinChunk && v <= 61) b->kchunk = v;
    }
    b->step_blocks = batch_kernel_slots(b->K, b->g27, b->device);
    // 2 fix-up workgroups per CU (3.60 s for the 64-simulation ensemble, against 3.63-3.65 s
    // with the 4 or 6 per CU that fit)
    {
        int ncu = 0;
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, b->device) == hipSuccess)
            b->fix_blocks = std::min(fixup_kernel_slots(b->K, b->g27, b->device, true), 2 * ncu);
    }
    auto bail = [&](int code) {
        batch_free(b);
        return code;
    };
    if (b->step_blocks < 1 || b->fix_blocks < 1)
        return bail(fail(PF_E_HIP, "occupancy query for the batch step or fix-up kernel failed"));
    if (const char* fb = getenv("CUPF_FIX_BLOCKS")) {
        const int v = atoi(fb);
        if (v >= 1 && v < b->fix_blocks) b->fix_blocks = v;
    }
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(PF_E_HIP, "hipStreamCreate failed"));
    for (pf_ctx* c : b->ctxs)
        b->wl_cap += (size_t)((c->p.fnz - 2 + pf_batch::kMinChunk - 1) / pf_batch::kMinChunk) *
                     c->items_chunk;
    int rc;
    if ((rc = dalloc(&b->sims, (size_t)n)) || (rc = dalloc(&b->wl_sim, b->wl_cap + 1)) ||
        (rc = dalloc(&b->wl_item, b->wl_cap + 1)) || (rc = dalloc(&b->wl_need, b->wl_cap + 1)) ||
        (rc = dalloc(&b->wl_act, b->wl_cap + 1)) || (rc = dalloc(&b->cnt, 10)))
        return bail(rc);
    for (hipEvent_t& e : b->lead)
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
            return bail(fail(PF_E_HIP, "hipEventCreate failed"));
    if (hipHostMalloc(&b->seen, 2 * sizeof(unsigned long long)) != hipSuccess) {
        b->seen = nullptr;
        return bail(fail(PF_E_NOMEM, "hipHostMalloc of the step variant report failed"));
    }
    b->seen[0] = b->seen[1] = 0ull;
    if (const char* sv = getenv("CUPF_STEP_VARIANT")) b->variant_force = atoi(sv) ? 1 : 0;
    *out = b;
    return PF_OK;
}

int pf_batch_destroy(pf_batch* b)
{
    if (!b) return PF_OK;
    (void)hipStreamSynchronize(b->stream);
    batch_free(b);
    return PF_OK;
}

int pf_batch_set_kchunk(pf_batch* b, int kchunk)
{
    if (!b) return fail(PF_E_INVALID, "null batch");
    if (kchunk < pf_batch::kMinChunk || kchunk > 61) return fail(PF_E_INVALID, "kchunk must be 2..61");
    b->kchunk = kchunk;
    return PF_OK;
}

int pf_batch_get_kchunk(pf_batch* b, int* kchunk)
{
    if (!b || !kchunk) return fail(PF_E_INVALID, "null argument");
    *kchunk = b->kchunk;
    return PF_OK;
}

int pf_batch_step(pf_batch* b, int nsteps, const int* first_step, int flags)
{
    if (!b || !first_step) return fail(PF_E_INVALID, "null argument");
    if (nsteps < 0) return fail(PF_E_INVALID, "nsteps must be >= 0");
    if (flags & PF_STEP_NUCLEATE) return fail(PF_E_INVALID, "batched steps exclude nucleation");
    for (pf_ctx* c : b->ctxs) {
        if (!c->have_coords || !c->have_orient || !c->have_state)
            return fail(PF_E_STATE, "batch member needs coordinates, orientations and a state");
        if (c->p.thermal_type == PF_THERMAL_PARABOLOID && !c->psi_mac)
            return fail(PF_E_STATE, "paraboloid thermal field needs psi_mac");
    }
    {
        // the batch list holds every member's worst case at any depth >= kMinChunk
        long long need = 0;
        for (pf_ctx* c : b->ctxs) {
            if (int rc = check_capacity(c)) return rc;
            need += worst_list_items(c->p.fnx, c->p.fny, c->p.fnz, pf_batch::kMinChunk, kPartialMin);
        }
        if (need > (long long)b->wl_cap)
            return fail(PF_E_STATE, "batch work list capacity below its worst case");
    }
    HIPCHK(hipSetDevice(b->device));
    const int use_iuc = (flags & PF_STEP_INITIAL_UC) ? 1 : 0;
    const bool use_graph = (flags & PF_BATCH_GRAPH) != 0;
    std::vector<int> fs(first_step, first_step + b->ctxs.size());
    int left = nsteps;
    if (!(flags & PF_BATCH_AUTOTUNE)) return batch_run(b, left, fs, use_iuc, use_graph);
    // The chunk depth changes only how the work is cut, never the results.
    // Every kTuneEvery steps, time a few real steps with each candidate and
    // keep the fastest for the rest of the segment: the best depth moves as
    // the simulations evolve (overflow start-up, then thin fronts).  The
    // candidates are timed in two rounds, the second in reverse order, so a
    // cost that drifts steadily over the tuning steps weighs on all alike.
    constexpr int kTuneEvery = 400, kTuneSteps = 4;
    const int cand[] = {2, 4, 8, 16};
    constexpr int nc = (int)(sizeof(cand) / sizeof(cand[0]));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    int rc = PF_OK;
    while (left > 0 && rc == PF_OK) {
        if (left >= 2 * nc * (kTuneSteps + 1) + 16) {
            float tot[nc] = {};
            for (int round = 0; round < 2 && rc == PF_OK; ++round) {
                for (int u = 0; u < nc && rc == PF_OK; ++u) {
                    const int c = round == 0 ? u : nc - 1 - u;
                    b->kchunk = cand[c];
                    if ((rc = batch_run(b, 1, fs, use_iuc, false))) break;  // warm the depth
                    left -= 1;
                    if (hipEventRecord(e0, b->stream) != hipSuccess) {
                        rc = fail(PF_E_HIP, "event");
                        break;
                    }
                    if ((rc = batch_run(b, kTuneSteps, fs, use_iuc, false))) break;
                    left -= kTuneSteps;
                    float ms = 0.0f;
                    if (hipEventRecord(e1, b->stream) != hipSuccess ||
                        hipEventSynchronize(e1) != hipSuccess ||
                        hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                        rc = fail(PF_E_HIP, "autotune timing");
                        break;
                    }
                    tot[c] += ms;
                }
            }
            if (rc) break;
            int best = 0;
            for (int c = 1; c < nc; ++c) best = tot[c] < tot[best] ? c : best;
            b->kchunk = cand[best];
        }
        const int seg = std::min(left, kTuneEvery);
        rc = batch_run(b, seg, fs, use_iuc, use_graph);
        left -= seg;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

int pf_sync(pf_ctx* c)
{
    int rc = ctx_check(c);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    return PF_OK;
}

}  // extern "C"
