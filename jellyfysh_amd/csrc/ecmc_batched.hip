// ecmc_batched.hip -- thousands of independent ECMC replicas of uneven size advanced by one
// persistent kernel per size class.
//
// Each workgroup (WPR wavefronts) pulls replicas from a queue sorted largest first (LPT)
// and runs `legs` legs on each.  A replica's FP32 positions and its cell lists live in LDS;
// its FP64 positions stay in HBM, where only the active unit's row changes.  One leg:
//   1. horizon h = min(displacement to the next cell wall, rest of the chain), FP64;
//   2. every candidate (occupants of the 27 nearby cells, or all units when the box holds
//      fewer than three cells per side) and factor gets an FP32 lower bound on its event
//      displacement, with a proven downward margin; only those whose bound is <= h enter
//      a contender list in LDS;
//   3. the contenders are inverted exactly in FP64 (correctly rounded pow/log) and a
//      workgroup min-reduce over (displacement, kind, target, factor) picks the event;
//   4. thread 0 commits it (lifting, wall snap, or end of chain with a Philox-drawn new
//      active unit) and updates the cell lists.
// Because every bound is <= the exact FP64 value, the committed event is the one a full
// FP64 evaluation of every candidate selects; the audit mode checks that on the device.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <algorithm>
#include <vector>

#include "../../include/jellyfysh_amd.h"
#include "ecmc_common.h"
#include "ecmc_math.h"
#include "ecmc_potentials.h"

#include "ecmc_batched_kernel.h"

namespace ecmc {

// Work items per replica and step (the step's legs split into this many chunks at most).
constexpr int WORK_CHUNKS = 16;

// Start of a step for one class: every member free with no chunk done, claim cursor at 0.
__global__ void reset_work(int32_t *state, int count, int32_t *q) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) state[i] = 0;
    if (i == 0) q[0] = 0;
}

}  // namespace ecmc

using namespace ecmc;

struct ecmc_batch {
    ecmc_batch_desc desc;
    int n_rep = 0;
    std::vector<int32_t> n;
    std::vector<ReplicaDev> rep_host;
    ReplicaDev *d_rep = nullptr;
    double *d_pos = nullptr, *d_bounds = nullptr;
    int64_t total_units = 0;
    // size classes
    int n_classes = 0;
    std::vector<int32_t> class_max_n, class_wpr, class_count, class_nmax, class_cellmax;
    std::vector<int32_t *> d_order, d_state;
    int32_t *d_queue = nullptr;
    std::vector<hipStream_t> streams;
    hipStream_t main_stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<hipEvent_t> fork;
    // per class: timing events around its kernel on the class's stream (direct launches only:
    // events recorded inside a captured graph carry no timestamps) and its kernel time of the
    // last ecmc_batch_run call
    std::vector<hipEvent_t> cls_ev0, cls_ev1;
    std::vector<double> cls_ms;
    std::vector<int32_t> cls_launches;
    ecmc_batch_event *d_log = nullptr;
    int32_t log_legs = 0;
    unsigned long long *d_audit = nullptr;
    int32_t mode = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int32_t graph_legs = -1;
    bool class_timing = false;
    int32_t max_blocks = 0;   // limit of every class's persistent grid (0 = as many as fit)
};

// Add the kernel time of every class of the step that just completed.
static int add_class_times(ecmc_batch *b) {
    for (int c = 0; c < b->n_classes; c++) {
        if (b->class_count[c] == 0) continue;
        float f = 0.f;
        if (hipEventElapsedTime(&f, b->cls_ev0[c], b->cls_ev1[c]) != hipSuccess)
            return ecmc_fail(ECMC_ERR_HIP, "class event timing failed");
        b->cls_ms[c] += f;
        b->cls_launches[c] += 1;
    }
    return ECMC_OK;
}

static int launch_class(ecmc_batch *b, int c, int legs, hipStream_t st) {
    if (b->class_count[c] == 0) return ECMC_OK;
    BatchArgs A;
    memset(&A, 0, sizeof(A));
    A.rep = b->d_rep;
    A.pos64 = b->d_pos;
    A.bounds = b->d_bounds;
    A.order = b->d_order[c];
    A.state = b->d_state[c];
    A.count = b->class_count[c];
    // the step's legs in at most WORK_CHUNKS chunks per replica: a replica is free again after
    // each chunk, so blocks keep busy to the end of the step whatever the replica count
    const int nch = std::max(1, std::min(WORK_CHUNKS, legs));
    A.chunk_legs = std::max(1, (legs + nch - 1) / nch);
    A.nchunks = legs > 0 ? (legs + A.chunk_legs - 1) / A.chunk_legs : 0;
    A.nmax = b->class_nmax[c];
    A.cellmax = b->class_cellmax[c];
    A.legs = legs;
    A.mode = b->mode;
    A.log_legs = b->log_legs;
    A.seed_lo = (uint32_t)b->desc.seed;
    A.beta = b->desc.beta;
    A.chain_length = b->desc.chain_length;
    A.fac.n_factors = b->desc.n_factors;
    for (int f = 0; f < 2; f++) { A.fac.power[f] = b->desc.power[f]; A.fac.prefactor[f] = b->desc.prefactor[f]; }
    A.queue = b->d_queue + 4 * c;
    hipLaunchKernelGGL(reset_work, dim3((A.count + 255) / 256), dim3(256), 0, st, A.state, A.count, A.queue);
    A.log = b->d_log;
    A.audit = b->d_audit;
    int wpr = b->class_wpr[c];
    size_t lds = batched_lds_bytes(A.nmax, A.cellmax, wpr);
    int dev = 0, cus = 0;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    int per_cu = 1;
#define ECMC_LAUNCH(W)                                                                                     \
    {                                                                                                      \
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batched_kernel<W>, 64 * W, lds);              \
        if (per_cu < 1) return ecmc_fail(ECMC_ERR_CONFIGURATION, "replica does not fit in LDS");           \
        int grid = std::min(A.count, per_cu * std::max(cus, 1));                                           \
        if (b->max_blocks > 0) grid = std::min(grid, b->max_blocks);                                       \
        hipLaunchKernelGGL(batched_kernel<W>, dim3(grid), dim3(64 * W), lds, st, A);                       \
    }
    switch (wpr) {
    case 1: ECMC_LAUNCH(1); break;
    case 2: ECMC_LAUNCH(2); break;
    case 4: ECMC_LAUNCH(4); break;
    case 8: ECMC_LAUNCH(8); break;
    default: return ecmc_fail(ECMC_ERR_ARGUMENT, "wavefronts per replica must be 1, 2, 4 or 8");
    }
#undef ECMC_LAUNCH
    if (hipGetLastError() != hipSuccess) return ecmc_fail(ECMC_ERR_HIP, "batched kernel launch failed");
    return ECMC_OK;
}

// One step: reset the work queues, then every size class on its own stream.
static int enqueue_step(ecmc_batch *b, int legs, hipStream_t main, bool timed) {
    hipEventRecord(b->fork[0], main);
    // largest replicas first: their blocks are resident from the start, and the smaller
    // classes fill the slots left beside them
    for (int c = b->n_classes - 1; c >= 0; c--) {
        hipStreamWaitEvent(b->streams[c], b->fork[0], 0);
        if (timed) hipEventRecord(b->cls_ev0[c], b->streams[c]);
        int rc = launch_class(b, c, legs, b->streams[c]);
        if (rc) return rc;
        if (timed) hipEventRecord(b->cls_ev1[c], b->streams[c]);
        hipEventRecord(b->fork[c + 1], b->streams[c]);
        hipStreamWaitEvent(main, b->fork[c + 1], 0);
    }
    return ECMC_OK;
}

extern "C" int ecmc_batch_create(const ecmc_batch_desc *desc, const int32_t *n_per_replica,
                                 const double *system_length, const double *positions, ecmc_batch **out) {
    if (!desc || !n_per_replica || !system_length || !positions || !out) return ecmc_fail(ECMC_ERR_ARGUMENT, "null argument");
    *out = nullptr;
    if (desc->n_replicas < 1) return ecmc_fail(ECMC_ERR_ARGUMENT, "at least one replica");
    if (desc->n_factors < 1 || desc->n_factors > 2) return ecmc_fail(ECMC_ERR_CONFIGURATION, "one or two factors");
    for (int f = 0; f < desc->n_factors; f++) {
        double p = desc->power[f];
        if (!(p == 6.0 || p == 12.0)) return ecmc_fail(ECMC_ERR_CONFIGURATION, "factor powers 6 and 12 are supported");
        if (desc->prefactor[f] == 0.0) return ecmc_fail(ECMC_ERR_CONFIGURATION, "zero prefactor");
    }
    if (!(desc->cutoff > 0.0) || !(desc->beta > 0.0) || !(desc->chain_length > 0.0))
        return ecmc_fail(ECMC_ERR_CONFIGURATION, "cutoff, beta and chain_length must be positive");
    ecmc_batch *b = new ecmc_batch();
    b->desc = *desc;
    b->n_rep = desc->n_replicas;
    b->n.assign(n_per_replica, n_per_replica + b->n_rep);
    std::vector<double> bounds;
    int64_t off = 0;
    b->rep_host.resize(b->n_rep);
    std::vector<double> soa;
    const double *pin = positions;
    for (int r = 0; r < b->n_rep; r++) {
        int nr = b->n[r];
        double L = system_length[r];
        if (nr < 2 || nr > 16384 || !(L > 0.0)) { delete b; return ecmc_fail(ECMC_ERR_CONFIGURATION, "replica size out of range"); }
        ReplicaDev &R = b->rep_host[r];
        memset(&R, 0, sizeof(R));
        R.n = nr;
        R.off = off;
        R.L = L;
        R.cutoff = std::min(desc->cutoff, L / 2.0);
        int cps = (int)(L / R.cutoff);
        if (cps >= 3) {
            if (cps > B_MAX_CPS) { delete b; return ecmc_fail(ECMC_ERR_CONFIGURATION, "more than 8 cells per side"); }
            R.cps = cps;
            R.side = L / cps;
            R.bound_off = (int32_t)bounds.size();
            for (int c = 0; c < cps; c++) {
                double lo = c * R.side;
                if (lo > 0.0) {
                    while ((int)(lo / R.side) == c) lo = nextafter(lo, -INFINITY);
                    while ((int)(lo / R.side) < c) lo = nextafter(lo, INFINITY);
                }
                bounds.push_back(lo);
            }
            bounds.push_back(L);
        } else {
            R.cps = 0;
            R.side = L;
            R.bound_off = 0;
        }
        R.active = 0;
        R.direction = 0;
        R.chain_left = desc->chain_length;
        R.leg = 0;
        for (int i = 0; i < nr; i++)
            for (int q = 0; q < 3; q++) {
                double x = pin[3 * i + q];
                if (!(x >= 0.0 && x < L)) { delete b; return ecmc_fail(ECMC_ERR_STATE, "initial position outside [0, L)"); }
            }
        soa.resize((size_t)(off + 3 * nr));
        for (int i = 0; i < nr; i++)
            for (int q = 0; q < 3; q++) soa[(size_t)off + (size_t)q * nr + i] = pin[3 * i + q];
        pin += 3 * (size_t)nr;
        off += 3 * nr;
    }
    b->total_units = off / 3;
    if (bounds.empty()) bounds.push_back(0.0);
    // default size classes (replaced by ecmc_batch_set_classes)
    b->n_classes = 0;
    int32_t def_max[3] = {256, 768, 16384}, def_wpr[3] = {1, 2, 4};
    int rc = 0;
    if (hipMalloc((void **)&b->d_rep, sizeof(ReplicaDev) * b->n_rep) != hipSuccess ||
        hipMalloc((void **)&b->d_pos, sizeof(double) * soa.size()) != hipSuccess ||
        hipMalloc((void **)&b->d_bounds, sizeof(double) * bounds.size()) != hipSuccess ||
        hipMalloc((void **)&b->d_queue, sizeof(int32_t) * 4 * 8) != hipSuccess ||
        hipMalloc((void **)&b->d_audit, sizeof(unsigned long long)) != hipSuccess) {
        ecmc_batch_destroy(b);
        return ecmc_fail(ECMC_ERR_HIP, "hipMalloc failed");
    }
    hipMemcpy(b->d_rep, b->rep_host.data(), sizeof(ReplicaDev) * b->n_rep, hipMemcpyHostToDevice);
    hipMemcpy(b->d_pos, soa.data(), sizeof(double) * soa.size(), hipMemcpyHostToDevice);
    hipMemcpy(b->d_bounds, bounds.data(), sizeof(double) * bounds.size(), hipMemcpyHostToDevice);
    int32_t qzero[4 * 8] = {0};
    hipMemcpy(b->d_queue, qzero, sizeof(qzero), hipMemcpyHostToDevice);
    unsigned long long zero = 0;
    hipMemcpy(b->d_audit, &zero, sizeof(zero), hipMemcpyHostToDevice);
    hipStreamCreate(&b->main_stream);
    hipEventCreate(&b->ev0);
    hipEventCreate(&b->ev1);
    rc = ecmc_batch_set_classes(b, 3, def_max, def_wpr);
    if (rc) { ecmc_batch_destroy(b); return rc; }
    *out = b;
    return ECMC_OK;
}

extern "C" int ecmc_batch_set_classes(ecmc_batch *b, int32_t n_classes, const int32_t *class_max_n,
                                      const int32_t *class_wpr) {
    if (!b || n_classes < 1 || n_classes > 8 || !class_max_n || !class_wpr) return ecmc_fail(ECMC_ERR_ARGUMENT, "bad classes");
    for (int c = 0; c < n_classes; c++) {
        int w = class_wpr[c];
        if (w != 1 && w != 2 && w != 4 && w != 8) return ecmc_fail(ECMC_ERR_ARGUMENT, "wavefronts per replica must be 1, 2, 4 or 8");
        if (c > 0 && class_max_n[c] <= class_max_n[c - 1]) return ecmc_fail(ECMC_ERR_ARGUMENT, "class bounds must increase");
    }
    if (class_max_n[n_classes - 1] < *std::max_element(b->n.begin(), b->n.end()))
        return ecmc_fail(ECMC_ERR_ARGUMENT, "the last class must hold the largest replica");
    for (auto p : b->d_order) hipFree(p);
    for (auto p : b->d_state) hipFree(p);
    for (auto s : b->streams) hipStreamDestroy(s);
    for (auto e : b->fork) hipEventDestroy(e);
    for (auto e : b->cls_ev0) hipEventDestroy(e);
    for (auto e : b->cls_ev1) hipEventDestroy(e);
    b->d_state.clear();
    b->d_order.clear(); b->streams.clear(); b->fork.clear(); b->cls_ev0.clear(); b->cls_ev1.clear();
    if (b->graph_exec) { hipGraphExecDestroy(b->graph_exec); b->graph_exec = nullptr; }
    if (b->graph) { hipGraphDestroy(b->graph); b->graph = nullptr; }
    b->n_classes = n_classes;
    b->class_max_n.assign(class_max_n, class_max_n + n_classes);
    b->class_wpr.assign(class_wpr, class_wpr + n_classes);
    b->class_count.assign(n_classes, 0);
    b->class_nmax.assign(n_classes, 2);
    b->class_cellmax.assign(n_classes, 0);
    std::vector<std::vector<int32_t>> members(n_classes);
    for (int r = 0; r < b->n_rep; r++) {
        int c = 0;
        while (b->n[r] > class_max_n[c]) c++;
        members[c].push_back(r);
    }
    for (int c = 0; c < n_classes; c++) {
        auto &m = members[c];
        std::stable_sort(m.begin(), m.end(), [&](int x, int y) { return b->n[x] > b->n[y]; });  // LPT
        b->class_count[c] = (int32_t)m.size();
        for (int r : m) {
            b->class_nmax[c] = std::max(b->class_nmax[c], b->n[r]);
            int cps = b->rep_host[r].cps;
            b->class_cellmax[c] = std::max(b->class_cellmax[c], cps * cps * cps);
        }
        int32_t *d = nullptr;
        if (hipMalloc((void **)&d, sizeof(int32_t) * std::max<size_t>(m.size(), 1)) != hipSuccess)
            return ecmc_fail(ECMC_ERR_HIP, "hipMalloc failed");
        if (!m.empty()) hipMemcpy(d, m.data(), sizeof(int32_t) * m.size(), hipMemcpyHostToDevice);
        b->d_order.push_back(d);
        int32_t *state = nullptr;
        if (hipMalloc((void **)&state, sizeof(int32_t) * std::max<size_t>(m.size(), 1)) != hipSuccess)
            return ecmc_fail(ECMC_ERR_HIP, "hipMalloc failed");
        b->d_state.push_back(state);
        hipStream_t s;
        hipStreamCreate(&s);
        b->streams.push_back(s);
    }
    for (int c = 0; c <= n_classes; c++) {
        hipEvent_t e;
        hipEventCreateWithFlags(&e, hipEventDisableTiming);
        b->fork.push_back(e);
    }
    for (int c = 0; c < n_classes; c++) {
        hipEvent_t e0, e1;
        hipEventCreate(&e0);
        hipEventCreate(&e1);
        b->cls_ev0.push_back(e0);
        b->cls_ev1.push_back(e1);
    }
    b->cls_ms.assign(n_classes, 0.0);
    b->cls_launches.assign(n_classes, 0);
    return ECMC_OK;
}

extern "C" void ecmc_batch_destroy(ecmc_batch *b) {
    if (!b) return;
    if (b->graph_exec) hipGraphExecDestroy(b->graph_exec);
    if (b->graph) hipGraphDestroy(b->graph);
    for (auto p : b->d_order) hipFree(p);
    for (auto p : b->d_state) hipFree(p);
    for (auto s : b->streams) hipStreamDestroy(s);
    for (auto e : b->fork) hipEventDestroy(e);
    for (auto e : b->cls_ev0) hipEventDestroy(e);
    for (auto e : b->cls_ev1) hipEventDestroy(e);
    hipFree(b->d_rep); hipFree(b->d_pos); hipFree(b->d_bounds); hipFree(b->d_queue); hipFree(b->d_log);
    hipFree(b->d_audit);
    if (b->ev0) hipEventDestroy(b->ev0);
    if (b->ev1) hipEventDestroy(b->ev1);
    if (b->main_stream) hipStreamDestroy(b->main_stream);
    delete b;
}

extern "C" int ecmc_batch_set_mode(ecmc_batch *b, int32_t mode) {
    if (!b || mode < 0 || mode > 2) return ecmc_fail(ECMC_ERR_ARGUMENT, "mode must be 0, 1 or 2");
    b->mode = mode;
    if (b->graph_exec) { hipGraphExecDestroy(b->graph_exec); b->graph_exec = nullptr; }
    if (b->graph) { hipGraphDestroy(b->graph); b->graph = nullptr; }
    return ECMC_OK;
}

extern "C" int ecmc_batch_set_log(ecmc_batch *b, int32_t legs_per_replica) {
    if (!b || legs_per_replica < 0) return ecmc_fail(ECMC_ERR_ARGUMENT, "bad argument");
    hipFree(b->d_log);
    b->d_log = nullptr;
    b->log_legs = legs_per_replica;
    if (legs_per_replica > 0 &&
        hipMalloc((void **)&b->d_log, sizeof(ecmc_batch_event) * (size_t)legs_per_replica * b->n_rep) != hipSuccess)
        return ecmc_fail(ECMC_ERR_HIP, "hipMalloc of the log failed");
    if (b->graph_exec) { hipGraphExecDestroy(b->graph_exec); b->graph_exec = nullptr; }
    if (b->graph) { hipGraphDestroy(b->graph); b->graph = nullptr; }
    return ECMC_OK;
}

extern "C" int ecmc_batch_run(ecmc_batch *b, int32_t legs_per_replica) {
    if (!b || legs_per_replica < 0) return ecmc_fail(ECMC_ERR_ARGUMENT, "bad argument");
    hipEventRecord(b->ev0, b->main_stream);
    int rc = enqueue_step(b, legs_per_replica, b->main_stream, b->class_timing);
    if (rc) return rc;
    hipEventRecord(b->ev1, b->main_stream);
    if (hipStreamSynchronize(b->main_stream) != hipSuccess) return ecmc_fail(ECMC_ERR_HIP, "batched kernel failed");
    std::fill(b->cls_ms.begin(), b->cls_ms.end(), 0.0);
    std::fill(b->cls_launches.begin(), b->cls_launches.end(), 0);
    return b->class_timing ? add_class_times(b) : ECMC_OK;
}

// Capture one step (queue reset + every class) into a hipGraph, then replay it n_steps times.
extern "C" int ecmc_batch_run_graph(ecmc_batch *b, int32_t legs_per_replica, int32_t n_steps) {
    if (!b || legs_per_replica < 0 || n_steps < 0) return ecmc_fail(ECMC_ERR_ARGUMENT, "bad argument");
    if (!b->graph_exec || b->graph_legs != legs_per_replica) {
        if (b->graph_exec) { hipGraphExecDestroy(b->graph_exec); b->graph_exec = nullptr; }
        if (b->graph) { hipGraphDestroy(b->graph); b->graph = nullptr; }
        if (hipStreamBeginCapture(b->main_stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
            return ecmc_fail(ECMC_ERR_HIP, "stream capture failed");
        int rc = enqueue_step(b, legs_per_replica, b->main_stream, false);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(b->main_stream, &g);
        if (rc) { if (g) hipGraphDestroy(g); return rc; }
        if (e != hipSuccess) return ecmc_fail(ECMC_ERR_HIP, "stream capture failed");
        b->graph = g;
        if (hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0) != hipSuccess)
            return ecmc_fail(ECMC_ERR_HIP, "graph instantiation failed");
        b->graph_legs = legs_per_replica;
    }
    std::fill(b->cls_ms.begin(), b->cls_ms.end(), 0.0);
    std::fill(b->cls_launches.begin(), b->cls_launches.end(), 0);   // replays are not class-timed
    hipEventRecord(b->ev0, b->main_stream);
    for (int s = 0; s < n_steps; s++)
        if (hipGraphLaunch(b->graph_exec, b->main_stream) != hipSuccess) return ecmc_fail(ECMC_ERR_HIP, "graph launch failed");
    hipEventRecord(b->ev1, b->main_stream);
    if (hipStreamSynchronize(b->main_stream) != hipSuccess) return ecmc_fail(ECMC_ERR_HIP, "batched graph failed");
    return ECMC_OK;
}

extern "C" int ecmc_batch_set_max_blocks(ecmc_batch *b, int32_t max_blocks) {
    if (!b || max_blocks < 0) return ecmc_fail(ECMC_ERR_ARGUMENT, "bad argument");
    b->max_blocks = max_blocks;
    if (b->graph_exec) { hipGraphExecDestroy(b->graph_exec); b->graph_exec = nullptr; }
    if (b->graph) { hipGraphDestroy(b->graph); b->graph = nullptr; }
    return ECMC_OK;
}

extern "C" int ecmc_batch_set_class_timing(ecmc_batch *b, int32_t on) {
    if (!b) return ecmc_fail(ECMC_ERR_ARGUMENT, "null batch");
    b->class_timing = on != 0;
    return ECMC_OK;
}

extern "C" int ecmc_batch_class_ms(const ecmc_batch *b, int32_t cls, double *ms, int32_t *launches) {
    if (!b || !ms || !launches) return ecmc_fail(ECMC_ERR_ARGUMENT, "null argument");
    if (cls < 0 || cls >= b->n_classes) return ecmc_fail(ECMC_ERR_ARGUMENT, "no such class");
    *ms = b->cls_ms[cls];
    *launches = b->cls_launches[cls];
    return ECMC_OK;
}

extern "C" int ecmc_batch_last_ms(const ecmc_batch *b, double *ms) {
    if (!b || !ms) return ecmc_fail(ECMC_ERR_ARGUMENT, "null argument");
    float f = 0.f;
    if (hipEventElapsedTime(&f, b->ev0, b->ev1) != hipSuccess) return ecmc_fail(ECMC_ERR_HIP, "event timing failed");
    *ms = f;
    return ECMC_OK;
}

static int pull_replicas(const ecmc_batch *b, std::vector<ReplicaDev> &rep) {
    rep.resize(b->n_rep);
    if (hipMemcpy(rep.data(), b->d_rep, sizeof(ReplicaDev) * b->n_rep, hipMemcpyDeviceToHost) != hipSuccess)
        return ecmc_fail(ECMC_ERR_HIP, "hipMemcpy failed");
    for (int r = 0; r < b->n_rep; r++)
        if (rep[r].err == 3) return ecmc_fail(ECMC_ERR_STATE, "a replica's chunk was held by two workgroups");
        else if (rep[r].err) return ecmc_fail(ECMC_ERR_STATE, "a unit was missing from its cell's occupant list");

    return ECMC_OK;
}

extern "C" int ecmc_batch_states(const ecmc_batch *b, int32_t *active, int32_t *direction, double *chain_left,
                                 uint64_t *leg, double *est8, int64_t *stats4) {
    if (!b) return ecmc_fail(ECMC_ERR_ARGUMENT, "null batch");
    std::vector<ReplicaDev> rep;
    int rc = pull_replicas(b, rep);
    if (rc) return rc;
    for (int r = 0; r < b->n_rep; r++) {
        if (active) active[r] = rep[r].active;
        if (direction) direction[r] = rep[r].direction;
        if (chain_left) chain_left[r] = rep[r].chain_left;
        if (leg) leg[r] = rep[r].leg;
        if (est8) memcpy(est8 + 8 * (size_t)r, rep[r].est, sizeof(double) * 8);
        if (stats4) memcpy(stats4 + 4 * (size_t)r, rep[r].stats, sizeof(int64_t) * 4);
    }
    return ECMC_OK;
}

extern "C" int ecmc_batch_positions(const ecmc_batch *b, double *positions) {
    if (!b || !positions) return ecmc_fail(ECMC_ERR_ARGUMENT, "null argument");
    std::vector<double> soa((size_t)3 * b->total_units);
    if (hipMemcpy(soa.data(), b->d_pos, sizeof(double) * soa.size(), hipMemcpyDeviceToHost) != hipSuccess)
        return ecmc_fail(ECMC_ERR_HIP, "hipMemcpy failed");
    double *o = positions;
    for (int r = 0; r < b->n_rep; r++) {
        const ReplicaDev &R = b->rep_host[r];
        for (int i = 0; i < R.n; i++)
            for (int q = 0; q < 3; q++) *o++ = soa[(size_t)R.off + (size_t)q * R.n + i];
    }
    return ECMC_OK;
}

extern "C" int ecmc_batch_log(const ecmc_batch *b, ecmc_batch_event *host_log) {
    if (!b || !host_log || !b->d_log) return ecmc_fail(ECMC_ERR_ARGUMENT, "no log configured");
    if (hipMemcpy(host_log, b->d_log, sizeof(ecmc_batch_event) * (size_t)b->log_legs * b->n_rep,
                  hipMemcpyDeviceToHost) != hipSuccess)
        return ecmc_fail(ECMC_ERR_HIP, "hipMemcpy failed");
    return ECMC_OK;
}

// Snapshot layout: header, the replicas' device records (state, Philox leg counters,
// estimators, statistics), the SoA positions.
struct BatchSnapshotHeader {
    uint64_t magic;
    int32_t n_rep, version;
    int64_t total_units;
    ecmc_batch_desc desc;
};
static const uint64_t SNAPSHOT_MAGIC = 0x45434d4342415443ull;   // "ECMCBATC"

static size_t snapshot_bytes(const ecmc_batch *b) {
    return sizeof(BatchSnapshotHeader) + sizeof(ReplicaDev) * (size_t)b->n_rep + sizeof(double) * 3 * (size_t)b->total_units;
}

extern "C" int ecmc_batch_snapshot(const ecmc_batch *b, void *buffer, int64_t *bytes) {
    if (!b || !bytes) return ecmc_fail(ECMC_ERR_ARGUMENT, "null argument");
    const size_t need = snapshot_bytes(b);
    if (!buffer) {
        *bytes = (int64_t)need;
        return ECMC_OK;
    }
    if (*bytes < (int64_t)need) return ecmc_fail(ECMC_ERR_ARGUMENT, "snapshot buffer too small");
    if (hipDeviceSynchronize() != hipSuccess) return ecmc_fail(ECMC_ERR_HIP, "device synchronisation failed");
    BatchSnapshotHeader h;
    memset(&h, 0, sizeof(h));
    h.magic = SNAPSHOT_MAGIC;
    h.n_rep = b->n_rep;
    h.version = 1;
    h.total_units = b->total_units;
    h.desc = b->desc;
    char *o = (char *)buffer;
    memcpy(o, &h, sizeof(h));
    o += sizeof(h);
    if (hipMemcpy(o, b->d_rep, sizeof(ReplicaDev) * b->n_rep, hipMemcpyDeviceToHost) != hipSuccess)
        return ecmc_fail(ECMC_ERR_HIP, "hipMemcpy failed");
    o += sizeof(ReplicaDev) * b->n_rep;
    if (hipMemcpy(o, b->d_pos, sizeof(double) * 3 * b->total_units, hipMemcpyDeviceToHost) != hipSuccess)
        return ecmc_fail(ECMC_ERR_HIP, "hipMemcpy failed");
    *bytes = (int64_t)need;
    return ECMC_OK;
}

extern "C" int ecmc_batch_restore(ecmc_batch *b, const void *buffer, int64_t bytes) {
    if (!b || !buffer) return ecmc_fail(ECMC_ERR_ARGUMENT, "null argument");
    if (bytes != (int64_t)snapshot_bytes(b)) return ecmc_fail(ECMC_ERR_ARGUMENT, "snapshot size does not match the batch");
    BatchSnapshotHeader h;
    memcpy(&h, buffer, sizeof(h));
    if (h.magic != SNAPSHOT_MAGIC || h.version != 1) return ecmc_fail(ECMC_ERR_ARGUMENT, "not a batch snapshot");
    if (h.n_rep != b->n_rep || h.total_units != b->total_units || memcmp(&h.desc, &b->desc, sizeof(h.desc)) != 0)
        return ecmc_fail(ECMC_ERR_CONFIGURATION, "the snapshot belongs to another batch configuration");
    const char *in = (const char *)buffer + sizeof(h);
    std::vector<ReplicaDev> rep((size_t)b->n_rep);
    memcpy(rep.data(), in, sizeof(ReplicaDev) * b->n_rep);
    for (int r = 0; r < b->n_rep; r++) {
        const ReplicaDev &x = rep[r], &y = b->rep_host[r];
        if (x.n != y.n || x.off != y.off || x.L != y.L || x.cps != y.cps || x.bound_off != y.bound_off)
            return ecmc_fail(ECMC_ERR_CONFIGURATION, "the snapshot's replica sizes or boxes differ from the batch");
    }
    in += sizeof(ReplicaDev) * b->n_rep;
    if (hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(b->d_rep, rep.data(), sizeof(ReplicaDev) * b->n_rep, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->d_pos, in, sizeof(double) * 3 * b->total_units, hipMemcpyHostToDevice) != hipSuccess)
        return ecmc_fail(ECMC_ERR_HIP, "hipMemcpy failed");
    return ECMC_OK;
}

extern "C" int ecmc_batch_audit(const ecmc_batch *b, uint64_t *violations) {
    if (!b || !violations) return ecmc_fail(ECMC_ERR_ARGUMENT, "null argument");
    unsigned long long v = 0;
    if (hipMemcpy(&v, b->d_audit, sizeof(v), hipMemcpyDeviceToHost) != hipSuccess) return ecmc_fail(ECMC_ERR_HIP, "hipMemcpy failed");
    *violations = v;
    return ECMC_OK;
}
