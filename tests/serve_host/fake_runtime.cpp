// A stand-in for the HIP runtime and the engine entry points that the serving loop
// (lc2registration_amd/csrc/lc2e_serve.cpp) calls, so that its scheduling can be tested on
// the host: "device" memory is host memory, copies are immediate, a launch computes fake
// results at once and its completion event reports "not ready" for a set number of
// queries. Test infrastructure only; no GPU is touched.
#include <hip/hip_runtime_api.h>

#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstdint>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "lc2e.h"

struct lc2e_engine {
    int dummy = 0;
};
struct lc2e_pair {
    int id = 0;
    unsigned long long gen = 1;  // launch-state generation (graphs)
    int mpp = 1;               // matrices per pose
    double share = 0.1;        // device share per pose
    int fail_code = 0;         // < 0: every launch of this pair fails with it
};

namespace {

std::mutex g_mu;
thread_local std::string g_err;
size_t g_budget = 0, g_used = 0;  // bytes, 0 = unlimited
int g_query_delay = 0;            // "not ready" answers before an event completes
int g_budget_poses = 0;           // > 0: launches of more poses fail with LC2E_ERR_BUDGET
int g_host_fail_at = -1;          // >= 0: that many pinned allocations succeed, the next fails
struct Launch {
    int pair, n;
    std::vector<float> first;  // first float of each pose's matrices
};
std::vector<Launch> g_log;
int g_in_flight = 0, g_max_in_flight = 0;

struct FakeEvent {
    std::atomic<int> left{0};
    bool pending = false;  // recorded, not yet seen complete (guarded by g_mu)
    hipStream_t stream = nullptr;
};
std::vector<FakeEvent*> g_events;

// Stream capture: operations enqueued on a capturing stream are recorded instead of run,
// and a graph launch runs them again in order.
struct FakeGraph {
    std::vector<std::function<void()>> ops;
};
std::map<void*, FakeGraph*> g_capture;  // stream -> graph being captured (guarded by g_mu)
int g_graph_replays = 0;

FakeGraph* capturing(void* stream) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_capture.find(stream);
    return it == g_capture.end() ? nullptr : it->second;
}

}  // namespace

extern "C" {

// --- test control ----------------------------------------------------------------------
lc2e_pair* fake_pair(int id, int mpp, double share, int fail_code) {
    auto* p = new lc2e_pair();
    p->id = id;
    p->mpp = mpp;
    p->share = share;
    p->fail_code = fail_code;
    return p;
}
void fake_free_pair(lc2e_pair* p) { delete p; }
lc2e_engine* fake_engine() { return new lc2e_engine(); }
void fake_free_engine(lc2e_engine* e) { delete e; }
void fake_config(int query_delay, long long budget_bytes, int budget_poses) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_query_delay = query_delay;
    g_budget = (size_t)budget_bytes;
    g_budget_poses = budget_poses;
    g_log.clear();
    g_in_flight = g_max_in_flight = 0;
    g_graph_replays = 0;
}
int fake_launches() {
    std::lock_guard<std::mutex> lk(g_mu);
    return (int)g_log.size();
}
// launch k: pair id, pose count, and up to `cap` first floats of its poses
int fake_launch(int k, int* pair, int* n, float* first, int cap) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (k < 0 || k >= (int)g_log.size()) return -1;
    *pair = g_log[k].pair;
    *n = g_log[k].n;
    for (int i = 0; i < g_log[k].n && i < cap; ++i) first[i] = g_log[k].first[i];
    return 0;
}
int fake_graph_replays() {
    std::lock_guard<std::mutex> lk(g_mu);
    return g_graph_replays;
}
void fake_bump_generation(lc2e_pair* p) { ++p->gen; }
int fake_max_in_flight() {
    std::lock_guard<std::mutex> lk(g_mu);
    return g_max_in_flight;
}
void fake_fail_host_malloc(int after) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_host_fail_at = after;
}
long long fake_used() {
    std::lock_guard<std::mutex> lk(g_mu);
    return (long long)g_used;
}

// --- engine entry points the serving loop uses -----------------------------------------
const char* lc2e_last_error(void) { return g_err.c_str(); }

// --- and those the registrar (lc2e_sessions.cpp) uses ----------------------------------
// A pair "created" from two fake volumes charges g_pair_bytes to the budget; its id is the
// creation count. Its matrices hold the pose parameters in their first floats.
namespace {
long long g_pair_bytes = 0;
int g_pairs_created = 0, g_pairs_live = 0, g_pairs_max_live = 0;
}
void fake_config_pairs(long long pair_bytes) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_pair_bytes = pair_bytes;
    g_pairs_created = g_pairs_live = g_pairs_max_live = 0;
}
int fake_pairs_max_live() {
    std::lock_guard<std::mutex> lk(g_mu);
    return g_pairs_max_live;
}
int fake_pairs_live() {
    std::lock_guard<std::mutex> lk(g_mu);
    return g_pairs_live;
}
int lc2e_pair_create(lc2e_engine*, lc2e_volume* fixed, lc2e_volume*, const float*, int, int, int,
                     int, lc2e_pair** out) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_budget && g_used + (size_t)g_pair_bytes > g_budget) {
        g_err = "fake HBM budget exceeded by a pair";
        return LC2E_ERR_BUDGET;
    }
    g_used += (size_t)g_pair_bytes;
    auto* p = new lc2e_pair();
    p->id = 1000 + ++g_pairs_created;
    // frames: the fixed volume's "address" encodes the frame count (fake volumes are tags)
    p->mpp = (int)(reinterpret_cast<uintptr_t>(fixed) & 0xff);
    if (p->mpp < 1) p->mpp = 1;
    p->share = 0.05;
    ++g_pairs_live;
    if (g_pairs_live > g_pairs_max_live) g_pairs_max_live = g_pairs_live;
    *out = p;
    return LC2E_OK;
}
void lc2e_pair_destroy(lc2e_pair* p) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(g_mu);
    g_used -= (size_t)g_pair_bytes;
    --g_pairs_live;
    delete p;
}
int lc2e_pair_set_per_frame(lc2e_pair* p, int on) {
    if (!on) p->mpp = 1;
    return LC2E_OK;
}
int lc2e_pose_to_matrix(lc2e_pair*, const double* params, int n, float out12[12]) {
    if (n < 1 || n > 12) {
        g_err = "fake pose_to_matrix: bad parameter count";
        return LC2E_ERR_INVALID;
    }
    for (int k = 0; k < 12; ++k) out12[k] = k < n ? (float)params[k] : 0.0f;
    return LC2E_OK;
}

int lc2e_pair_get_info(lc2e_pair* p, lc2e_pair_info* info) {
    std::memset(info, 0, sizeof(*info));
    info->per_frame_matrices = p->mpp > 1;
    info->n_frames = p->mpp;
    return LC2E_OK;
}

// Fake results per pose: sum of the pose's matrix floats as Σmeasure, 1 as Σw (0 if any
// float is NaN), LC2 = Σmeasure / Σw with NaN -> 0.
static void fake_compute(lc2e_pair* p, int n, const float* d_mats, double* d_out);

int lc2e_eval_batch_async(lc2e_pair* p, int n, const float* d_mats, double* d_out, void* stream) {
    if (p->fail_code < 0) {
        g_err = "fake launch failure of pair " + std::to_string(p->id);
        return p->fail_code;
    }
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (g_budget_poses > 0 && n > g_budget_poses) {
            g_err = "fake budget";
            return LC2E_ERR_BUDGET;
        }
    }
    if (FakeGraph* g = capturing(stream)) {
        g->ops.push_back([=] { fake_compute(p, n, d_mats, d_out); });
        return LC2E_OK;
    }
    fake_compute(p, n, d_mats, d_out);
    return LC2E_OK;
}

static void fake_compute(lc2e_pair* p, int n, const float* d_mats, double* d_out) {
    {
        std::lock_guard<std::mutex> lk(g_mu);
        Launch l{p->id, n, {}};
        for (int i = 0; i < n; ++i) l.first.push_back(d_mats[(size_t)i * p->mpp * 12]);
        g_log.push_back(l);
    }
    for (int i = 0; i < n; ++i) {
        double s = 0.0, w = 1.0;
        for (int k = 0; k < p->mpp * 12; ++k) {
            const float v = d_mats[(size_t)i * p->mpp * 12 + k];
            if (std::isnan(v)) w = 0.0;
            else s += v;
        }
        const double l = s / w;
        d_out[3 * i] = w > 0.0 ? s : 0.0;
        d_out[3 * i + 1] = w;
        d_out[3 * i + 2] = std::isfinite(l) ? l : 0.0;
    }
}

// --- HIP runtime subset ----------------------------------------------------------------
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned int) {
    *s = reinterpret_cast<hipStream_t>(new int(0));
    return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s) {
    delete reinterpret_cast<int*>(s);
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t s) {
    // everything on the stream is complete
    std::lock_guard<std::mutex> lk(g_mu);
    for (FakeEvent* ev : g_events)
        if (ev->stream == s && ev->pending) {
            ev->pending = false;
            ev->left = 0;
            --g_in_flight;
        }
    return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) {
    auto* ev = new FakeEvent();
    std::lock_guard<std::mutex> lk(g_mu);
    g_events.push_back(ev);
    *e = reinterpret_cast<hipEvent_t>(ev);
    return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t e) {
    auto* ev = reinterpret_cast<FakeEvent*>(e);
    std::lock_guard<std::mutex> lk(g_mu);
    for (size_t i = 0; i < g_events.size(); ++i)
        if (g_events[i] == ev) {
            g_events.erase(g_events.begin() + (long)i);
            break;
        }
    if (ev->pending) --g_in_flight;
    delete ev;
    return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s) {
    auto* ev = reinterpret_cast<FakeEvent*>(e);
    std::lock_guard<std::mutex> lk(g_mu);
    ev->left = g_query_delay;
    ev->stream = s;
    ev->pending = true;
    ++g_in_flight;
    if (g_in_flight > g_max_in_flight) g_max_in_flight = g_in_flight;
    return hipSuccess;
}
hipError_t hipEventQuery(hipEvent_t e) {
    auto* ev = reinterpret_cast<FakeEvent*>(e);
    if (ev->left.fetch_sub(1) > 0) return hipErrorNotReady;
    std::lock_guard<std::mutex> lk(g_mu);
    if (ev->pending) {
        ev->pending = false;
        --g_in_flight;
    }
    return hipSuccess;
}
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t bytes, hipMemcpyKind, hipStream_t s) {
    if (FakeGraph* g = capturing(s)) {
        g->ops.push_back([=] { std::memcpy(dst, src, bytes); });
        return hipSuccess;
    }
    std::memcpy(dst, src, bytes);
    return hipSuccess;
}
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_capture.count(s)) return hipErrorInvalidValue;
    g_capture[s] = new FakeGraph();
    return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* graph) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_capture.find(s);
    if (it == g_capture.end()) return hipErrorInvalidValue;
    *graph = reinterpret_cast<hipGraph_t>(it->second);
    g_capture.erase(it);
    return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* exec, hipGraph_t graph, hipGraphNode_t*, char*, size_t) {
    *exec = reinterpret_cast<hipGraphExec_t>(graph);
    return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t exec, hipStream_t) {
    auto* g = reinterpret_cast<FakeGraph*>(exec);
    for (auto& op : g->ops) op();
    std::lock_guard<std::mutex> lk(g_mu);
    ++g_graph_replays;
    return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t) { return hipSuccess; }
hipError_t hipGraphDestroy(hipGraph_t graph) {
    delete reinterpret_cast<FakeGraph*>(graph);
    return hipSuccess;
}
hipError_t hipHostMalloc(void** p, size_t bytes, unsigned int) {
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (g_host_fail_at >= 0 && g_host_fail_at-- == 0) {
            *p = nullptr;
            return hipErrorOutOfMemory;
        }
    }
    *p = std::malloc(bytes ? bytes : 1);
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipHostGetDevicePointer(void** dev, void* host, unsigned int) {
    *dev = host;  // one address space here
    return hipSuccess;
}
hipError_t hipHostFree(void* p) {
    std::free(p);
    return hipSuccess;
}
const char* hipGetErrorString(hipError_t) { return "fake error"; }

}  // extern "C"

namespace lc2e {
void set_last_error(const std::string& msg) { g_err = msg; }
int engine_alloc(lc2e_engine*, void** p, size_t bytes) {
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (g_budget && g_used + bytes > g_budget) {
            g_err = "fake HBM budget exceeded";
            return LC2E_ERR_BUDGET;
        }
        g_used += bytes;
    }
    *p = std::malloc(bytes ? bytes : 1);
    return LC2E_OK;
}
void engine_release(lc2e_engine*, void* p, size_t bytes) {
    std::free(p);
    std::lock_guard<std::mutex> lk(g_mu);
    g_used -= bytes;
}
int engine_device(lc2e_engine*) { return 0; }
double launch_share(lc2e_pair* p, int n) { return p->share * n; }
unsigned long long pair_generation(lc2e_pair* p) { return p->gen; }
int pair_prepare(lc2e_pair*, int) { return LC2E_OK; }
}  // namespace lc2e
