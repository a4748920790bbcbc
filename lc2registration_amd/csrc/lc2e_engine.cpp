// C ABI of the LC2 engine (include/lc2e.h): resident volumes, pairs, batched
// evaluation on one HIP stream, HBM budget accounting, optional HIP-graph replay.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "lc2e.h"
#include "lc2e_geometry.hpp"
#include "lc2e_kernels.hpp"

namespace {

thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

#define HIP_TRY(expr)                                                                      \
    do {                                                                                   \
        hipError_t _e = (expr);                                                            \
        if (_e != hipSuccess)                                                              \
            return fail(LC2E_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), \
                        __FILE__, __LINE__);                                               \
    } while (0)

int dtype_size(int dt) {
    switch (dt) {
        case LC2E_U8: return 1;
        case LC2E_I16: return 2;
        case LC2E_F16: return 2;
        case LC2E_F32: return 4;
        default: return 0;
    }
}

}  // namespace

struct lc2e_engine {
    int device = 0;
    int cus = 0;  // compute units of the device
    hipStream_t stream = nullptr;
    // the device's scratch array for lc2e_eval_batch_noalloc (null if its lookup failed,
    // then scratch_error says why)
    void* scratch = nullptr;
    std::string scratch_error;
    size_t budget = 0;  // 0 = unlimited
    size_t used = 0;
    std::mutex mu;      // guards `used`: a serving loop allocates from its own thread

    int alloc(void** p, size_t bytes) {
        {
            std::lock_guard<std::mutex> lk(mu);
            if (budget && used + bytes > budget)
                return fail(LC2E_ERR_BUDGET, "HBM budget exceeded: %zu used + %zu requested > %zu",
                            used, bytes, budget);
            used += bytes;  // reserved before the allocation, returned if it fails
        }
        hipError_t err = hipSetDevice(device);
        if (err == hipSuccess) err = hipMalloc(p, bytes ? bytes : 1);
        if (err != hipSuccess) {
            std::lock_guard<std::mutex> lk(mu);
            used -= bytes;
            return fail(LC2E_ERR_HIP, "hipMalloc of %zu bytes failed: %s", bytes, hipGetErrorString(err));
        }
        return LC2E_OK;
    }
    void release(void* p, size_t bytes) {
        if (p) {
            (void)hipFree(p);
            std::lock_guard<std::mutex> lk(mu);
            used -= bytes;
        }
    }
};

struct lc2e_volume {
    lc2e_engine* e = nullptr;
    void* d = nullptr;
    size_t bytes = 0;
    int f16 = 0;
    lc2e_geom g{};
};

struct GraphEntry {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

struct lc2e_pair {
    lc2e_engine* e = nullptr;
    lc2e_volume* fixed = nullptr;
    lc2e_volume* moving = nullptr;
    lc2e_pair_info info{};
    lc2e::VGeom v{};
    lc2e::SGeom sg{};
    float* stat_img = nullptr;
    float* stat_grad = nullptr;
    double* ustat = nullptr;  // [2][V]: patch mean and centred energy of a static ultrasound
    size_t vbytes = 0;
    // 3D z-chunk: forced for every launch (lc2e_pair_set_tuning), else the value tuned for
    // the launch's batch size (lc2e_pair_autotune / lc2e_pair_set_tuning_n), else a model
    int zchunk_forced = 0;
    std::map<int, int> tuned;
    std::map<int, int> model_cache;  // the launch model's z-chunk per batch size
    int last_zchunk = 0;
    // batch buffers
    int cap = 0;
    int part_bpp = 0;  // partial sums per pose the partials buffer holds
    float* d_mats = nullptr;
    double* d_partials = nullptr;
    float* h_mats = nullptr;   // pinned, mapped and coherent
    double* h_out = nullptr;   // pinned, mapped and coherent
    float* h_mats_dev = nullptr;   // their device addresses
    double* h_out_dev = nullptr;
    size_t cap_bytes = 0;
    bool use_graph = false;
    std::map<int, GraphEntry> graphs;
    // changes whenever captured launches of this pair go stale (buffers or decomposition
    // changed); unique across pairs, so a graph cached by address is never reused for
    // another pair
    unsigned long long gen = 0;
    // mutual information (patchSize 0): extremes of the static image, per-pose extremes,
    // joint histograms and results, sized for mi_cap poses of mi_bins counters
    int* mi_ext_stat = nullptr;
    bool mi_stat_ready = false;
    int mi_cap = 0;
    size_t mi_bins = 0;
    int* mi_ext = nullptr;
    unsigned* mi_hist = nullptr;
    double* mi_out = nullptr;
    int resident_per_cu = 0;  // fused work-groups one CU holds at once (0: not queried yet)
    // separable path (radius beyond the fused kernels): grids and box-sum buffers, made at
    // pair creation
    bool sep = false;
    float* sep_grids = nullptr;    // [2][V] f32: moving image, its |Sobel|
    double* sep_sums = nullptr;    // [2][channels][V] f64
    size_t sep_grid_bytes = 0, sep_sum_bytes = 0;
    // kernel timing (eager path)
    bool profile = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double prof_ms = 0.0;
    long long prof_launches = 0;

    int mats_per_pose() const { return info.per_frame_matrices ? info.n_frames : 1; }
};

namespace {

std::atomic<unsigned long long> g_generation{1};

void drop_graphs(lc2e_pair* p) {
    for (auto& kv : p->graphs) {
        if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
        if (kv.second.graph) (void)hipGraphDestroy(kv.second.graph);
    }
    p->graphs.clear();
    p->gen = g_generation.fetch_add(1);
}

// Work decomposition of one launch.
struct Tiling {
    int zchunk = 1, ntx = 0, nty = 0, nzc = 0;
    int tiles = 0;  // work-groups per pose
    int bpp = 0;    // partial sums per pose
};

int round_zchunk(int c) { return (c + lc2e::kZBlock - 1) / lc2e::kZBlock * lc2e::kZBlock; }

lc2e::FusedArgs make_args(lc2e_pair* p, const float* d_mats, double* d_partials, const Tiling& t);

// The pair's evaluation kernels for n poses: the fused kernel, or the separable path for
// radii beyond it.
hipError_t launch_eval(lc2e_pair* p, const lc2e::FusedArgs& a, int n, hipStream_t st) {
    if (p->sep)
        return lc2e::launch_sep(a, p->info.grid_ndim, p->info.src_ndim, p->info.high_res_is_moving, n,
                                st);
    return lc2e::launch_fused(a, p->info.grid_ndim, p->info.src_ndim, p->info.high_res_is_moving, n,
                              st);
}

// Fused work-groups one CU holds at once for this pair's kernel (queried once, then kept).
int resident_per_cu(lc2e_pair* p) {
    if (p->sep) return 1;  // the separable path's launches fill the device by themselves
    if (p->resident_per_cu <= 0) {
        const lc2e::FusedArgs a = make_args(p, nullptr, nullptr, Tiling{});
        int per_cu = 0;
        if (lc2e::fused_resident_per_cu(a, p->info.grid_ndim, p->info.src_ndim,
                                        p->info.high_res_is_moving, &per_cu) != hipSuccess ||
            per_cu <= 0)
            per_cu = 1;
        p->resident_per_cu = per_cu;
    }
    return p->resident_per_cu;
}

// z-chunk for n poses when nothing is tuned. A work-group of chunk c marches
// min(c, planes left) + 2R + 2 plane steps. Work-groups go to the XCDs round-robin by id
// and, within an XCD, to the first free slot in id order (map_block: tile-major, the poses
// of a tile consecutive; tiles ordered by z-chunk), each CU holding resident_per_cu of them.
// The model replays that dispatch with a min-heap of slot finishing times per XCD and takes
// the latest finish; launches of more than 32 rounds fall back to rounds x steps. Minimise
// over the chunk lengths (multiples of kZBlock; ties: longer chunks). Against exhaustive
// timing on MI355X (3D affine grid 192x192x160 and CT grid 256^3, 1 to 32 poses) it picks
// within 4% of the fastest chunk in all but one case (3.5%); a rounds-only model lost up to
// 16%. Cached per batch size.
int model_zchunk(lc2e_pair* p, int n) {
    auto hit = p->model_cache.find(n);
    if (hit != p->model_cache.end()) return hit->second;
    const int r = p->info.patch_radius;
    const int ez = p->v.Z - 2 * r;
    int ntx, nty, nzc;
    if (ez <= 0 || lc2e::fused_blocks_3d(p->v, r, ez, &ntx, &nty, &nzc) <= 0) return lc2e::kZBlock;
    const long long cols = (long long)ntx * nty;
    const int cus = p->e->cus > 0 ? p->e->cus : 256;
    const int per_xcd = std::max(1, cus / lc2e::kXcds * resident_per_cu(p));
    const int cmax = round_zchunk(ez);
    constexpr double kOverhead = 2.0;  // per work-group, in plane steps (fitted)
    int best = cmax;
    double best_cost = -1.0;
    std::vector<double> heap;
    for (int c = cmax; c >= lc2e::kZBlock; c -= lc2e::kZBlock) {
        const int chunks = (ez + c - 1) / c;
        const long long tiles = cols * chunks;
        const long long tiles8 = (tiles + lc2e::kXcds - 1) / lc2e::kXcds * lc2e::kXcds;
        const long long rounds = (tiles8 / lc2e::kXcds * n + per_xcd - 1) / per_xcd;
        double cost = 0.0;
        if (rounds > 32) {
            cost = (double)rounds * (c + 2 * r + 2 + kOverhead);
        } else {
            for (int x = 0; x < lc2e::kXcds; ++x) {
                heap.assign(per_xcd, 0.0);  // a min-heap of slot finishing times
                for (long long t = x; t < tiles8; t += lc2e::kXcds) {
                    const double len = t >= tiles ? 0.5
                        : (double)std::min<long long>(c, ez - (t / cols) * c) + 2 * r + 2 + kOverhead;
                    for (int k = 0; k < n; ++k) {
                        std::pop_heap(heap.begin(), heap.end(), std::greater<double>());
                        heap.back() += len;
                        std::push_heap(heap.begin(), heap.end(), std::greater<double>());
                    }
                }
                cost = std::max(cost, *std::max_element(heap.begin(), heap.end()));
            }
        }
        if (best_cost < 0.0 || cost < best_cost) {
            best_cost = cost;
            best = c;
        }
    }
    p->model_cache[n] = best;
    return best;
}

Tiling tiling_for(lc2e_pair* p, int n) {
    Tiling t;
    const int r = p->info.patch_radius;
    if (p->sep) {  // one partial per 256 output voxels
        t.tiles = t.bpp = lc2e::sep_blocks(p->v, r, p->info.grid_ndim);
        return t;
    }
    if (p->info.grid_ndim == 3) {
        auto it = p->tuned.find(n);
        t.zchunk = round_zchunk(p->zchunk_forced > 0 ? p->zchunk_forced
                                : it != p->tuned.end() ? it->second : model_zchunk(p, n));
        t.tiles = lc2e::fused_blocks_3d(p->v, r, t.zchunk, &t.ntx, &t.nty, &t.nzc);
        t.bpp = lc2e::fused_partials_3d(p->v, r);
        p->last_zchunk = t.zchunk;
    } else {
        // one work-group and one partial per tile and run of kFrameRun frames (fixed, so
        // that the summation order does not depend on the batch)
        const int tf = lc2e::fused_blocks_2d(p->v, r, &t.ntx, &t.nty);
        t.zchunk = p->v.Z < lc2e::kFrameRun ? p->v.Z : lc2e::kFrameRun;
        t.nzc = tf > 0 ? (p->v.Z + t.zchunk - 1) / t.zchunk : 0;
        t.tiles = t.bpp = t.ntx * t.nty * t.nzc;
    }
    return t;
}

void free_batch(lc2e_pair* p) {
    drop_graphs(p);
    lc2e_engine* e = p->e;
    if (p->d_mats) e->release(p->d_mats, (size_t)p->cap * p->mats_per_pose() * 12 * sizeof(float));
    if (p->d_partials) e->release(p->d_partials, (size_t)p->cap * p->part_bpp * 2 * sizeof(double));
    if (p->h_mats) (void)hipHostFree(p->h_mats);
    if (p->h_out) (void)hipHostFree(p->h_out);
    p->d_mats = nullptr;
    p->d_partials = nullptr;
    p->h_mats = nullptr;
    p->h_out = nullptr;
    p->h_mats_dev = nullptr;
    p->h_out_dev = nullptr;
    p->cap = 0;
    p->part_bpp = 0;
}

void free_mi(lc2e_pair* p) {
    lc2e_engine* e = p->e;
    if (p->mi_ext) e->release(p->mi_ext, (size_t)p->mi_cap * 2 * sizeof(int));
    if (p->mi_hist) e->release(p->mi_hist, (size_t)p->mi_cap * p->mi_bins * sizeof(unsigned));
    if (p->mi_out) e->release(p->mi_out, (size_t)p->mi_cap * sizeof(double));
    p->mi_ext = nullptr;
    p->mi_hist = nullptr;
    p->mi_out = nullptr;
    p->mi_cap = 0;
    p->mi_bins = 0;
}

// Batch buffers for n poses whose launches use up to bpp work-groups per pose.
int ensure_capacity(lc2e_pair* p, int n, int bpp) {
    if (bpp < 1) bpp = 1;
    if (n <= p->cap && bpp <= p->part_bpp) return LC2E_OK;
    int cap = p->cap > 1 ? p->cap : 1;
    while (cap < n) cap *= 2;
    if (bpp < p->part_bpp) bpp = p->part_bpp;
    free_batch(p);
    lc2e_engine* e = p->e;
    const int mpp = p->mats_per_pose();
    int rc;
    if ((rc = e->alloc((void**)&p->d_mats, (size_t)cap * mpp * 12 * sizeof(float))) < 0) return rc;
    if ((rc = e->alloc((void**)&p->d_partials, (size_t)cap * bpp * 2 * sizeof(double))) < 0) {
        e->release(p->d_mats, (size_t)cap * mpp * 12 * sizeof(float));
        p->d_mats = nullptr;
        return rc;
    }
    p->cap = cap;  // free_batch returns the device bytes by these sizes
    p->part_bpp = bpp;
    void* hm = nullptr;
    void* ho = nullptr;
    // mapped and coherent: small launches read their matrices from here and every launch's
    // reduction writes its results here, with no copy command (enqueue_batch)
    const unsigned fl = hipHostMallocMapped | hipHostMallocCoherent;
    hipError_t he = hipHostMalloc(&hm, (size_t)cap * mpp * 12 * sizeof(float), fl);
    if (he == hipSuccess) {
        p->h_mats = (float*)hm;
        he = hipHostMalloc(&ho, (size_t)cap * 3 * sizeof(double), fl);
        if (he == hipSuccess) p->h_out = (double*)ho;
    }
    void* dm = nullptr;
    void* dout = nullptr;
    if (he == hipSuccess) he = hipHostGetDevicePointer(&dm, hm, 0);
    if (he == hipSuccess) he = hipHostGetDevicePointer(&dout, ho, 0);
    if (he == hipSuccess) {
        p->h_mats_dev = (float*)dm;
        p->h_out_dev = (double*)dout;
    }
    if (he != hipSuccess) {
        free_batch(p);  // no half-made buffers: the next call allocates again
        return fail(LC2E_ERR_HIP, "pinned staging allocation of %d poses failed: %s", cap,
                    hipGetErrorString(he));
    }
    return LC2E_OK;
}

lc2e::FusedArgs make_args(lc2e_pair* p, const float* d_mats, double* d_partials, const Tiling& t) {
    lc2e::FusedArgs a{};
    a.v = p->v;
    a.s = p->sg;
    a.src = p->moving->d;
    a.src_f16 = p->moving->f16;
    a.stat_img = p->stat_img;
    a.stat_grad = p->stat_grad;
    const size_t V = p->vbytes / sizeof(float);
    a.ustat_mean = p->ustat;
    a.ustat_cyy = p->ustat ? p->ustat + V : nullptr;
    a.mats = d_mats;
    a.partials = d_partials;
    a.wmap = nullptr;
    a.mmap = nullptr;
    a.r = p->info.patch_radius;
    a.zchunk = t.zchunk;
    a.ntx = t.ntx;
    a.nty = t.nty;
    a.nzc = t.nzc;
    a.blocks_per_pose = t.bpp;
    a.tiles = t.tiles;
    a.mats_per_pose = p->mats_per_pose();
    if (p->sep) {
        const size_t V = p->vbytes / sizeof(float);
        a.sep_img = p->sep_grids;
        a.sep_grad = p->sep_grids + V;
        a.sep_a = p->sep_sums;
        a.sep_b = p->sep_sums + (size_t)lc2e::kSepChannels(p->info.high_res_is_moving != 0) * V;
    }
    return a;
}

// Enqueue: the staged matrices to the device, the fused kernel, the reduction into the
// pinned result buffer. Each launch command costs several microseconds, so small launches
// (up to kDirectPoses poses of one matrix each, on a single volume or image: each
// work-group reads its pose's 12 floats once) read the pinned matrices in place instead of
// copying them, and the reduction always writes the results in place: one or two commands
// per batch instead of four.
constexpr int kDirectPoses = 8;
int enqueue_batch(lc2e_pair* p, int n, const Tiling& t, hipStream_t st) {
    const int mpp = p->mats_per_pose();
    const bool direct = !p->sep && mpp == 1 && n <= kDirectPoses &&
                        (p->info.grid_ndim == 3 || p->v.Z == 1);
    if (!direct)
        HIP_TRY(hipMemcpyAsync(p->d_mats, p->h_mats, (size_t)n * mpp * 12 * sizeof(float),
                               hipMemcpyHostToDevice, st));
    if (t.tiles > 0) {
        lc2e::FusedArgs a = make_args(p, direct ? p->h_mats_dev : p->d_mats, p->d_partials, t);
        // timed launches are eager (run_batch), the events bracket the fused kernel
        const bool timed = p->profile;
        if (timed) HIP_TRY(hipEventRecord(p->ev0, st));
        HIP_TRY(launch_eval(p, a, n, st));
        if (timed) HIP_TRY(hipEventRecord(p->ev1, st));
        HIP_TRY(lc2e::launch_reduce(p->d_partials, t.bpp, n, p->h_out_dev, st));
    } else {
        // no patch centre fits in the grid: every sum is zero
        HIP_TRY(lc2e::launch_reduce(p->d_partials, 0, n, p->h_out_dev, st));
    }
    return LC2E_OK;
}

// Runs a batch of n poses staged in h_mats; the buffers must already fit tiling t.
int run_batch(lc2e_pair* p, int n, const Tiling& t) {
    hipStream_t st = p->e->stream;
    if (p->use_graph && !p->profile) {
        auto it = p->graphs.find(n);
        if (it == p->graphs.end()) {
            GraphEntry g;
            HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            int rc = enqueue_batch(p, n, t, st);
            hipGraph_t graph = nullptr;
            hipError_t ce = hipStreamEndCapture(st, &graph);
            if (rc >= 0 && ce == hipSuccess) ce = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
            if (rc < 0 || ce != hipSuccess) {
                if (graph) (void)hipGraphDestroy(graph);
                if (rc < 0) return rc;
                return fail(LC2E_ERR_HIP, "graph capture failed: %s", hipGetErrorString(ce));
            }
            g.graph = graph;
            it = p->graphs.emplace(n, g).first;
        }
        HIP_TRY(hipGraphLaunch(it->second.exec, st));
    } else {
        int rc = enqueue_batch(p, n, t, st);
        if (rc < 0) return rc;
    }
    HIP_TRY(hipStreamSynchronize(st));
    if (p->profile && t.tiles > 0) {
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, p->ev0, p->ev1));
        p->prof_ms += ms;
        p->prof_launches += 1;
    }
    return LC2E_OK;
}

}  // namespace

namespace lc2e {
// Sets the calling thread's lc2e_last_error() text (the serving loop hands a worker
// thread's message to the client that collects the request).
void set_last_error(const std::string& msg) { g_last_error = msg; }
// Device memory charged to the engine's HBM budget, for the serving loop's launch slots.
int engine_alloc(lc2e_engine* e, void** p, size_t bytes) { return e->alloc(p, bytes); }
void engine_release(lc2e_engine* e, void* p, size_t bytes) { e->release(p, bytes); }
int engine_device(lc2e_engine* e) { return e->device; }
// The share of the device one fused launch of n poses occupies: its work-groups over the
// work-groups all CUs hold at once (>= 1: it fills the GPU by itself). The serving loop
// runs launches side by side only while their shares add up to at most 1.
// Generation of the pair's launch state: graphs captured under another generation are stale.
unsigned long long pair_generation(lc2e_pair* p) { return p->gen; }
// Make the pair's batch buffers fit n poses now, so that a following
// lc2e_eval_batch_async of n poses allocates nothing (e.g. under stream capture).
int pair_prepare(lc2e_pair* p, int n) {
    const Tiling t = tiling_for(p, n);
    return ensure_capacity(p, n, t.bpp);
}
double launch_share(lc2e_pair* p, int n) {
    const Tiling t = tiling_for(p, n);
    if (t.tiles <= 0 || n <= 0) return 0.0;
    const int per_cu = resident_per_cu(p);
    const int prun = p->sep ? 1 : pose_run(p->v, p->info.grid_ndim, p->info.src_ndim,
                                           p->mats_per_pose(), t.tiles, n);
    const double wgs = (double)((t.tiles + kXcds - 1) / kXcds * kXcds) * ((n + prun - 1) / prun);
    const int cus = p->e->cus > 0 ? p->e->cus : 256;
    return wgs / ((double)cus * per_cu);
}
}  // namespace lc2e

extern "C" {

const char* lc2e_last_error(void) { return g_last_error.c_str(); }

int lc2e_create(int device, size_t hbm_budget_bytes, lc2e_engine** out) {
    if (!out) return fail(LC2E_ERR_INVALID, "null output handle");
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev)
        return fail(LC2E_ERR_INVALID, "device %d out of range (%d devices)", device, ndev);
    auto* e = new lc2e_engine();
    e->device = device;
    e->budget = hbm_budget_bytes;
    hipError_t err = hipSetDevice(device);
    if (err == hipSuccess)
        err = hipDeviceGetAttribute(&e->cus, hipDeviceAttributeMultiprocessorCount, device);
    if (err == hipSuccess) err = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
    if (err != hipSuccess) {
        delete e;
        return fail(LC2E_ERR_HIP, "stream creation failed: %s", hipGetErrorString(err));
    }
    // the allocation-free entry point's device scratch: its address on this device, looked
    // up once per device (with the device current) and kept
    const hipError_t se = lc2e::scratch_address(device, &e->scratch);
    if (se != hipSuccess) {
        e->scratch = nullptr;
        e->scratch_error = hipGetErrorString(se);
    }
    *out = e;
    return LC2E_OK;
}

void lc2e_destroy(lc2e_engine* e) {
    if (!e) return;
    (void)hipSetDevice(e->device);
    (void)hipStreamSynchronize(e->stream);
    (void)hipStreamDestroy(e->stream);
    delete e;
}

int lc2e_memory(lc2e_engine* e, size_t* used, size_t* budget) {
    if (!e) return fail(LC2E_ERR_INVALID, "null engine");
    std::lock_guard<std::mutex> lk(e->mu);
    if (used) *used = e->used;
    if (budget) *budget = e->budget;
    return LC2E_OK;
}

int lc2e_volume_upload(lc2e_engine* e, const void* host, lc2e_dtype dtype, const lc2e_geom* geom,
                       lc2e_dtype storage, lc2e_volume** out) {
    if (!e || !host || !geom || !out) return fail(LC2E_ERR_INVALID, "null argument");
    const int dsz = dtype_size(dtype);
    if (!dsz) return fail(LC2E_ERR_INVALID, "unknown dtype %d", (int)dtype);
    for (int a = 0; a < 3; ++a) {
        if (geom->dims[a] <= 0) return fail(LC2E_ERR_INVALID, "dims must be positive");
        // the samplers form row and voxel offsets with 24-bit multiplies
        if (geom->dims[a] >= (1 << 24))
            return fail(LC2E_ERR_UNSUPPORTED, "dimension %d exceeds 2^24 - 1", geom->dims[a]);
        if (!(geom->spacing[a] > 0.0f)) return fail(LC2E_ERR_INVALID, "spacing must be positive");
    }
    const long long n = (long long)geom->dims[0] * geom->dims[1] * geom->dims[2];
    if (n >= (1LL << 31)) return fail(LC2E_ERR_UNSUPPORTED, "volume exceeds 2^31 voxels");
    int f16;
    if (storage == LC2E_STORE_AUTO) {
        f16 = (dtype == LC2E_U8 || dtype == LC2E_F16) ? 1 : 0;
        if (dtype == LC2E_I16) {
            // every integer in [-2048, 2048] is exact in half precision (11-bit significand),
            // which covers CT in Hounsfield units: half storage is then lossless
            const short* h = static_cast<const short*>(host);
            f16 = 1;
            for (long long i = 0; i < n && f16; ++i)
                if (h[i] < -2048 || h[i] > 2048) f16 = 0;
        }
    }
    else if (storage == LC2E_F16) f16 = 1;
    else if (storage == LC2E_F32) f16 = 0;
    else return fail(LC2E_ERR_INVALID, "storage must be F16, F32 or AUTO");
    // the samplers address a volume and its zeroed tail with 32-bit byte offsets
    const long long pad = lc2e::volume_pad_elems(geom->dims[0], geom->dims[1]);
    if ((n + pad) * (f16 ? 2 : 4) > 0xffffffffLL)
        return fail(LC2E_ERR_UNSUPPORTED,
                    "volume of %lld voxels (plus a %lld-voxel zeroed tail) exceeds the 4 GiB "
                    "that %s storage can address%s", n, pad, f16 ? "f16" : "f32",
                    f16 ? "" : "; store it as F16");
    // The samplers rely on finite voxels (a zero-weight corner must contribute exactly 0).
    if (dtype == LC2E_F32) {
        const float lim = f16 ? 65504.0f : 3.4028235e38f;
        const float* h = static_cast<const float*>(host);
        for (long long i = 0; i < n; ++i)
            if (!(std::fabs(h[i]) <= lim))
                return fail(LC2E_ERR_INVALID, "value %g at %lld is not finite%s", h[i], i,
                            f16 ? " in half storage" : "");
    } else if (dtype == LC2E_F16) {
        const unsigned short* h = static_cast<const unsigned short*>(host);
        for (long long i = 0; i < n; ++i)
            if ((h[i] & 0x7c00u) == 0x7c00u)
                return fail(LC2E_ERR_INVALID, "half value 0x%04x at %lld is not finite", h[i], i);
    }
    auto* v = new lc2e_volume();
    v->e = e;
    v->g = *geom;
    v->f16 = f16;
    const size_t esz = f16 ? 2 : 4;
    v->bytes = (size_t)(n + pad) * esz;
    int rc = e->alloc(&v->d, v->bytes);
    if (rc < 0) {
        delete v;
        return rc;
    }
    {
        // zeroed tail (lc2e::volume_pad_elems): the samplers' reads past the last voxel
        hipError_t pe = hipMemsetAsync(static_cast<char*>(v->d) + (size_t)n * esz, 0,
                                       (size_t)pad * esz, e->stream);
        if (pe != hipSuccess) {
            e->release(v->d, v->bytes);
            delete v;
            return fail(LC2E_ERR_HIP, "volume padding failed: %s", hipGetErrorString(pe));
        }
    }
    // stage the raw bytes on the device, convert there
    void* staging = nullptr;
    hipError_t err = hipMalloc(&staging, (size_t)n * dsz);
    if (err == hipSuccess)
        err = hipMemcpyAsync(staging, host, (size_t)n * dsz, hipMemcpyHostToDevice, e->stream);
    if (err == hipSuccess) err = lc2e::launch_convert(staging, (int)dtype, v->d, f16, n, e->stream);
    if (err == hipSuccess) err = hipStreamSynchronize(e->stream);
    if (staging) (void)hipFree(staging);
    if (err != hipSuccess) {
        e->release(v->d, v->bytes);
        delete v;
        return fail(LC2E_ERR_HIP, "volume upload failed: %s", hipGetErrorString(err));
    }
    *out = v;
    return LC2E_OK;
}

void lc2e_volume_destroy(lc2e_volume* v) {
    if (!v) return;
    (void)hipStreamSynchronize(v->e->stream);
    v->e->release(v->d, v->bytes);
    delete v;
}

int lc2e_pair_create(lc2e_engine* e, lc2e_volume* fixed, lc2e_volume* moving,
                     const float resampled_spacing[3], int high_res_is_moving,
                     int center_overlaid, int patch_radius, int dim, lc2e_pair** out) {
    if (!e || !fixed || !moving || !resampled_spacing || !out)
        return fail(LC2E_ERR_INVALID, "null argument");
    if (fixed->e != e || moving->e != e) return fail(LC2E_ERR_INVALID, "volume of another engine");
    if (dim != 2 && dim != 3) return fail(LC2E_ERR_INVALID, "dim must be 2 or 3");
    if (patch_radius < 1) return fail(LC2E_ERR_INVALID, "patch radius must be >= 1 (0 selects MI, unsupported)");
    if (patch_radius > lc2e::kMaxRSep)
        return fail(LC2E_ERR_UNSUPPORTED, "patch radius %d > %d", patch_radius, lc2e::kMaxRSep);
    // radii beyond the fused kernels, and 3D sources whose slices hold 2^24 voxels or more
    // (beyond the fused samplers' 24-bit offsets), take the separable path
    const bool big_slices = moving->g.dims[2] > 1 &&
                            (long long)moving->g.dims[0] * moving->g.dims[1] >= lc2e::kMaxFusedSlice;
    const bool sep = (dim == 3 && patch_radius > lc2e::kMaxR3) ||
                     (dim == 2 && patch_radius > lc2e::kMaxR2) || big_slices;
    const int src_ndim = moving->g.dims[2] > 1 ? 3 : 2;
    if (dim == 3 && (src_ndim != 3 || fixed->g.dims[2] < 1))
        return fail(LC2E_ERR_INVALID, "3D LC2 needs a 3D moving volume");
    for (int a = 0; a < dim; ++a)
        if (!(resampled_spacing[a] > 0.0f)) return fail(LC2E_ERR_INVALID, "resampled spacing must be positive");

    lc2e::PairGeometry pg = lc2e::make_pair_geometry(fixed->g, moving->g, resampled_spacing, dim,
                                                     src_ndim, center_overlaid != 0);
    const long long V = (long long)pg.virt.dims[0] * pg.virt.dims[1] * pg.virt.dims[2];
    if (V >= (1LL << 31)) return fail(LC2E_ERR_UNSUPPORTED, "virtual grid exceeds 2^31 voxels");

    auto* p = new lc2e_pair();
    p->gen = g_generation.fetch_add(1);
    p->e = e;
    p->fixed = fixed;
    p->moving = moving;
    p->info.virt = pg.virt;
    for (int a = 0; a < 3; ++a) {
        p->info.fixed_center[a] = pg.fixed_center[a];
        p->info.shift[a] = pg.shift[a];
    }
    p->info.grid_ndim = dim;
    p->info.src_ndim = src_ndim;
    p->info.high_res_is_moving = high_res_is_moving ? 1 : 0;
    p->info.patch_radius = patch_radius;
    p->info.n_frames = dim == 2 ? pg.virt.dims[2] : 1;
    p->info.per_frame_matrices = 0;
    const lc2e_geom& vg = pg.virt;
    p->v = lc2e::VGeom{vg.dims[0], vg.dims[1], vg.dims[2], vg.spacing[0], vg.spacing[1],
                       vg.spacing[2], vg.origin[0], vg.origin[1], vg.origin[2]};
    const lc2e_geom& mg = moving->g;
    p->sg = lc2e::make_sgeom(mg.dims, mg.spacing, mg.origin);
    p->vbytes = (size_t)V * sizeof(float);
    int rc = e->alloc((void**)&p->stat_img, p->vbytes);
    if (rc < 0) {
        delete p;
        return rc;
    }
    if (sep) {
        p->sep = true;
        p->sep_grid_bytes = 2 * (size_t)V * sizeof(float);
        p->sep_sum_bytes = 2 * (size_t)lc2e::kSepChannels(high_res_is_moving != 0) * V * sizeof(double);
        if ((rc = e->alloc((void**)&p->sep_grids, p->sep_grid_bytes)) < 0 ||
            (rc = e->alloc((void**)&p->sep_sums, p->sep_sum_bytes)) < 0) {
            if (p->sep_grids) e->release(p->sep_grids, p->sep_grid_bytes);
            e->release(p->stat_img, p->vbytes);
            delete p;
            return rc;
        }
    }
    const lc2e_geom& fg = fixed->g;
    lc2e::SGeom fsg = lc2e::make_sgeom(fg.dims, fg.spacing, fg.origin);
    hipError_t err = lc2e::launch_warp_static(fixed->d, fixed->f16, fsg, p->v, dim, src_ndim,
                                              p->stat_img, e->stream);
    if (err == hipSuccess && high_res_is_moving) {
        // the fixed image is the ultrasound: its patch statistics are pose-independent
        rc = e->alloc((void**)&p->ustat, 2 * (size_t)V * sizeof(double));
        if (rc < 0) {
            lc2e_pair_destroy(p);
            return rc;
        }
        err = lc2e::launch_ustats(p->stat_img, p->v, dim, patch_radius, p->ustat, p->ustat + V,
                                  e->stream);
    }
    if (err == hipSuccess && !high_res_is_moving) {
        // the fixed image is the high-resolution one: its |Sobel| is pose-independent
        rc = e->alloc((void**)&p->stat_grad, p->vbytes);
        if (rc < 0) {
            lc2e_pair_destroy(p);
            return rc;
        }
        err = lc2e::launch_sobel(p->stat_img, p->v, dim, p->stat_grad, e->stream);
    }
    if (err == hipSuccess) err = hipStreamSynchronize(e->stream);
    if (err != hipSuccess) {
        lc2e_pair_destroy(p);
        return fail(LC2E_ERR_HIP, "pair set-up failed: %s", hipGetErrorString(err));
    }
    *out = p;
    return LC2E_OK;
}

void lc2e_pair_destroy(lc2e_pair* p) {
    if (!p) return;
    (void)hipStreamSynchronize(p->e->stream);
    free_batch(p);
    free_mi(p);
    if (p->mi_ext_stat) p->e->release(p->mi_ext_stat, 2 * sizeof(int));
    if (p->ev0) (void)hipEventDestroy(p->ev0);
    if (p->ev1) (void)hipEventDestroy(p->ev1);
    p->e->release(p->stat_img, p->vbytes);
    if (p->stat_grad) p->e->release(p->stat_grad, p->vbytes);
    if (p->ustat) p->e->release(p->ustat, 2 * (p->vbytes / sizeof(float)) * sizeof(double));
    if (p->sep_grids) p->e->release(p->sep_grids, p->sep_grid_bytes);
    if (p->sep_sums) p->e->release(p->sep_sums, p->sep_sum_bytes);
    delete p;
}

int lc2e_pair_get_info(lc2e_pair* p, lc2e_pair_info* info) {
    if (!p || !info) return fail(LC2E_ERR_INVALID, "null argument");
    *info = p->info;
    return LC2E_OK;
}

int lc2e_pair_set_per_frame(lc2e_pair* p, int per_frame) {
    if (!p) return fail(LC2E_ERR_INVALID, "null pair");
    if (p->info.grid_ndim != 2 && per_frame)
        return fail(LC2E_ERR_INVALID, "per-frame matrices need a 2D frame stack");
    free_batch(p);
    p->info.per_frame_matrices = per_frame ? 1 : 0;
    return LC2E_OK;
}

int lc2e_pose_to_matrix(lc2e_pair* p, const double* params, int n, float out12[12]) {
    if (!p || !params || !out12) return fail(LC2E_ERR_INVALID, "null argument");
    int rc = lc2e::pose_to_matrix(p->info, params, n, out12);
    if (rc < 0) return fail(rc, "pose_to_matrix: %d parameters not valid for this pair", n);
    return LC2E_OK;
}

int lc2e_eval_batch(lc2e_pair* p, int n, const float* mats, double* lc2, double* sum_measure,
                    double* sum_w) {
    if (!p || !mats || n < 0) return fail(LC2E_ERR_INVALID, "invalid argument");
    if (n == 0) return LC2E_OK;
    HIP_TRY(hipSetDevice(p->e->device));
    const int mpp = p->mats_per_pose();
    const int chunk_max = 65535;
    int status = LC2E_OK;
    for (int base = 0; base < n; base += chunk_max) {
        const int m = (n - base) < chunk_max ? (n - base) : chunk_max;
        const Tiling t = tiling_for(p, m);
        int rc = ensure_capacity(p, m, t.bpp);
        if (rc < 0) return rc;
        std::memcpy(p->h_mats, mats + (size_t)base * mpp * 12, (size_t)m * mpp * 12 * sizeof(float));
        rc = run_batch(p, m, t);
        if (rc < 0) return rc;
        for (int i = 0; i < m; ++i) {
            const double sm = p->h_out[3 * i], sw = p->h_out[3 * i + 1], l = p->h_out[3 * i + 2];
            if (lc2) lc2[base + i] = l;
            if (sum_measure) sum_measure[base + i] = sm;
            if (sum_w) sum_w[base + i] = sw;
            if (!(sw > 0.0) || !std::isfinite(sm / sw)) status = LC2E_OK_DEGENERATE;
        }
    }
    return status;
}

int lc2e_eval(lc2e_pair* p, const float* m12, double* lc2) {
    return lc2e_eval_batch(p, 1, m12, lc2, nullptr, nullptr);
}

int lc2e_eval_batch_noalloc(lc2e_pair* p, int n, const float* mats, double* lc2,
                            double* sum_measure, double* sum_w) {
    if (!p || !mats || n < 0) return fail(LC2E_ERR_INVALID, "invalid argument");
    if (n == 0) return LC2E_OK;
    lc2e_engine* e = p->e;
    if (!e->scratch)
        return fail(LC2E_ERR_HIP, "device scratch unavailable: %s", e->scratch_error.c_str());
    // One device array serves every call on the device: calls are serialised here, so a
    // call that overlaps another one on the same device waits for it.
    static std::mutex device_locks[64];
    std::lock_guard<std::mutex> lk(device_locks[e->device & 63]);
    HIP_TRY(hipSetDevice(e->device));
    hipStream_t st = e->stream;
    const int mpp = p->mats_per_pose();
    const Tiling t = tiling_for(p, n);
    // scratch layout: results [n][3] f64, matrices [n][mpp][12] f32, partials [n][bpp][2] f64
    auto up256 = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t out_b = up256((size_t)n * 3 * sizeof(double));
    const size_t mat_b = up256((size_t)n * mpp * 12 * sizeof(float));
    const size_t part_b = (size_t)n * (t.tiles > 0 ? t.bpp : 0) * 2 * sizeof(double);
    if (out_b + mat_b + part_b > lc2e::kScratchBytes)
        return fail(LC2E_ERR_INVALID,
                    "%d poses need %zu bytes of device scratch; the allocation-free entry point "
                    "has %zu (use lc2e_eval_batch or fewer poses per call)",
                    n, out_b + mat_b + part_b, lc2e::kScratchBytes);
    char* base = static_cast<char*>(e->scratch);
    double* d_out = reinterpret_cast<double*>(base);
    float* d_mats = reinterpret_cast<float*>(base + out_b);
    double* d_part = reinterpret_cast<double*>(base + out_b + mat_b);
    HIP_TRY(hipMemcpyAsync(d_mats, mats, (size_t)n * mpp * 12 * sizeof(float),
                           hipMemcpyHostToDevice, st));
    if (t.tiles > 0) {
        // every partial slot is written by the fused kernel before k_reduce reads it, so the
        // previous call's contents need no clearing
        const lc2e::FusedArgs a = make_args(p, d_mats, d_part, t);
        HIP_TRY(launch_eval(p, a, n, st));
    }
    HIP_TRY(lc2e::launch_reduce(d_part, t.tiles > 0 ? t.bpp : 0, n, d_out, st));
    // results back through a fixed host buffer, a chunk of poses at a time
    int status = LC2E_OK;
    constexpr int kChunk = 128;
    double h[kChunk * 3];
    for (int b = 0; b < n; b += kChunk) {
        const int m = n - b < kChunk ? n - b : kChunk;
        HIP_TRY(hipMemcpyAsync(h, d_out + (size_t)b * 3, (size_t)m * 3 * sizeof(double),
                               hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        for (int i = 0; i < m; ++i) {
            const double sm = h[3 * i], sw = h[3 * i + 1], l = h[3 * i + 2];
            if (lc2) lc2[b + i] = l;
            if (sum_measure) sum_measure[b + i] = sm;
            if (sum_w) sum_w[b + i] = sw;
            if (!(sw > 0.0) || !std::isfinite(sm / sw)) status = LC2E_OK_DEGENERATE;
        }
    }
    return status;
}

int lc2e_eval_batch_async(lc2e_pair* p, int n, const float* d_mats, double* d_out, void* stream) {
    if (!p || !d_mats || !d_out || n <= 0) return fail(LC2E_ERR_INVALID, "invalid argument");
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : p->e->stream;
    // partials are per pair; size them for n
    const Tiling t = tiling_for(p, n);
    int rc = ensure_capacity(p, n, t.bpp);
    if (rc < 0) return rc;
    if (t.tiles > 0) {
        lc2e::FusedArgs a = make_args(p, d_mats, p->d_partials, t);
        HIP_TRY(launch_eval(p, a, n, st));
    }
    HIP_TRY(lc2e::launch_reduce(p->d_partials, t.tiles > 0 ? t.bpp : 0, n, d_out, st));
    return LC2E_OK;
}

int lc2e_eval_mi_batch(lc2e_pair* p, int n, const float* mats, int bins_us, int bins_hr,
                       double* nmi) {
    if (!p || !mats || !nmi || n < 0) return fail(LC2E_ERR_INVALID, "invalid argument");
    if (n == 0) return LC2E_OK;
    if (bins_us < 2 || bins_hr < 2 || bins_us > lc2e::kMiMaxBins || bins_hr > lc2e::kMiMaxBins)
        return fail(LC2E_ERR_INVALID, "histogram bins must lie in [2, %d]", lc2e::kMiMaxBins);
    HIP_TRY(hipSetDevice(p->e->device));
    lc2e_engine* e = p->e;
    hipStream_t st = e->stream;
    const int mpp = p->mats_per_pose();
    const size_t nb = (size_t)bins_us * bins_hr;
    int rc;
    if (!p->mi_ext_stat) {
        if ((rc = e->alloc((void**)&p->mi_ext_stat, 2 * sizeof(int))) < 0) return rc;
        p->mi_stat_ready = false;
    }
    const int chunk_max = 64;
    for (int base = 0; base < n; base += chunk_max) {
        const int m = (n - base) < chunk_max ? (n - base) : chunk_max;
        const Tiling t = tiling_for(p, m);
        if ((rc = ensure_capacity(p, m, t.bpp)) < 0) return rc;
        if (m > p->mi_cap || nb != p->mi_bins) {
            free_mi(p);
            const int cap = m > 8 ? chunk_max : 8;
            if ((rc = e->alloc((void**)&p->mi_ext, (size_t)cap * 2 * sizeof(int))) < 0) return rc;
            p->mi_cap = cap;
            p->mi_bins = nb;
            if ((rc = e->alloc((void**)&p->mi_hist, (size_t)cap * nb * sizeof(unsigned))) < 0 ||
                (rc = e->alloc((void**)&p->mi_out, (size_t)cap * sizeof(double))) < 0) {
                free_mi(p);
                return rc;
            }
        }
        std::memcpy(p->h_mats, mats + (size_t)base * mpp * 12, (size_t)m * mpp * 12 * sizeof(float));
        HIP_TRY(hipMemcpyAsync(p->d_mats, p->h_mats, (size_t)m * mpp * 12 * sizeof(float),
                               hipMemcpyHostToDevice, st));
        lc2e::FusedArgs a = make_args(p, p->d_mats, p->d_partials, t);
        // the static image is U when the moving image is the high-resolution one
        const int us_static = p->info.high_res_is_moving;
        HIP_TRY(lc2e::launch_mi(a, p->info.grid_ndim, p->info.src_ndim, us_static, m, bins_us,
                                bins_hr, p->stat_img,
                                p->mi_ext_stat, p->mi_stat_ready ? 1 : 0, p->mi_ext, p->mi_hist,
                                p->mi_out, st));
        p->mi_stat_ready = true;
        HIP_TRY(hipMemcpyAsync(p->h_out, p->mi_out, (size_t)m * sizeof(double),
                               hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        for (int i = 0; i < m; ++i) nmi[base + i] = p->h_out[i];
    }
    return LC2E_OK;
}

int lc2e_sync(lc2e_engine* e) {
    if (!e) return fail(LC2E_ERR_INVALID, "null engine");
    HIP_TRY(hipSetDevice(e->device));
    HIP_TRY(hipDeviceSynchronize());
    return LC2E_OK;
}

int lc2e_debug_fetch(lc2e_pair* p, lc2e_stage stage, const float* m12, void* host) {
    if (!p || !host || !m12) return fail(LC2E_ERR_INVALID, "null argument");
    HIP_TRY(hipSetDevice(p->e->device));
    hipStream_t st = p->e->stream;
    const long long V = (long long)p->v.X * p->v.Y * p->v.Z;
    const int mpp = p->mats_per_pose();
    const Tiling t = tiling_for(p, 1);
    int rc = ensure_capacity(p, 1, t.bpp);
    if (rc < 0) return rc;
    HIP_TRY(hipMemcpy(p->d_mats, m12, (size_t)mpp * 12 * sizeof(float), hipMemcpyHostToDevice));
    lc2e::FusedArgs a = make_args(p, p->d_mats, p->d_partials, t);
    const bool movi = p->info.high_res_is_moving != 0;
    if (stage == LC2E_STAGE_WARP_I || stage == LC2E_STAGE_WARP_US || stage == LC2E_STAGE_GRAD) {
        const bool want_moving = (stage == LC2E_STAGE_WARP_I) == movi || (stage == LC2E_STAGE_GRAD && movi);
        float* buf = nullptr;
        float* buf2 = nullptr;
        HIP_TRY(hipMalloc(&buf, V * sizeof(float)));
        hipError_t err = hipSuccess;
        const float* src = p->stat_img;
        if (stage == LC2E_STAGE_GRAD && !movi) {
            src = p->stat_grad;
        } else if (want_moving) {
            err = lc2e::launch_warp_moving(a, p->info.grid_ndim, p->info.src_ndim, buf, st);
            src = buf;
            if (err == hipSuccess && stage == LC2E_STAGE_GRAD) {
                err = hipMalloc(&buf2, V * sizeof(float));
                if (err == hipSuccess) err = lc2e::launch_sobel(buf, p->v, p->info.grid_ndim, buf2, st);
                src = buf2;
            }
        }
        if (err == hipSuccess) err = hipMemcpyAsync(host, src, V * sizeof(float), hipMemcpyDeviceToHost, st);
        if (err == hipSuccess) err = hipStreamSynchronize(st);
        (void)hipFree(buf);
        if (buf2) (void)hipFree(buf2);
        if (err != hipSuccess) return fail(LC2E_ERR_HIP, "debug fetch failed: %s", hipGetErrorString(err));
        return LC2E_OK;
    }
    if (stage != LC2E_STAGE_W && stage != LC2E_STAGE_MEASURE)
        return fail(LC2E_ERR_INVALID, "unknown stage %d", (int)stage);
    double* maps = nullptr;
    HIP_TRY(hipMalloc(&maps, 2 * V * sizeof(double)));
    hipError_t err = hipMemsetAsync(maps, 0, 2 * V * sizeof(double), st);
    a.wmap = maps;
    a.mmap = maps + V;
    if (err == hipSuccess && t.tiles > 0)
        err = launch_eval(p, a, 1, st);
    if (err == hipSuccess)
        err = hipMemcpyAsync(host, stage == LC2E_STAGE_W ? a.wmap : a.mmap, V * sizeof(double),
                             hipMemcpyDeviceToHost, st);
    if (err == hipSuccess) err = hipStreamSynchronize(st);
    (void)hipFree(maps);
    if (err != hipSuccess) return fail(LC2E_ERR_HIP, "debug fetch failed: %s", hipGetErrorString(err));
    return LC2E_OK;
}

int lc2e_pair_set_tuning(lc2e_pair* p, int zchunk) {
    if (!p || zchunk < 0) return fail(LC2E_ERR_INVALID, "invalid argument");
    (void)hipStreamSynchronize(p->e->stream);
    drop_graphs(p);  // captured launches carry the old decomposition
    p->zchunk_forced = zchunk > 0 ? round_zchunk(zchunk) : 0;
    return LC2E_OK;
}

int lc2e_pair_set_tuning_n(lc2e_pair* p, int n, int zchunk) {
    if (!p || n <= 0 || zchunk < 0) return fail(LC2E_ERR_INVALID, "invalid argument");
    (void)hipStreamSynchronize(p->e->stream);
    drop_graphs(p);
    if (zchunk == 0) p->tuned.erase(n);
    else p->tuned[n] = round_zchunk(zchunk);
    return LC2E_OK;
}

int lc2e_pair_autotune(lc2e_pair* p, int n, const float* mats, int iters, int* best_zchunk,
                       double* best_ms) {
    if (!p || !mats || n <= 0 || iters <= 0) return fail(LC2E_ERR_INVALID, "invalid argument");
    HIP_TRY(hipSetDevice(p->e->device));
    if (p->info.grid_ndim != 3 || p->sep) {
        if (best_zchunk) *best_zchunk = 1;
        if (best_ms) *best_ms = 0.0;
        return LC2E_OK;
    }
    const int ez = p->v.Z - 2 * p->info.patch_radius;
    // chunks shorter than the column, then the whole column in one chunk
    std::vector<int> cand;
    for (int zc : {8, 16, 24, 32, 48, 64, 96, 128, 192, 256})
        if (zc < ez) cand.push_back(zc);
    cand.push_back(ez > 0 ? round_zchunk(ez) : lc2e::kZBlock);
    struct Events {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        ~Events() {
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
        }
    } ev;
    HIP_TRY(hipEventCreate(&ev.e0));
    HIP_TRY(hipEventCreate(&ev.e1));
    hipEvent_t e0 = ev.e0, e1 = ev.e1;
    int best = 0;
    double bms = 1e300;
    hipStream_t st = p->e->stream;
    drop_graphs(p);
    const int forced = p->zchunk_forced;
    for (int zc : cand) {
        p->zchunk_forced = zc;
        const Tiling t = tiling_for(p, n);
        p->zchunk_forced = forced;
        int rc = ensure_capacity(p, n, t.bpp);
        if (rc < 0) return rc;
        HIP_TRY(hipMemcpy(p->d_mats, mats, (size_t)n * p->mats_per_pose() * 12 * sizeof(float),
                          hipMemcpyHostToDevice));
        lc2e::FusedArgs a = make_args(p, p->d_mats, p->d_partials, t);
        HIP_TRY(lc2e::launch_fused(a, 3, p->info.src_ndim, p->info.high_res_is_moving, n, st));
        HIP_TRY(hipEventRecord(e0, st));
        for (int it = 0; it < iters; ++it)
            HIP_TRY(lc2e::launch_fused(a, 3, p->info.src_ndim, p->info.high_res_is_moving, n, st));
        HIP_TRY(hipEventRecord(e1, st));
        HIP_TRY(hipEventSynchronize(e1));
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        if (ms / (double)iters < bms) {
            bms = ms / (double)iters;
            best = zc;
        }
    }
    p->zchunk_forced = forced;  // a forced z-chunk keeps precedence over the tuned value
    p->tuned[n] = best;
    if (best_zchunk) *best_zchunk = best;
    if (best_ms) *best_ms = bms;
    return LC2E_OK;
}

int lc2e_pair_get_tuning(lc2e_pair* p, int* zchunk) {
    if (!p || !zchunk) return fail(LC2E_ERR_INVALID, "invalid argument");
    *zchunk = p->zchunk_forced > 0 ? p->zchunk_forced
              : p->last_zchunk > 0 ? p->last_zchunk
              : (p->info.grid_ndim == 3 ? round_zchunk(model_zchunk(p, 1)) : 1);
    return LC2E_OK;
}

int lc2e_pair_profile(lc2e_pair* p, int enable) {
    if (!p) return fail(LC2E_ERR_INVALID, "null pair");
    if ((enable != 0) != p->profile) (void)hipStreamSynchronize(p->e->stream);
    if (enable && !p->ev0) {
        HIP_TRY(hipEventCreate(&p->ev0));
        HIP_TRY(hipEventCreate(&p->ev1));
    }
    p->profile = enable != 0;
    p->prof_ms = 0.0;
    p->prof_launches = 0;
    return LC2E_OK;
}

int lc2e_pair_profile_read(lc2e_pair* p, double* kernel_ms, long long* launches) {
    if (!p) return fail(LC2E_ERR_INVALID, "null pair");
    if (kernel_ms) *kernel_ms = p->prof_ms;
    if (launches) *launches = p->prof_launches;
    return LC2E_OK;
}

int lc2e_pair_use_graph(lc2e_pair* p, int enable) {
    if (!p) return fail(LC2E_ERR_INVALID, "null pair");
    if (!enable) drop_graphs(p);
    p->use_graph = enable != 0;
    return LC2E_OK;
}

}  // extern "C"
