// Serving loop (include/lc2serve.h). One worker thread owns a few launch slots, each with
// its own HIP stream and device/pinned staging buffers. It merges the queued requests of a
// pair into one lc2e_eval_batch_async launch on a free slot, keeps up to `slots` launches
// of different pairs in flight on the device at once while their work-groups fit the
// device together (a launch that fills the GPU runs alone), and completes requests as
// their launches finish. Requests that arrive meanwhile are merged into the next launch of
// their pair (continuous batching at launch granularity).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "lc2serve.h"

namespace lc2e {
// lc2e_engine.cpp
void set_last_error(const std::string& msg);
int engine_alloc(lc2e_engine* e, void** p, size_t bytes);
void engine_release(lc2e_engine* e, void* p, size_t bytes);
int engine_device(lc2e_engine* e);
double launch_share(lc2e_pair* p, int n);
unsigned long long pair_generation(lc2e_pair* p);
int pair_prepare(lc2e_pair* p, int n);
}  // namespace lc2e

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    long long ticket = 0;
    lc2e_pair* pair = nullptr;
    int n = 0;
    std::vector<float> mats;
    std::vector<double> lc2, sm, sw;
    int status = LC2E_OK;
    std::string error;  // lc2e_last_error() text when status < 0
    bool done = false;
    Clock::time_point arrived;
};
using RequestPtr = std::shared_ptr<Request>;

int pair_mats_per_pose(lc2e_pair* p) {
    lc2e_pair_info info;
    if (lc2e_pair_get_info(p, &info) < 0) return -1;
    return info.per_frame_matrices ? info.n_frames : 1;
}

// A launch captured as a HIP graph on a slot: matrix upload and kernels for one pair and
// batch size, valid while the pair's launch state keeps its generation.
struct SlotGraph {
    unsigned long long gen = 0;
    int uses = 0;  // launches seen for this key (the first one runs eagerly)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

void destroy_graph(SlotGraph& g) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
    g.exec = nullptr;
    g.graph = nullptr;
}

// A launch slot: a stream, a completion event and staging buffers sized on demand.
struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    size_t mats_cap = 0;  // floats
    int out_cap = 0;      // poses
    float* d_mats = nullptr;
    float* h_mats = nullptr;      // pinned
    double* h_out = nullptr;      // pinned, mapped and coherent: the reduction writes here
    double* h_out_dev = nullptr;  // its device address
    bool busy = false;
    lc2e_pair* pair = nullptr;
    std::vector<RequestPtr> batch;
    int total = 0;
    double share = 0.0;  // of the device (lc2e::launch_share)
    std::map<std::pair<lc2e_pair*, int>, SlotGraph> graphs;  // by (pair, poses)
};

void fill_error(Request& r, int rc, const std::string& err) {
    r.lc2.assign(r.n, 0.0);
    r.sm.assign(r.n, 0.0);
    r.sw.assign(r.n, 0.0);
    r.status = rc;
    r.error = err;
}

// Results of poses [off, off + r.n) of a finished launch into the request. Its status is
// its own: OK_DEGENERATE only if one of its poses had no weight (lc2e_eval_batch's rule).
void fill(Request& r, const double* out, int off) {
    r.lc2.resize(r.n);
    r.sm.resize(r.n);
    r.sw.resize(r.n);
    r.status = LC2E_OK;
    for (int i = 0; i < r.n; ++i) {
        const double* o = out + 3 * (size_t)(off + i);
        r.sm[i] = o[0];
        r.sw[i] = o[1];
        r.lc2[i] = o[2];
        if (!(o[1] > 0.0) || !std::isfinite(o[0] / o[1])) r.status = LC2E_OK_DEGENERATE;
    }
}

}  // namespace

struct lc2s_server {
    lc2e_engine* engine = nullptr;
    int device = 0;
    int max_poses = 64;
    std::chrono::duration<double, std::micro> window{0.0};
    std::vector<Slot> slots;
    // worker thread only
    std::set<lc2e_pair*> in_flight;  // pairs with a launch on the device
    long long extra_launches = 0;    // per-request launches of split merged launches
    std::atomic<bool> use_graphs{false};       // replay captured launches (lc2s_use_graphs)
    std::atomic<long long> graph_launches{0};

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<RequestPtr> queue;             // FIFO of pending requests
    std::map<long long, RequestPtr> tickets;  // accepted, not yet collected
    long long next_ticket = 1;
    bool stopping = false;
    long long launches = 0, completed = 0, poses = 0;
    int busy_max = 0;
    std::thread worker;

    ~lc2s_server() {
        (void)hipSetDevice(device);
        for (Slot& s : slots) {
            if (s.stream) (void)hipStreamSynchronize(s.stream);
            free_buffers(s);
            if (s.done) (void)hipEventDestroy(s.done);
            if (s.stream) (void)hipStreamDestroy(s.stream);
        }
    }

    void free_buffers(Slot& s) {
        for (auto& kv : s.graphs) destroy_graph(kv.second);  // they captured these buffers
        s.graphs.clear();
        if (s.d_mats) lc2e::engine_release(engine, s.d_mats, s.mats_cap * sizeof(float));
        if (s.h_mats) (void)hipHostFree(s.h_mats);
        if (s.h_out) (void)hipHostFree(s.h_out);
        s.d_mats = nullptr;
        s.h_mats = nullptr;
        s.h_out = nullptr;
        s.h_out_dev = nullptr;
        s.mats_cap = 0;
        s.out_cap = 0;
    }

    // Staging for `floats` matrix floats and `n` poses (the slot is idle). Device bytes are
    // charged to the engine's HBM budget. On failure the slot holds no staging at all, so
    // the next call allocates again.
    int reserve(Slot& s, size_t floats, int n) {
        if (floats <= s.mats_cap && n <= s.out_cap) return LC2E_OK;
        const size_t fc = std::max(floats, s.mats_cap);
        const int nc = std::max(n, s.out_cap);
        free_buffers(s);
        int rc = lc2e::engine_alloc(engine, (void**)&s.d_mats, fc * sizeof(float));
        if (rc < 0) {
            s.d_mats = nullptr;
        } else {
            s.mats_cap = fc;
            s.out_cap = nc;
        }
        if (rc >= 0) {
            void* hm = nullptr;
            void* ho = nullptr;
            void* hod = nullptr;
            if (hipHostMalloc(&hm, fc * sizeof(float), hipHostMallocDefault) == hipSuccess) s.h_mats = (float*)hm;
            if (hipHostMalloc(&ho, (size_t)nc * 3 * sizeof(double),
                              hipHostMallocMapped | hipHostMallocCoherent) == hipSuccess) {
                s.h_out = (double*)ho;
                if (hipHostGetDevicePointer(&hod, ho, 0) == hipSuccess) s.h_out_dev = (double*)hod;
            }
            if (!s.h_mats || !s.h_out || !s.h_out_dev) {
                lc2e::set_last_error("serving: pinned staging allocation failed");
                rc = LC2E_ERR_HIP;
            }
        }
        if (rc < 0) free_buffers(s);
        return rc;
    }

    int queued_poses(lc2e_pair* p) const {
        int t = 0;
        for (const auto& r : queue)
            if (r->pair == p) t += r->n;
        return t;
    }

    // Under the lock: the pair of the oldest queued request whose pair has no launch in
    // flight and whose batching window has closed (or whose queued poses fill a launch).
    // `held`/`wake`: a request is held by its window, and the earliest time one closes.
    lc2e_pair* pick(Clock::time_point now, Clock::time_point& wake, bool& held) {
        std::set<lc2e_pair*> seen;
        for (const auto& r : queue) {
            if (in_flight.count(r->pair) || !seen.insert(r->pair).second) continue;
            if (window.count() > 0.0 && !stopping) {
                const auto until = r->arrived + std::chrono::duration_cast<Clock::duration>(window);
                if (now < until && queued_poses(r->pair) < max_poses) {
                    if (!held || until < wake) wake = until;
                    held = true;
                    continue;
                }
            }
            return r->pair;
        }
        return nullptr;
    }

    // Under the lock: the poses take() would put into one launch of the pair.
    int peek_total(lc2e_pair* pair) const {
        int total = 0;
        for (const auto& r : queue) {
            if (r->pair != pair) continue;
            if (total > 0 && total + r->n > max_poses) break;
            total += r->n;
        }
        return total;
    }

    // Under the lock: move the pair's oldest queued requests, in arrival order, into the
    // slot while they fit one launch (a request larger than max_poses goes alone). The
    // first request that does not fit ends the launch, so a pair's requests are launched
    // in the order they arrived.
    void take(Slot& s, lc2e_pair* pair) {
        s.batch.clear();
        s.total = 0;
        for (auto it = queue.begin(); it != queue.end();) {
            Request& r = **it;
            if (r.pair != pair) {
                ++it;
                continue;
            }
            if (!s.batch.empty() && s.total + r.n > max_poses) break;
            s.total += r.n;
            s.batch.push_back(*it);
            it = queue.erase(it);
        }
        s.pair = pair;
        s.busy = true;
    }

    // Outside the lock: upload, launch and download on the slot's stream, then record its
    // event. The slot's staging and the pair's partial sums are not in use by the device.
    // The launch itself: upload and lc2e_eval_batch_async on the slot's stream; the
    // reduction writes the results straight into the slot's mapped host buffer (one copy
    // command less per launch).
    int enqueue_ops(Slot& s, int n, size_t floats) {
        if (hipMemcpyAsync(s.d_mats, s.h_mats, floats * sizeof(float), hipMemcpyHostToDevice,
                           s.stream) != hipSuccess) {
            lc2e::set_last_error("serving: matrix upload failed");
            return LC2E_ERR_HIP;
        }
        return lc2e_eval_batch_async(s.pair, n, s.d_mats, s.h_out_dev, s.stream);
    }

    // A (pair, n) launch runs eagerly the first time (which also sizes the pair's buffers
    // and does one-time kernel set-up), is captured as a HIP graph the second time and
    // replayed after that, until the pair's launch state changes (generation).
    int enqueue(Slot& s, int n, size_t floats) {
        int rc = LC2E_OK;
        SlotGraph* g = nullptr;
        if (use_graphs) {
            if ((rc = lc2e::pair_prepare(s.pair, n)) < 0) return rc;  // no allocation when captured
            g = &s.graphs[{s.pair, n}];
            const unsigned long long gen = lc2e::pair_generation(s.pair);
            if (g->gen != gen) {
                destroy_graph(*g);
                g->gen = gen;
                g->uses = 0;
            }
            if (!g->exec && g->uses >= 1) {
                hipGraph_t graph = nullptr;
                if (hipStreamBeginCapture(s.stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
                    lc2e::set_last_error("serving: stream capture failed");
                    return LC2E_ERR_HIP;
                }
                rc = enqueue_ops(s, n, floats);
                hipError_t ce = hipStreamEndCapture(s.stream, &graph);
                if (rc >= 0 && ce == hipSuccess) ce = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
                if (rc < 0 || ce != hipSuccess) {
                    if (graph) (void)hipGraphDestroy(graph);
                    g->exec = nullptr;
                    if (rc >= 0) {
                        lc2e::set_last_error("serving: graph capture failed");
                        rc = LC2E_ERR_HIP;
                    }
                    return rc;
                }
                g->graph = graph;
            }
            ++g->uses;
        }
        if (g && g->exec) {
            if (hipGraphLaunch(g->exec, s.stream) != hipSuccess) {
                lc2e::set_last_error("serving: graph launch failed");
                return LC2E_ERR_HIP;
            }
            ++graph_launches;
        } else if ((rc = enqueue_ops(s, n, floats)) < 0) {
            return rc;
        }
        if (hipEventRecord(s.done, s.stream) != hipSuccess) {
            lc2e::set_last_error("serving: completion event failed");
            return LC2E_ERR_HIP;
        }
        return LC2E_OK;
    }

    // Outside the lock: start the slot's launch. Returns true if the slot's requests are
    // already complete (an error, or a merged launch split for the budget).
    bool launch(Slot& s) {
        const int mpp = pair_mats_per_pose(s.pair);
        int rc = mpp < 0 ? LC2E_ERR_INVALID : reserve(s, (size_t)s.total * mpp * 12, s.total);
        if (rc >= 0) {
            size_t off = 0;
            for (auto& r : s.batch) {
                std::memcpy(s.h_mats + off, r->mats.data(), r->mats.size() * sizeof(float));
                off += r->mats.size();
            }
            rc = enqueue(s, s.total, off);
            if (rc >= 0) return false;
        }
        (void)hipStreamSynchronize(s.stream);
        if (rc == LC2E_ERR_BUDGET && s.batch.size() > 1) {
            run_split(s, mpp);
        } else {
            const std::string err = lc2e_last_error();
            for (auto& r : s.batch) fill_error(*r, rc, err);
        }
        return true;
    }

    // The merged launch's buffers do not fit the HBM budget: evaluate each request of the
    // slot on its own, synchronously, so that none fails for what it was merged with.
    void run_split(Slot& s, int mpp) {
        for (auto& r : s.batch) {
            int rc = reserve(s, (size_t)r->n * mpp * 12, r->n);
            if (rc >= 0) {
                std::memcpy(s.h_mats, r->mats.data(), r->mats.size() * sizeof(float));
                rc = enqueue(s, r->n, (size_t)r->n * mpp * 12);
            }
            if (rc >= 0 && hipStreamSynchronize(s.stream) != hipSuccess) {
                lc2e::set_last_error("serving: launch failed");
                rc = LC2E_ERR_HIP;
            }
            if (rc < 0) {
                (void)hipStreamSynchronize(s.stream);
                fill_error(*r, rc, lc2e_last_error());
            } else {
                fill(*r, s.h_out, 0);
            }
            ++extra_launches;
        }
    }

    void run() {
        (void)hipSetDevice(device);
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            // 1. fill free slots with the oldest dispatchable pairs, while the launches in
            // flight and the next one together fit the device (a launch that fills it alone
            // runs alone); the oldest request that does not fit waits, so none is passed over
            std::vector<Slot*> started;
            Clock::time_point wake{};
            bool held = false;
            double busy_share = 0.0;
            for (Slot& s : slots) busy_share += s.busy ? s.share : 0.0;
            for (Slot& s : slots) {
                if (s.busy) continue;
                lc2e_pair* pair = pick(Clock::now(), wake, held);
                if (!pair) break;
                const double share = lc2e::launch_share(pair, peek_total(pair));
                if (busy_share > 0.0 && busy_share + share > 1.0) break;
                take(s, pair);
                s.share = share;
                busy_share += share;
                in_flight.insert(pair);
                started.push_back(&s);
            }
            int busy = 0;
            for (Slot& s : slots) busy += s.busy ? 1 : 0;
            busy_max = std::max(busy_max, busy);
            lk.unlock();
            std::vector<Slot*> finished;
            for (Slot* s : started)
                if (launch(*s)) finished.push_back(s);
            // 2. collect the launches the device has finished
            for (Slot& s : slots) {
                if (!s.busy || std::find(finished.begin(), finished.end(), &s) != finished.end())
                    continue;
                const hipError_t q = hipEventQuery(s.done);
                if (q == hipErrorNotReady) continue;
                if (q != hipSuccess) {
                    const std::string err = std::string("serving: launch failed: ") + hipGetErrorString(q);
                    for (auto& r : s.batch) fill_error(*r, LC2E_ERR_HIP, err);
                } else {
                    int off = 0;
                    for (auto& r : s.batch) {
                        fill(*r, s.h_out, off);
                        off += r->n;
                    }
                }
                finished.push_back(&s);
            }
            lk.lock();
            for (Slot* s : finished) {
                for (auto& r : s->batch) r->done = true;
                ++launches;
                completed += (long long)s->batch.size();
                poses += s->total;
                in_flight.erase(s->pair);
                s->batch.clear();
                s->busy = false;
                --busy;
            }
            launches += extra_launches;
            extra_launches = 0;
            if (!finished.empty()) cv_done.notify_all();
            // 3. sleep until there is something to do
            if (stopping && queue.empty() && busy == 0) return;
            if (busy > 0) {
                // launches on the device: poll their events, or wake early for new work
                cv_work.wait_for(lk, std::chrono::microseconds(50));
            } else if (held) {
                cv_work.wait_until(lk, wake);
            } else if (queue.empty()) {
                cv_work.wait(lk, [&] { return stopping || !queue.empty(); });
            }
        }
    }
};

extern "C" {

int lc2s_create(lc2e_engine* e, int max_poses, int slots, double window_us, lc2s_server** out) {
    if (!e || !out || max_poses <= 0 || slots < 1 || slots > LC2S_MAX_SLOTS || window_us < 0.0) {
        lc2e::set_last_error("lc2s_create: invalid argument");
        return LC2E_ERR_INVALID;
    }
    auto* s = new lc2s_server();
    s->engine = e;
    s->device = lc2e::engine_device(e);
    s->max_poses = max_poses;
    s->window = std::chrono::duration<double, std::micro>(window_us);
    s->slots.resize(slots);
    if (hipSetDevice(s->device) != hipSuccess) {
        delete s;
        lc2e::set_last_error("lc2s_create: hipSetDevice failed");
        return LC2E_ERR_HIP;
    }
    for (Slot& sl : s->slots) {
        if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess) {
            delete s;
            lc2e::set_last_error("lc2s_create: stream or event creation failed");
            return LC2E_ERR_HIP;
        }
    }
    s->worker = std::thread([s] { s->run(); });
    *out = s;
    return LC2E_OK;
}

void lc2s_destroy(lc2s_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stopping = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    delete s;
}

int lc2s_submit(lc2s_server* s, lc2e_pair* p, int n, const float* mats, long long* ticket) {
    if (!s || !p || !mats || !ticket || n <= 0) {
        lc2e::set_last_error("lc2s_submit: invalid argument");
        return LC2E_ERR_INVALID;
    }
    const int mpp = pair_mats_per_pose(p);
    if (mpp < 0) return LC2E_ERR_INVALID;
    auto r = std::make_shared<Request>();
    r->pair = p;
    r->n = n;
    r->mats.assign(mats, mats + (size_t)n * mpp * 12);
    r->arrived = Clock::now();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stopping) {
            lc2e::set_last_error("lc2s_submit: server is stopping");
            return LC2E_ERR_INVALID;
        }
        r->ticket = s->next_ticket++;
        s->queue.push_back(r);
        s->tickets[r->ticket] = r;
        *ticket = r->ticket;
    }
    s->cv_work.notify_all();
    return LC2E_OK;
}

int lc2s_wait(lc2s_server* s, long long ticket, double timeout_ms, double* lc2,
              double* sum_measure, double* sum_w) {
    if (!s) {
        lc2e::set_last_error("lc2s_wait: invalid argument");
        return LC2E_ERR_INVALID;
    }
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end()) {
        lc2e::set_last_error("lc2s_wait: unknown ticket " + std::to_string(ticket));
        return LC2E_ERR_INVALID;
    }
    RequestPtr r = it->second;
    auto ready = [&] { return r->done; };
    if (timeout_ms < 0.0) {
        s->cv_done.wait(lk, ready);
    } else if (!s->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), ready)) {
        lc2e::set_last_error("lc2s_wait: time-out, request " + std::to_string(ticket) + " still queued");
        return LC2E_ERR_UNSUPPORTED;
    }
    s->tickets.erase(ticket);
    lk.unlock();
    if (lc2) std::memcpy(lc2, r->lc2.data(), sizeof(double) * r->n);
    if (sum_measure) std::memcpy(sum_measure, r->sm.data(), sizeof(double) * r->n);
    if (sum_w) std::memcpy(sum_w, r->sw.data(), sizeof(double) * r->n);
    if (r->status < 0) lc2e::set_last_error(r->error);
    return r->status;
}

int lc2s_stats(lc2s_server* s, lc2s_counters* c) {
    if (!s || !c) {
        lc2e::set_last_error("lc2s_stats: invalid argument");
        return LC2E_ERR_INVALID;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    c->launches = s->launches;
    c->requests = s->completed;
    c->poses = s->poses;
    c->queued = (long long)s->queue.size();
    c->max_in_flight = s->busy_max;
    c->graph_launches = s->graph_launches;
    return LC2E_OK;
}

int lc2s_use_graphs(lc2s_server* s, int enable) {
    if (!s) {
        lc2e::set_last_error("lc2s_use_graphs: invalid argument");
        return LC2E_ERR_INVALID;
    }
    s->use_graphs = enable != 0;
    return LC2E_OK;
}

}  // extern "C"
