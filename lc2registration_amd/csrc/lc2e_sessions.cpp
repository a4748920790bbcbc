// Whole registrations as serving requests (include/lc2serve.h, lc2s_registrar): admission
// by session count and HBM budget, one host thread per running registration, and every
// candidate set of its optimizer evaluated through the serving loop as one request.
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "lc2e_geometry.hpp"
#include "lc2serve.h"
#include "optim.hpp"

namespace lc2e {
void set_last_error(const std::string& msg);
}

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

struct Job {
    long long ticket = 0;
    lc2s_registration reg{};
    std::vector<float> frames;
    Clock::time_point submitted, admitted;
    lc2e_pair* pair = nullptr;
    bool done = false;
    lc2s_registration_result res{};
    std::string error;
};
using JobPtr = std::shared_ptr<Job>;

// A failed evaluation inside the optimizer: its status and lc2e_last_error() text.
struct EvalError : std::runtime_error {
    int status;
    EvalError(int s, const std::string& m) : std::runtime_error(m), status(s) {}
};

}  // namespace

struct lc2s_registrar {
    lc2s_server* srv = nullptr;
    lc2e_engine* e = nullptr;
    int max_sessions = 1;
    std::mutex mu;
    std::condition_variable cv_admit;  // a session ended, a job arrived, or stopping
    std::condition_variable cv_done;   // a job completed
    std::deque<JobPtr> waiting;        // admission queue, submission order
    std::map<long long, JobPtr> jobs;  // not yet collected
    long long next_ticket = 1;
    int running = 0, max_running = 0;
    long long submitted = 0, completed = 0, failed = 0, budget_waits = 0, ended = 0;
    bool stopping = false;
    std::thread admitter;
    std::vector<std::thread> sessions;

    // Pose vector -> the matrices of one pose (mats_per_pose x 12 floats).
    void matrices(const Job& j, const double* x, float* out) const {
        const lc2s_registration& r = j.reg;
        if (r.n_frames > 0) {
            const float T[3] = {(float)x[0], (float)x[1], (float)x[2]};
            const float R[3] = {(float)x[3], (float)x[4], (float)x[5]};
            const float S[3] = {1.0f, 1.0f, 1.0f}, H[3] = {0.0f, 0.0f, 0.0f};
            const float zero[3] = {0.0f, 0.0f, 0.0f};
            double A[9], t[3];
            lc2e::affine_3d_fp64(T, R, S, H, r.frames_centre, zero, A, t);
            for (int k = 0; k < r.n_frames; ++k) {
                const float* F = j.frames.data() + (size_t)k * 12;
                float* M = out + (size_t)k * 12;
                for (int i = 0; i < 3; ++i) {
                    for (int c = 0; c < 4; ++c) {
                        double s = c == 3 ? t[i] : 0.0;
                        for (int m = 0; m < 3; ++m) s += A[i * 3 + m] * (double)F[m * 4 + c];
                        M[i * 4 + c] = (float)s;
                    }
                }
            }
            return;
        }
        const int rc = lc2e_pose_to_matrix(j.pair, x, r.n_params, out);
        if (rc < 0) throw EvalError(rc, lc2e_last_error());
    }

    void session(JobPtr j) {
        const lc2s_registration& r = j->reg;
        int status = LC2E_OK;
        std::string err;
        try {
            lc2e_pair_info info{};
            int rc = lc2e_pair_get_info(j->pair, &info);
            if (rc >= 0 && r.n_frames > 0) {
                if (info.n_frames != r.n_frames)
                    throw EvalError(LC2E_ERR_INVALID, "registration: n_frames differs from the fixed stack's frames");
                rc = lc2e_pair_set_per_frame(j->pair, 1);
            }
            if (rc < 0) throw EvalError(rc, lc2e_last_error());
            const int mpp = r.n_frames > 0 ? r.n_frames : 1;
            int evals = 0;
            std::vector<float> mats;
            std::vector<double> lc2;
            lc2opt::BatchObjective obj = [&](int n, int d, const double* X, double* F) {
                mats.resize((size_t)n * mpp * 12);
                lc2.resize((size_t)n);
                for (int i = 0; i < n; ++i) matrices(*j, X + (size_t)i * d, mats.data() + (size_t)i * mpp * 12);
                long long t = 0;
                int rc2 = lc2s_submit(srv, j->pair, n, mats.data(), &t);
                if (rc2 >= 0) rc2 = lc2s_wait(srv, t, -1.0, lc2.data(), nullptr, nullptr);
                if (rc2 < 0) throw EvalError(rc2, lc2e_last_error());
                for (int i = 0; i < n; ++i)
                    F[i] = lc2[i] > 0.0 ? -std::log(lc2[i]) : std::numeric_limits<double>::infinity();
                evals += n;
            };
            const int d = r.n_params;
            std::vector<double> lb(r.lower, r.lower + d), ub(r.upper, r.upper + d), x(d);
            for (int i = 0; i < d; ++i) x[i] = std::min(std::max(r.x0[i], lb[i]), ub[i]);
            double f = std::numeric_limits<double>::infinity();
            if (r.direct_maxeval > 0) {
                const lc2opt::Result g = lc2opt::direct_l(obj, lb, ub, r.direct_maxeval, 1e-4, true);
                if (g.code < 0) throw EvalError(LC2E_ERR_INVALID, "registration: DIRECT-L: " + g.message);
                x = g.x;
                f = g.f;
            }
            const lc2opt::Result b = lc2opt::bobyqa(obj, x, lb, ub, lc2opt::default_initial_step(x, lb, ub),
                                                    r.xtol_rel, r.maxeval);
            if (b.code < 0) throw EvalError(LC2E_ERR_INVALID, "registration: BOBYQA: " + b.message);
            if (b.f <= f) {
                x = b.x;
                f = b.f;
            }
            for (int i = 0; i < d; ++i) j->res.x[i] = x[i];
            j->res.lc2 = std::isfinite(f) ? std::exp(-f) : 0.0;
            j->res.evaluations = evals;
        } catch (const EvalError& ex) {
            status = ex.status;
            err = ex.what();
        } catch (const std::exception& ex) {
            status = LC2E_ERR_HIP;
            err = ex.what();
        }
        lc2e_pair_destroy(j->pair);
        j->pair = nullptr;
        std::lock_guard<std::mutex> lk(mu);
        j->res.status = status;
        j->res.run_ms = ms_between(j->admitted, Clock::now());
        j->error = err;
        j->done = true;
        --running;
        ++ended;
        if (status < 0) ++failed;
        else ++completed;
        cv_done.notify_all();
        cv_admit.notify_all();
    }

    void finish_unadmitted(const JobPtr& j, int status, const std::string& err) {
        // with mu held
        j->res.status = status;
        j->res.admit_ms = ms_between(j->submitted, Clock::now());
        j->error = err;
        j->done = true;
        ++failed;
        cv_done.notify_all();
    }

    // Admission in submission order: the head job starts when a session is free and its
    // pair fits the budget; while it does not fit and others run, it waits for one to end.
    void admit_loop() {
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            cv_admit.wait(lk, [&] { return (!waiting.empty() && running < max_sessions) ||
                                           (stopping && waiting.empty()); });
            if (waiting.empty()) return;  // stopping
            JobPtr j = waiting.front();
            const long long ended_before = ended;
            const int running_before = running;
            lk.unlock();
            const lc2s_registration& r = j->reg;
            lc2e_pair* p = nullptr;
            const int rc = lc2e_pair_create(e, r.fixed, r.moving, r.resampled_spacing,
                                            r.high_res_is_moving, r.center_overlaid,
                                            r.patch_radius, r.dim, &p);
            const std::string err = rc < 0 ? lc2e_last_error() : "";
            lk.lock();
            if (rc == LC2E_ERR_BUDGET && (running_before > 0 || running > 0)) {
                // wait for a running registration to end and free its pair, then retry
                ++budget_waits;
                cv_admit.wait(lk, [&] { return ended != ended_before || running == 0; });
                continue;
            }
            waiting.pop_front();
            if (rc < 0) {
                finish_unadmitted(j, rc, err);
                continue;
            }
            j->pair = p;
            j->admitted = Clock::now();
            j->res.admit_ms = ms_between(j->submitted, j->admitted);
            ++running;
            max_running = std::max(max_running, running);
            sessions.emplace_back(&lc2s_registrar::session, this, j);
        }
    }
};

extern "C" {

int lc2s_registrar_create(lc2s_server* s, lc2e_engine* e, int max_sessions, lc2s_registrar** out) {
    if (!s || !e || !out || max_sessions < 1) {
        lc2e::set_last_error("lc2s_registrar_create: invalid argument");
        return LC2E_ERR_INVALID;
    }
    auto* r = new lc2s_registrar();
    r->srv = s;
    r->e = e;
    r->max_sessions = max_sessions;
    r->admitter = std::thread(&lc2s_registrar::admit_loop, r);
    *out = r;
    return LC2E_OK;
}

void lc2s_registrar_destroy(lc2s_registrar* r) {
    if (!r) return;
    {
        std::lock_guard<std::mutex> lk(r->mu);
        r->stopping = true;
    }
    r->cv_admit.notify_all();
    r->admitter.join();
    // the admitter has started every queued job; wait for the sessions
    for (std::thread& t : r->sessions) t.join();
    delete r;
}

int lc2s_register_submit(lc2s_registrar* r, const lc2s_registration* reg, long long* ticket) {
    if (!r || !reg || !ticket || !reg->fixed || !reg->moving || reg->n_params < 1 ||
        reg->n_params > 12 || reg->n_frames < 0 || (reg->n_frames > 0 && (!reg->frames || reg->n_params != 6))) {
        lc2e::set_last_error("lc2s_register_submit: invalid argument");
        return LC2E_ERR_INVALID;
    }
    for (int i = 0; i < reg->n_params; ++i)
        if (!(reg->lower[i] <= reg->upper[i])) {
            lc2e::set_last_error("lc2s_register_submit: a lower bound exceeds its upper bound");
            return LC2E_ERR_INVALID;
        }
    auto j = std::make_shared<Job>();
    j->reg = *reg;
    if (reg->n_frames > 0) {
        j->frames.assign(reg->frames, reg->frames + (size_t)reg->n_frames * 12);
        j->reg.frames = j->frames.data();
    }
    j->submitted = Clock::now();
    {
        std::lock_guard<std::mutex> lk(r->mu);
        if (r->stopping) {
            lc2e::set_last_error("lc2s_register_submit: registrar is stopping");
            return LC2E_ERR_INVALID;
        }
        j->ticket = r->next_ticket++;
        r->jobs[j->ticket] = j;
        r->waiting.push_back(j);
        ++r->submitted;
        *ticket = j->ticket;
    }
    r->cv_admit.notify_all();
    return LC2E_OK;
}

int lc2s_register_wait(lc2s_registrar* r, long long ticket, double timeout_ms,
                       lc2s_registration_result* out) {
    if (!r) {
        lc2e::set_last_error("lc2s_register_wait: null registrar");
        return LC2E_ERR_INVALID;
    }
    std::unique_lock<std::mutex> lk(r->mu);
    auto it = r->jobs.find(ticket);
    if (it == r->jobs.end()) {
        lc2e::set_last_error("lc2s_register_wait: unknown ticket");
        return LC2E_ERR_INVALID;
    }
    JobPtr j = it->second;
    auto ready = [&] { return j->done; };
    if (timeout_ms < 0.0) {
        r->cv_done.wait(lk, ready);
    } else if (!r->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), ready)) {
        lc2e::set_last_error("lc2s_register_wait: time-out");
        return LC2E_ERR_UNSUPPORTED;
    }
    r->jobs.erase(ticket);
    if (out) *out = j->res;
    if (j->res.status < 0) lc2e::set_last_error(j->error);
    return j->res.status;
}

int lc2s_registrar_stats(lc2s_registrar* r, lc2s_registrar_counters* out) {
    if (!r || !out) {
        lc2e::set_last_error("lc2s_registrar_stats: invalid argument");
        return LC2E_ERR_INVALID;
    }
    std::lock_guard<std::mutex> lk(r->mu);
    out->submitted = r->submitted;
    out->completed = r->completed;
    out->failed = r->failed;
    out->running = r->running;
    out->max_running = r->max_running;
    out->budget_waits = r->budget_waits;
    return LC2E_OK;
}

}  // extern "C"
