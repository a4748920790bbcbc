// Host-side scenarios for the registrar (lc2s_registrar, lc2serve.h) over fake_runtime.cpp:
// registrations run together end exactly as they do alone, admission by session count and
// by HBM budget, a registration whose pair can never fit fails alone, and destruction
// completes what was accepted. The optimizers are the library's own (bobyqa.cpp,
// optim.cpp); a fake evaluation returns the sum of the pose's matrix floats. Prints one line
// per scenario and exits non-zero on the first failed check. Test infrastructure only.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "lc2serve.h"

extern "C" {
lc2e_engine* fake_engine();
void fake_free_engine(lc2e_engine* e);
void fake_config(int query_delay, long long budget_bytes, int budget_poses);
void fake_config_pairs(long long pair_bytes);
int fake_pairs_max_live();
int fake_pairs_live();
long long fake_used();
}

#define CHECK(cond)                                                                  \
    do {                                                                             \
        if (!(cond)) {                                                               \
            std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);   \
            std::exit(1);                                                            \
        }                                                                            \
    } while (0)

namespace {

// fake volumes are tags: the low byte of the fixed one is its frame count (1: no frames)
lc2e_volume* tag(uintptr_t v) { return reinterpret_cast<lc2e_volume*>(v); }

lc2s_registration job(int k) {
    lc2s_registration r{};
    r.fixed = tag(0x100 + 1);
    r.moving = tag(0x200);
    r.resampled_spacing[0] = r.resampled_spacing[1] = r.resampled_spacing[2] = 1.0f;
    r.patch_radius = 3;
    r.dim = 3;
    r.n_params = 3 + (k % 4);  // 3..6 parameters
    for (int i = 0; i < r.n_params; ++i) {
        r.lower[i] = 1.0;
        r.upper[i] = 2.0 + 0.25 * k;
        r.x0[i] = 1.2 + 0.05 * ((k + i) % 5);
    }
    r.xtol_rel = 1e-6;
    r.direct_maxeval = (k % 3 == 0) ? 30 : 0;
    return r;
}

lc2s_registration sweep_job(int k, std::vector<float>& frames) {
    lc2s_registration r = job(k);
    const int nf = 4;
    r.fixed = tag(0x300 + nf);
    r.n_frames = nf;
    frames.assign((size_t)nf * 12, 0.0f);
    for (int f = 0; f < nf; ++f) {  // frame f: identity rotation, offset along z
        float* F = frames.data() + (size_t)f * 12;
        F[0] = F[5] = F[10] = 1.0f;
        F[3] = 1.0f;
        F[7] = 2.0f;
        F[11] = 3.0f + (float)f;
    }
    r.frames = frames.data();
    r.frames_centre[0] = r.frames_centre[1] = r.frames_centre[2] = 10.0f;
    r.n_params = 6;
    for (int i = 0; i < 6; ++i) {
        r.lower[i] = -2.0;
        r.upper[i] = 2.0;
        r.x0[i] = 0.1 * (i + 1) - 0.3;
    }
    return r;
}

struct Out {
    lc2s_registration_result res{};
    int rc = 0;
};

// run registrations through one registrar with `sessions` at once; submitted all at once
// from `clients` threads, or (clients = 0) one after another
std::vector<Out> run(lc2e_engine* e, const std::vector<lc2s_registration>& jobs, int sessions,
                     int clients, lc2s_registrar_counters* counters = nullptr) {
    lc2s_server* s = nullptr;
    CHECK(lc2s_create(e, 64, 4, 0.0, &s) == LC2E_OK);
    lc2s_registrar* r = nullptr;
    CHECK(lc2s_registrar_create(s, e, sessions, &r) == LC2E_OK);
    std::vector<Out> out(jobs.size());
    if (clients == 0) {
        for (size_t j = 0; j < jobs.size(); ++j) {
            long long t = 0;
            CHECK(lc2s_register_submit(r, &jobs[j], &t) == LC2E_OK);
            out[j].rc = lc2s_register_wait(r, t, -1.0, &out[j].res);
        }
    } else {
        std::vector<std::thread> th;
        for (int c = 0; c < clients; ++c)
            th.emplace_back([&, c] {
                std::vector<std::pair<size_t, long long>> mine;
                for (size_t j = c; j < jobs.size(); j += clients) {
                    long long t = 0;
                    CHECK(lc2s_register_submit(r, &jobs[j], &t) == LC2E_OK);
                    mine.emplace_back(j, t);
                }
                for (auto& jt : mine) out[jt.first].rc = lc2s_register_wait(r, jt.second, -1.0, &out[jt.first].res);
            });
        for (auto& t : th) t.join();
    }
    if (counters) CHECK(lc2s_registrar_stats(r, counters) == LC2E_OK);
    lc2s_registrar_destroy(r);
    lc2s_destroy(s);
    return out;
}

bool same(const Out& a, const Out& b, int n) {
    if (a.rc != b.rc || a.res.status != b.res.status || a.res.evaluations != b.res.evaluations) return false;
    if (std::memcmp(&a.res.lc2, &b.res.lc2, sizeof(double)) != 0) return false;
    return std::memcmp(a.res.x, b.res.x, sizeof(double) * n) == 0;
}

void scenario_mix_equals_alone() {
    fake_config(2, 0, 0);
    fake_config_pairs(1000);
    lc2e_engine* e = fake_engine();
    std::vector<float> fr[3];
    std::vector<lc2s_registration> jobs;
    for (int k = 0; k < 9; ++k) jobs.push_back(job(k));
    for (int k = 0; k < 3; ++k) jobs.push_back(sweep_job(k, fr[k]));
    const std::vector<Out> alone = run(e, jobs, 1, 0);
    lc2s_registrar_counters c{};
    const std::vector<Out> mix = run(e, jobs, 4, 3, &c);
    for (size_t j = 0; j < jobs.size(); ++j) {
        CHECK(alone[j].rc == LC2E_OK && alone[j].res.evaluations > 2 * jobs[j].n_params);
        CHECK(same(alone[j], mix[j], jobs[j].n_params));
    }
    // the fake objective grows with every parameter: the optimum is on the upper bounds
    for (int j = 0; j < 9; ++j)
        for (int i = 0; i < jobs[j].n_params; ++i) CHECK(std::fabs(alone[j].res.x[i] - jobs[j].upper[i]) < 1e-6);
    CHECK(c.completed == (long long)jobs.size() && c.failed == 0 && c.max_running >= 2 && c.max_running <= 4);
    CHECK(fake_pairs_max_live() <= 4 && fake_pairs_live() == 0);
    fake_free_engine(e);
    std::printf("mix_equals_alone ok: %zu registrations, up to %d at once\n", jobs.size(), c.max_running);
}

void scenario_budget_serialises() {
    // room for the server's staging and one pair, not two
    lc2e_engine* e = fake_engine();
    std::vector<lc2s_registration> jobs;
    for (int k = 0; k < 5; ++k) jobs.push_back(job(k));
    fake_config(1, 0, 0);
    fake_config_pairs(1 << 20);
    const std::vector<Out> alone = run(e, jobs, 1, 0);
    fake_config(1, (3 << 20) / 2 + (1 << 16), 0);
    fake_config_pairs(1 << 20);
    lc2s_registrar_counters c{};
    const std::vector<Out> tight = run(e, jobs, 4, 2, &c);
    for (size_t j = 0; j < jobs.size(); ++j) CHECK(same(alone[j], tight[j], jobs[j].n_params));
    CHECK(c.completed == 5 && c.max_running == 1 && c.budget_waits >= 1);
    CHECK(fake_pairs_max_live() == 1 && fake_used() == 0);
    fake_free_engine(e);
    std::printf("budget_serialises ok: %lld deferred admissions\n", c.budget_waits);
}

void scenario_never_fits() {
    lc2e_engine* e = fake_engine();
    fake_config(1, 1 << 16, 0);   // smaller than one pair
    fake_config_pairs(1 << 20);
    std::vector<lc2s_registration> jobs = {job(1), job(2)};
    lc2s_registrar_counters c{};
    const std::vector<Out> out = run(e, jobs, 2, 0, &c);
    for (const Out& o : out) CHECK(o.rc == LC2E_ERR_BUDGET && o.res.status == LC2E_ERR_BUDGET);
    CHECK(c.failed == 2 && c.completed == 0 && c.max_running == 0);
    fake_free_engine(e);
    std::printf("never_fits ok\n");
}

void scenario_destroy_completes_accepted() {
    fake_config(3, 0, 0);
    fake_config_pairs(1000);
    lc2e_engine* e = fake_engine();
    lc2s_server* s = nullptr;
    CHECK(lc2s_create(e, 64, 2, 0.0, &s) == LC2E_OK);
    lc2s_registrar* r = nullptr;
    CHECK(lc2s_registrar_create(s, e, 2, &r) == LC2E_OK);
    std::vector<lc2s_registration> jobs;
    for (int k = 0; k < 6; ++k) jobs.push_back(job(k));
    for (auto& j : jobs) {
        long long t = 0;
        CHECK(lc2s_register_submit(r, &j, &t) == LC2E_OK);
    }
    lc2s_registrar_counters c{};
    lc2s_registrar_destroy(r);  // waits for all six
    (void)c;
    CHECK(fake_pairs_live() == 0);
    lc2s_destroy(s);
    // invalid requests are refused at submission
    CHECK(lc2s_create(e, 64, 1, 0.0, &s) == LC2E_OK);
    CHECK(lc2s_registrar_create(s, e, 1, &r) == LC2E_OK);
    lc2s_registration bad = job(0);
    bad.lower[0] = 3.0;  // above its upper bound
    long long t = 0;
    CHECK(lc2s_register_submit(r, &bad, &t) == LC2E_ERR_INVALID);
    lc2s_registration_result res{};
    CHECK(lc2s_register_wait(r, 12345, 0.0, &res) == LC2E_ERR_INVALID);
    lc2s_registrar_destroy(r);
    lc2s_destroy(s);
    fake_free_engine(e);
    std::printf("destroy_completes_accepted ok\n");
}

}  // namespace

int main() {
    scenario_mix_equals_alone();
    scenario_budget_serialises();
    scenario_never_fits();
    scenario_destroy_completes_accepted();
    std::printf("all registrar scenarios passed\n");
    return 0;
}
