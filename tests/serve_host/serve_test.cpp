// Host-side scenarios for the serving loop (lc2serve.h) over fake_runtime.cpp: merging,
// per-request results and status, error and budget handling, time-outs, admission by
// device share, and destruction with requests in flight. Prints one line per scenario
// and exits non-zero on the first failed check. Test infrastructure only.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "lc2serve.h"

extern "C" {
lc2e_pair* fake_pair(int id, int mpp, double share, int fail_code);
void fake_free_pair(lc2e_pair* p);
lc2e_engine* fake_engine();
void fake_free_engine(lc2e_engine* e);
void fake_config(int query_delay, long long budget_bytes, int budget_poses);
int fake_launches();
int fake_launch(int k, int* pair, int* n, float* first, int cap);
int fake_max_in_flight();
int fake_graph_replays();
void fake_bump_generation(lc2e_pair* p);
long long fake_used();
void fake_fail_host_malloc(int after);
}

#define CHECK(cond)                                                                  \
    do {                                                                             \
        if (!(cond)) {                                                               \
            std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);   \
            std::exit(1);                                                            \
        }                                                                            \
    } while (0)

namespace {

// n poses of mpp matrices; pose i's first float is `tag + i`, the rest 0.01 * k
std::vector<float> poses(int n, int mpp, float tag) {
    std::vector<float> m((size_t)n * mpp * 12);
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < mpp * 12; ++k) m[(size_t)i * mpp * 12 + k] = k ? 0.01f * (float)k : tag + (float)i;
    return m;
}

// what the fake launch returns for one pose
double expect_lc2(const float* m, int mpp) {
    double s = 0.0;
    for (int k = 0; k < mpp * 12; ++k) {
        if (std::isnan(m[k])) return 0.0;
        s += m[k];
    }
    return s;
}

struct Req {
    lc2e_pair* pair;
    int mpp;
    std::vector<float> mats;
    long long ticket = 0;
    std::vector<double> lc2;
    int status = 0;
};

void scenario_merge_and_equality() {
    fake_config(20, 0, 0);
    lc2e_engine* e = fake_engine();
    lc2e_pair* a = fake_pair(1, 1, 0.1, 0);
    lc2e_pair* b = fake_pair(2, 4, 0.1, 0);  // 4 matrices per pose (a frame stack)
    lc2s_server* s = nullptr;
    CHECK(lc2s_create(e, 16, 2, 0.0, &s) == LC2E_OK);
    std::vector<Req> reqs;
    for (int k = 0; k < 40; ++k) {
        const bool pa = k % 3 != 0;
        const int n = 1 + k % 5, mpp = pa ? 1 : 4;
        reqs.push_back({pa ? a : b, mpp, poses(n, mpp, 100.0f * k), 0, {}, 0});
    }
    reqs[7].mats[3] = NAN;  // a poisoned pose
    std::vector<std::thread> th;
    for (int c = 0; c < 4; ++c)
        th.emplace_back([&, c] {
            for (size_t k = c; k < reqs.size(); k += 4) {
                Req& r = reqs[k];
                const int n = (int)(r.mats.size() / (r.mpp * 12));
                CHECK(lc2s_submit(s, r.pair, n, r.mats.data(), &r.ticket) == LC2E_OK);
            }
            for (size_t k = c; k < reqs.size(); k += 4) {
                Req& r = reqs[k];
                const int n = (int)(r.mats.size() / (r.mpp * 12));
                r.lc2.assign(n, -1.0);
                r.status = lc2s_wait(s, r.ticket, -1.0, r.lc2.data(), nullptr, nullptr);
            }
        });
    for (auto& t : th) t.join();
    for (size_t k = 0; k < reqs.size(); ++k) {
        const Req& r = reqs[k];
        const int n = (int)r.lc2.size();
        bool degenerate = false;
        for (int i = 0; i < n; ++i) {
            const float* m = r.mats.data() + (size_t)i * r.mpp * 12;
            const double want = expect_lc2(m, r.mpp);
            if (want == 0.0 && std::isnan(m[3])) degenerate = true;
            CHECK(r.lc2[i] == want);
        }
        CHECK(r.status == (degenerate ? LC2E_OK_DEGENERATE : LC2E_OK));
    }
    lc2s_counters c{};
    CHECK(lc2s_stats(s, &c) == LC2E_OK);
    CHECK(c.requests == 40 && c.queued == 0);
    CHECK(c.launches == fake_launches() && c.launches < 40);
    CHECK(c.max_in_flight >= 1 && c.max_in_flight <= 2);
    // every launch holds at most 16 poses of one pair, and each client's requests of a
    // pair run in the order that client submitted them (pose tag = 100 * request + i;
    // client = request % 4)
    float first[64];
    float last[4][3];
    for (auto& row : last)
        for (float& v : row) v = -1.0f;
    for (int k = 0; k < fake_launches(); ++k) {
        int pid, n;
        CHECK(fake_launch(k, &pid, &n, first, 64) == 0);
        CHECK(n <= 16);
        for (int i = 0; i < n; ++i) {
            const int req = (int)(first[i] / 100.0f), client = req % 4;
            CHECK((req % 3 != 0) == (pid == 1));  // the request's own pair
            CHECK(first[i] > last[client][pid]);
            last[client][pid] = first[i];
        }
    }
    lc2s_destroy(s);
    CHECK(fake_used() == 0);  // the slots' staging went back to the budget
    fake_free_pair(a);
    fake_free_pair(b);
    fake_free_engine(e);
    std::printf("merge_and_equality ok: %lld launches for 40 requests, %d in flight at most\n",
                c.launches, c.max_in_flight);
}

void scenario_errors_and_budget() {
    fake_config(2, 0, 3);  // launches of more than 3 poses fail with the budget error
    lc2e_engine* e = fake_engine();
    lc2e_pair* ok = fake_pair(1, 1, 0.1, 0);
    lc2e_pair* bad = fake_pair(2, 1, 0.1, LC2E_ERR_HIP);
    lc2s_server* s = nullptr;
    CHECK(lc2s_create(e, 64, 1, 100000.0, &s) == LC2E_OK);  // a 0.1 s window: requests merge
    std::vector<std::vector<float>> m;
    std::vector<long long> t(4);
    for (int k = 0; k < 3; ++k) {
        m.push_back(poses(2, 1, 10.0f * k));
        CHECK(lc2s_submit(s, ok, 2, m.back().data(), &t[k]) == LC2E_OK);
    }
    m.push_back(poses(1, 1, 99.0f));
    CHECK(lc2s_submit(s, bad, 1, m.back().data(), &t[3]) == LC2E_OK);
    double l[2];
    CHECK(lc2s_wait(s, t[0], 0.0, l, nullptr, nullptr) == LC2E_ERR_UNSUPPORTED);  // held
    CHECK(std::strstr(lc2e_last_error(), "time-out") != nullptr);
    for (int k = 0; k < 3; ++k) {
        CHECK(lc2s_wait(s, t[k], -1.0, l, nullptr, nullptr) == LC2E_OK);
        CHECK(l[0] == expect_lc2(m[k].data(), 1) && l[1] == expect_lc2(m[k].data() + 12, 1));
    }
    CHECK(lc2s_wait(s, t[3], -1.0, l, nullptr, nullptr) == LC2E_ERR_HIP);
    CHECK(std::strstr(lc2e_last_error(), "fake launch failure of pair 2") != nullptr);
    CHECK(lc2s_wait(s, t[3], 0.0, l, nullptr, nullptr) == LC2E_ERR_INVALID);  // collected
    lc2s_counters c{};
    CHECK(lc2s_stats(s, &c) == LC2E_OK);
    // the 6-pose merged launch was refused by the budget, then run as three of 2 poses
    CHECK(fake_launches() == 3);
    CHECK(c.launches == 1 + 3 + 1);
    lc2s_destroy(s);
    fake_free_pair(ok);
    fake_free_pair(bad);
    fake_free_engine(e);
    std::printf("errors_and_budget ok\n");
}

void scenario_admission_by_share() {
    fake_config(200, 0, 0);
    lc2e_engine* e = fake_engine();
    std::vector<lc2e_pair*> small, big;
    for (int k = 0; k < 4; ++k) small.push_back(fake_pair(10 + k, 1, 0.2, 0));  // 1 pose: 0.2
    for (int k = 0; k < 3; ++k) big.push_back(fake_pair(20 + k, 1, 1.5, 0));    // fills it
    for (auto* group : {&small, &big}) {
        fake_config(200, 0, 0);
        lc2s_server* s = nullptr;
        CHECK(lc2s_create(e, 64, 8, 0.0, &s) == LC2E_OK);
        std::vector<std::vector<float>> m;
        std::vector<long long> t;
        for (size_t k = 0; k < group->size(); ++k) {
            m.push_back(poses(1, 1, (float)k));
            t.push_back(0);
            CHECK(lc2s_submit(s, (*group)[k], 1, m.back().data(), &t.back()) == LC2E_OK);
        }
        double l;
        for (long long tk : t) CHECK(lc2s_wait(s, tk, -1.0, &l, nullptr, nullptr) == LC2E_OK);
        lc2s_counters c{};
        CHECK(lc2s_stats(s, &c) == LC2E_OK);
        if (group == &small) CHECK(c.max_in_flight >= 2 && fake_max_in_flight() <= 4);
        else CHECK(c.max_in_flight == 1 && fake_max_in_flight() == 1);
        std::printf("admission_by_share ok: %s launches, %d in flight at most\n",
                    group == &small ? "small" : "large", c.max_in_flight);
        lc2s_destroy(s);
    }
    for (auto* p : small) fake_free_pair(p);
    for (auto* p : big) fake_free_pair(p);
    fake_free_engine(e);
}

void scenario_destroy_with_work() {
    fake_config(50, 0, 0);
    lc2e_engine* e = fake_engine();
    lc2e_pair* a = fake_pair(1, 1, 0.1, 0);
    lc2s_server* s = nullptr;
    CHECK(lc2s_create(e, 4, 2, 0.0, &s) == LC2E_OK);
    std::vector<float> m = poses(3, 1, 1.0f);
    for (int k = 0; k < 20; ++k) {
        long long t;
        CHECK(lc2s_submit(s, a, 3, m.data(), &t) == LC2E_OK);
    }
    lc2s_destroy(s);  // completes what was accepted, then stops
    CHECK(fake_launches() == 20);
    fake_free_pair(a);
    fake_free_engine(e);
    std::printf("destroy_with_work ok\n");
}

void scenario_staging_failure() {
    // the second pinned staging allocation of a slot fails: that launch's request fails,
    // the slot keeps no half-made staging, and the next request allocates it again
    fake_config(2, 0, 0);
    lc2e_engine* e = fake_engine();
    lc2e_pair* a = fake_pair(1, 1, 0.1, 0);
    lc2s_server* s = nullptr;
    CHECK(lc2s_create(e, 8, 1, 0.0, &s) == LC2E_OK);
    fake_fail_host_malloc(1);
    std::vector<float> m = poses(2, 1, 5.0f);
    long long t;
    double l[2];
    CHECK(lc2s_submit(s, a, 2, m.data(), &t) == LC2E_OK);
    CHECK(lc2s_wait(s, t, -1.0, l, nullptr, nullptr) == LC2E_ERR_HIP);
    CHECK(std::strstr(lc2e_last_error(), "pinned staging") != nullptr);
    CHECK(fake_used() == 0);  // the device staging went back too
    for (int k = 0; k < 2; ++k) {
        CHECK(lc2s_submit(s, a, 2, m.data(), &t) == LC2E_OK);
        CHECK(lc2s_wait(s, t, -1.0, l, nullptr, nullptr) == LC2E_OK);
        CHECK(l[0] == expect_lc2(m.data(), 1) && l[1] == expect_lc2(m.data() + 12, 1));
    }
    lc2s_destroy(s);
    CHECK(fake_used() == 0);
    fake_free_pair(a);
    fake_free_engine(e);
    std::printf("staging_failure ok\n");
}

void scenario_bad_arguments() {
    lc2e_engine* e = fake_engine();
    lc2s_server* s = nullptr;
    CHECK(lc2s_create(e, 0, 1, 0.0, &s) == LC2E_ERR_INVALID);
    CHECK(lc2s_create(e, 4, 0, 0.0, &s) == LC2E_ERR_INVALID);
    CHECK(lc2s_create(e, 4, LC2S_MAX_SLOTS + 1, 0.0, &s) == LC2E_ERR_INVALID);
    CHECK(lc2s_create(e, 4, 1, -1.0, &s) == LC2E_ERR_INVALID);
    CHECK(lc2s_create(e, 4, 1, 0.0, &s) == LC2E_OK);
    long long t;
    float m[12] = {0};
    CHECK(lc2s_submit(s, nullptr, 1, m, &t) == LC2E_ERR_INVALID);
    CHECK(lc2s_wait(s, 777, 0.0, nullptr, nullptr, nullptr) == LC2E_ERR_INVALID);
    CHECK(lc2s_stats(s, nullptr) == LC2E_ERR_INVALID);
    lc2s_destroy(s);
    fake_free_engine(e);
    std::printf("bad_arguments ok\n");
}

}  // namespace

// Graph replay: a (pair, poses) launch runs eagerly once, is captured the second time and
// replayed after that; results stay those of the eager launch; a new generation of the pair
// (buffers reallocated) is captured again; with graphs off every launch is eager.
void scenario_graph_replay() {
    fake_config(2, 0, 0);
    lc2e_engine* e = fake_engine();
    lc2e_pair* a = fake_pair(1, 1, 0.1, 0);
    lc2s_server* s = nullptr;
    CHECK(lc2s_create(e, 16, 1, 0.0, &s) == LC2E_OK);
    CHECK(lc2s_use_graphs(s, 1) == LC2E_OK);  // off by default
    auto run = [&](int k) {
        std::vector<float> m = poses(3, 1, 10.0f * k);
        long long t = 0;
        CHECK(lc2s_submit(s, a, 3, m.data(), &t) == LC2E_OK);
        std::vector<double> l(3, -1.0);
        CHECK(lc2s_wait(s, t, -1.0, l.data(), nullptr, nullptr) == LC2E_OK);
        for (int i = 0; i < 3; ++i) CHECK(l[i] == expect_lc2(m.data() + 12 * i, 1));
    };
    for (int k = 0; k < 6; ++k) run(k);
    lc2s_counters c{};
    CHECK(lc2s_stats(s, &c) == LC2E_OK);
    CHECK(c.graph_launches == 5 && fake_graph_replays() == 5 && fake_launches() == 6);
    fake_bump_generation(a);  // the pair's launch state changed: capture again
    for (int k = 6; k < 9; ++k) run(k);
    CHECK(lc2s_stats(s, &c) == LC2E_OK);
    CHECK(c.graph_launches == 7 && fake_launches() == 9);
    CHECK(lc2s_use_graphs(s, 0) == LC2E_OK);
    for (int k = 9; k < 12; ++k) run(k);
    CHECK(lc2s_stats(s, &c) == LC2E_OK);
    CHECK(c.graph_launches == 7 && fake_launches() == 12);
    lc2s_destroy(s);
    fake_free_pair(a);
    fake_free_engine(e);
    std::printf("graph_replay ok: %lld of 12 launches replayed\n", c.graph_launches);
}

int main() {
    scenario_merge_and_equality();
    scenario_errors_and_budget();
    scenario_admission_by_share();
    scenario_destroy_with_work();
    scenario_staging_failure();
    scenario_bad_arguments();
    scenario_graph_replay();
    std::printf("all serving scenarios passed\n");
    return 0;
}
