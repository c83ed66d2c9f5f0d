// CPU check of the continuous-batching server's threading
// (n-body_amd/csrc/serve.cpp), built with ThreadSanitizer by
// tests/test_serve_sanitizers.py.  The forest calls the server makes
// (nbg_batch_*) are served here by a host-memory forest with a simple
// deterministic step, so the server's own code runs unchanged: lane worker
// threads, admission, retirement, the non-finite guard, waiters with and
// without time-outs, statistics read while jobs run, destroy with a client
// still waiting, and job snapshots (checkpoint, suspend, resume) of queued,
// resident and finished jobs.  Exits nonzero on the first failed expectation (a
// data race aborts the run by itself).
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <random>
#include <thread>
#include <vector>

#include "host_util.hpp"

// ---------------------------------------------------------------- host forest

extern "C" {

hipError_t hipGetDevice(int *d) {
    *d = 0;
    return hipSuccess;
}
hipError_t hipSetDevice(int) { return hipSuccess; }

const char *nbg_last_error(void) { return nbg::g_err.c_str(); }

}  // extern "C"

namespace {

struct HostSystem {
    std::vector<float> s[7];  // x y z m vx vy vz
    float dt = 0.0f;
};

// The host step: a harmonic kick toward the origin, then the drift (each
// system alone; a body of negative mass turns its system non-finite).
void host_step(HostSystem &h) {
    const size_t n = h.s[3].size();
    for (size_t i = 0; i < n; ++i) {
        for (int d = 0; d < 3; ++d) {
            h.s[4 + d][i] = h.s[4 + d][i] - h.s[d][i] * h.dt;
            h.s[d][i] = h.s[d][i] + h.s[4 + d][i] * h.dt;
        }
        if (h.s[3][i] < 0.0f) h.s[0][i] = NAN;
    }
}

}  // namespace

struct nbg_batch {
    int64_t cap_n = 0;
    int32_t cap_s = 0;
    int32_t next = 0;
    std::map<int32_t, HostSystem> sys;
};

extern "C" {

int64_t nbg_batch_bytes(int64_t max_bodies, int32_t max_systems) { return 4096 + 326 * max_bodies + 64 * max_systems; }

int nbg_batch_create(nbg_ctx *, int64_t max_bodies, int32_t max_systems, int32_t, nbg_batch **out) {
    nbg_batch *b = new nbg_batch;
    b->cap_n = max_bodies;
    b->cap_s = max_systems;
    *out = b;
    return NBG_OK;
}

int nbg_batch_destroy(nbg_batch *b) {
    delete b;
    return NBG_OK;
}

int nbg_batch_enable_counts(nbg_batch *, int32_t) { return NBG_OK; }

int nbg_batch_add(nbg_batch *b, const nbg_params *p, int64_t n, const float *x, const float *y, const float *z,
                  const float *m, const float *vx, const float *vy, const float *vz, int32_t *id) {
    if ((int32_t)b->sys.size() >= b->cap_s) return nbg::fail(NBG_ERR_BUDGET, "no system slot");
    HostSystem h;
    const float *src[7] = {x, y, z, m, vx, vy, vz};
    for (int k = 0; k < 7; ++k) h.s[k].assign(src[k], src[k] + n);
    h.dt = p->dt;
    *id = b->next++;
    b->sys[*id] = std::move(h);
    return NBG_OK;
}

int nbg_batch_remove(nbg_batch *b, int32_t id) {
    b->sys.erase(id);
    return NBG_OK;
}

int nbg_batch_get(nbg_batch *b, int32_t id, float *x, float *y, float *z, float *vx, float *vy, float *vz, float *,
                  float *, float *) {
    auto it = b->sys.find(id);
    if (it == b->sys.end()) return nbg::fail(NBG_ERR_ARG, "no such system");
    float *dst[6] = {x, y, z, vx, vy, vz};
    const int map[6] = {0, 1, 2, 4, 5, 6};
    for (int k = 0; k < 6; ++k)
        if (dst[k]) std::copy(it->second.s[map[k]].begin(), it->second.s[map[k]].end(), dst[k]);
    return NBG_OK;
}

int nbg_batch_step(nbg_batch *b, int32_t nsteps) {
    for (int k = 0; k < nsteps; ++k)
        for (auto &kv : b->sys) host_step(kv.second);
    return NBG_OK;
}

int nbg_batch_synchronize(nbg_batch *) {
    std::this_thread::sleep_for(std::chrono::microseconds(50));  // a device step takes time
    return NBG_OK;
}

int nbg_batch_nonfinite(nbg_batch *b, int32_t *ids, int32_t cap, int32_t *count) {
    int32_t c = 0;
    for (auto &kv : b->sys) {
        bool bad = false;
        for (int k = 0; k < 7 && !bad; ++k)
            for (float v : kv.second.s[k])
                if (!std::isfinite(v)) { bad = true; break; }
        if (bad) {
            if (c < cap) ids[c] = kv.first;
            ++c;
        }
    }
    *count = c;
    return NBG_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- the test

static std::atomic<int> failures{0};
static std::atomic<int> polls_pending{0}, jobs_failed{0}, jobs_checked{0};
#define EXPECT(c, ...)                                                           \
    do {                                                                         \
        if (!(c)) {                                                              \
            std::fprintf(stderr, "%s:%d: expected %s: ", __FILE__, __LINE__, #c); \
            std::fprintf(stderr, __VA_ARGS__);                                   \
            std::fprintf(stderr, "\n");                                          \
            ++failures;                                                          \
        }                                                                        \
    } while (0)

struct Job {
    HostSystem h;
    int32_t steps = 0;
    bool blows_up = false;
};

static Job make_job(std::mt19937 &g, int64_t n, int32_t steps, bool blow) {
    std::uniform_real_distribution<float> u(-1.0f, 1.0f);
    Job j;
    for (auto &v : j.h.s) v.resize(n);
    for (int64_t i = 0; i < n; ++i) {
        for (int d = 0; d < 3; ++d) { j.h.s[d][i] = u(g); j.h.s[4 + d][i] = 0.1f * u(g); }
        j.h.s[3][i] = 1.0f;
    }
    if (blow) j.h.s[3][n / 2] = -1.0f;
    j.h.dt = 1e-3f;
    j.steps = steps;
    j.blows_up = blow;
    return j;
}

static int64_t submit(nbg_server *S, const Job &j) {
    nbg_params p{};  // the server reads only the solver, the key width and dt
    p.solver = NBG_LINEAR_BVH;
    p.morton_bits = 30;
    p.dt = j.h.dt;
    int64_t id = -1;
    const int r = nbg_server_submit(S, &p, (int64_t)j.h.s[3].size(), j.h.s[0].data(), j.h.s[1].data(),
                                    j.h.s[2].data(), j.h.s[3].data(), j.h.s[4].data(), j.h.s[5].data(),
                                    j.h.s[6].data(), j.steps, &id);
    EXPECT(r == NBG_OK, "submit: %s", nbg_last_error());
    return id;
}

// A client: submits its jobs with small gaps, polls some with a zero
// time-out first, then waits for all and checks each result against the
// same host steps applied alone.
static void client(nbg_server *S, int seed, int njobs) {
    std::mt19937 g(seed);
    std::uniform_int_distribution<int> nd(1, 3000), sd(1, 40);
    std::vector<Job> jobs;
    std::vector<int64_t> ids;
    for (int k = 0; k < njobs; ++k) {
        const bool blow = (seed + k) % 7 == 0;
        jobs.push_back(make_job(g, nd(g), sd(g), blow));
        ids.push_back(submit(S, jobs.back()));
        std::this_thread::sleep_for(std::chrono::microseconds(200 * (k % 3)));
    }
    for (size_t k = 0; k < jobs.size(); ++k) {
        const Job &j = jobs[k];
        const int64_t n = (int64_t)j.h.s[3].size();
        std::vector<float> out[6];
        for (auto &v : out) v.resize(n);
        double t3[3];
        if (k % 2 == 0) {  // poll first: may or may not be finished
            const int r = nbg_server_wait(S, ids[k], 0.0, out[0].data(), out[1].data(), out[2].data(),
                                          out[3].data(), out[4].data(), out[5].data(), t3);
            if (r == NBG_ERR_STATE && std::string(nbg_last_error()).find("not finished") != std::string::npos) {
                ++polls_pending;  // still running: wait below
            } else {
                EXPECT(j.blows_up ? r == NBG_ERR_DEVICE : r == NBG_OK, "poll of job %lld: %d %s",
                       (long long)ids[k], r, nbg_last_error());
                continue;
            }
        }
        const int r = nbg_server_wait(S, ids[k], 60.0, out[0].data(), out[1].data(), out[2].data(), out[3].data(),
                                      out[4].data(), out[5].data(), t3);
        if (j.blows_up) {
            EXPECT(r == NBG_ERR_DEVICE, "job %lld should fail alone: %d %s", (long long)ids[k], r, nbg_last_error());
            ++jobs_failed;
            continue;
        }
        EXPECT(r == NBG_OK, "job %lld: %d %s", (long long)ids[k], r, nbg_last_error());
        if (r != NBG_OK) continue;
        HostSystem ref = j.h;
        for (int s = 0; s < j.steps; ++s) host_step(ref);
        const int map[6] = {0, 1, 2, 4, 5, 6};
        for (int c = 0; c < 6; ++c)
            for (int64_t i = 0; i < n; ++i)
                if (out[c][i] != ref.s[map[c]][i]) {
                    EXPECT(false, "job %lld component %d body %lld", (long long)ids[k], c, (long long)i);
                    break;
                }
        ++jobs_checked;
        EXPECT(t3[2] >= t3[1] && t3[1] >= 0.0, "job %lld times %g %g %g", (long long)ids[k], t3[0], t3[1], t3[2]);
    }
}

static std::vector<char> take(nbg_server *S, int64_t id, bool leave, int *status = nullptr) {
    int64_t size = 0;
    int r = leave ? nbg_server_suspend(S, id, nullptr, 0, &size) : nbg_server_checkpoint(S, id, nullptr, 0, &size);
    EXPECT(r == NBG_OK && size > 112, "size query of job %lld: %d", (long long)id, r);
    std::vector<char> buf((size_t)size + 8);
    char *b = buf.data() + (8 - reinterpret_cast<uintptr_t>(buf.data()) % 8) % 8;  // 8-byte aligned
    r = leave ? nbg_server_suspend(S, id, b, size, &size) : nbg_server_checkpoint(S, id, b, size, &size);
    if (status) *status = r;
    else EXPECT(r == NBG_OK, "snapshot of job %lld: %d %s", (long long)id, r, nbg_last_error());
    return r == NBG_OK ? std::vector<char>(b, b + size) : std::vector<char>();
}

static int64_t resume(nbg_server *S, const std::vector<char> &snap, int expect = NBG_OK) {
    std::vector<uint64_t> al((snap.size() + 7) / 8);  // aligned copy
    std::memcpy(al.data(), snap.data(), snap.size());
    int64_t id = -1;
    const int r = nbg_server_resume(S, al.data(), (int64_t)snap.size(), &id);
    EXPECT(r == expect, "resume: %d (expected %d) %s", r, expect, nbg_last_error());
    return id;
}

static void snap_steps(const std::vector<char> &snap, int32_t *total, int32_t *done) {
    std::memcpy(total, snap.data() + 24, 4);
    std::memcpy(done, snap.data() + 28, 4);
}

// The full state after the job's steps, checked against the host steps
// applied alone without interruption.
static void expect_result(nbg_server *S, int64_t id, const Job &j, const char *what) {
    const int64_t n = (int64_t)j.h.s[3].size();
    std::vector<float> out[6];
    for (auto &v : out) v.resize(n);
    const int r = nbg_server_wait(S, id, 120.0, out[0].data(), out[1].data(), out[2].data(), out[3].data(),
                                  out[4].data(), out[5].data(), nullptr);
    EXPECT(r == NBG_OK, "%s: %d %s", what, r, nbg_last_error());
    if (r != NBG_OK) return;
    HostSystem ref = j.h;
    for (int s = 0; s < j.steps; ++s) host_step(ref);
    const int map[6] = {0, 1, 2, 4, 5, 6};
    for (int c = 0; c < 6; ++c)
        EXPECT(std::memcmp(out[c].data(), ref.s[map[c]].data(), n * sizeof(float)) == 0, "%s: component %d", what, c);
}

static void snapshot_checks(nbg_ctx *ctx) {
    nbg_server *S = nullptr;
    // one lane of 20,000 bodies: a 15,000-body job keeps a 9,000-body one queued
    EXPECT(nbg_server_create(ctx, nbg_batch_bytes(20000, 4), 4, NBG_SUM_FAST, &S) == NBG_OK, "create");
    std::mt19937 g(11);
    const Job big = make_job(g, 15000, 400, false), queued = make_job(g, 9000, 30, false);
    const int64_t ib = submit(S, big), iq = submit(S, queued);
    // a queued job: its submitted state, steps_done 0; suspended and resumed
    std::vector<char> sq = take(S, iq, true);
    int32_t tot = 0, done = -1;
    snap_steps(sq, &tot, &done);
    EXPECT(tot == 30 && done == 0, "queued snapshot steps %d/%d", done, tot);
    std::vector<float> o(9000);
    EXPECT(nbg_server_wait(S, iq, 0.0, o.data(), nullptr, nullptr, nullptr, nullptr, nullptr, nullptr) ==
               NBG_ERR_ARG, "a suspended job's id is forgotten");
    // a resident job: checkpoint (runs on) and later suspend; both resumed
    while (true) {
        double st[8];
        nbg_server_stats(S, st);
        if (st[1] >= 3) break;  // a few forest steps have run
        std::this_thread::yield();
    }
    std::vector<char> c1 = take(S, ib, false);
    snap_steps(c1, &tot, &done);
    EXPECT(tot == 400 && done > 0 && done < 400, "resident checkpoint steps %d/%d", done, tot);
    const int64_t ic = resume(S, c1);
    std::vector<char> s1 = take(S, ib, true);  // the original leaves
    int32_t done2 = 0;
    snap_steps(s1, &tot, &done2);
    EXPECT(done2 >= done && done2 < 400, "suspend after checkpoint steps %d then %d", done, done2);
    const int64_t is = resume(S, s1);
    const int64_t iq2 = resume(S, sq);
    // a resumed job checkpointed again carries the steps before its snapshot
    std::vector<char> c2 = take(S, ic, false);
    int32_t done3 = 0;
    snap_steps(c2, &tot, &done3);
    EXPECT(tot == 400 && done3 >= done, "checkpoint of a resumed job: %d/%d", done3, tot);
    const int64_t ic2 = resume(S, c2);
    // a finished job: checkpoint gives its final state (steps_done ==
    // steps_total), which has no steps left to resume; suspend is refused
    const Job one = make_job(g, 10, 1, false);
    const int64_t ifin = submit(S, one);
    bool finished = false;
    for (int k = 0; k < 20000 && !finished; ++k) {
        int r = 0;
        std::vector<char> f = take(S, ifin, false, &r);
        EXPECT(r == NBG_OK, "checkpoint of a 1-step job: %d %s", r, nbg_last_error());
        if (r != NBG_OK) break;
        snap_steps(f, &tot, &done);
        if (tot == done) {
            finished = true;
            HostSystem ref = one.h;
            host_step(ref);
            EXPECT(std::memcmp(f.data() + 112, ref.s[0].data(), 40) == 0, "final x in a finished job's snapshot");
            resume(S, f, NBG_ERR_ARG);
            take(S, ifin, true, &r);
            EXPECT(r == NBG_ERR_STATE, "suspend of a finished job: %d", r);
        } else {
            std::this_thread::sleep_for(std::chrono::milliseconds(1));
        }
    }
    EXPECT(finished, "the 1-step job finished");
    expect_result(S, ifin, one, "a finished job after its checkpoint");
    // damaged and foreign buffers are refused
    std::vector<char> bad = sq;
    bad[200] ^= 1;
    resume(S, bad, NBG_ERR_ARG);
    bad = sq;
    bad.resize(bad.size() - 4);
    resume(S, bad, NBG_ERR_ARG);
    bad = sq;
    bad[0] = 'X';
    resume(S, bad, NBG_ERR_ARG);
    // every resumed job ends with the uninterrupted bits
    expect_result(S, ic, big, "resumed checkpoint");
    expect_result(S, is, big, "resumed suspension");
    expect_result(S, ic2, big, "resumed checkpoint of a resumed job");
    expect_result(S, iq2, queued, "resumed queued job");
    nbg_server_destroy(S);
}

int main() {
    nbg_ctx ctx;
    snapshot_checks(&ctx);
    const int64_t budget = nbg_batch_bytes(20000, 16), xbytes = nbg_batch_bytes(6000, 32);
    for (int express = 0; express < 2; ++express) {
        nbg_server *S = nullptr;
        int r = express ? nbg_server_create_express(&ctx, budget + xbytes, 16, NBG_SUM_FAST, 1000, xbytes, 32, &S)
                        : nbg_server_create(&ctx, budget, 16, NBG_SUM_FAST, &S);
        EXPECT(r == NBG_OK && S, "create: %s", nbg_last_error());
        if (!S) return 1;
        std::atomic<bool> done{false};
        std::thread watcher([&] {  // statistics read while the workers run
            double st[8];
            while (!done.load()) {
                EXPECT(nbg_server_stats(S, st) == NBG_OK, "stats");
                std::this_thread::sleep_for(std::chrono::microseconds(300));
            }
        });
        std::vector<std::thread> clients;
        for (int c = 0; c < 4; ++c) clients.emplace_back(client, S, 100 * express + c + 1, 12);
        for (auto &t : clients) t.join();
        done = true;
        watcher.join();
        double st[8];
        nbg_server_stats(S, st);
        EXPECT(st[0] == 48.0, "jobs done %g", st[0]);
        // destroy with a client blocked on a long job
        std::mt19937 g(7);
        Job longjob = make_job(g, 2000, 1000000, false);
        const int64_t id = submit(S, longjob);
        std::atomic<int> waited{1};
        std::atomic<bool> entering{false};
        std::thread blocked([&] {
            std::vector<float> o(2000);
            entering = true;
            waited = nbg_server_wait(S, id, -1.0, o.data(), nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
        });
        // no call may follow destroy, so give the client time to be inside
        // its wait before the server goes away
        while (!entering.load()) std::this_thread::yield();
        std::this_thread::sleep_for(std::chrono::milliseconds(200));
        EXPECT(nbg_server_destroy(S) == NBG_OK, "destroy");
        blocked.join();
        EXPECT(waited.load() == NBG_ERR_STATE, "the blocked waiter got %d", waited.load());
    }
    std::printf("jobs checked %d, failed alone %d, polls that found a job running %d\n", jobs_checked.load(),
                jobs_failed.load(), polls_pending.load());
    EXPECT(polls_pending.load() > 0 && jobs_failed.load() > 0, "the run exercised polls and non-finite jobs");
    if (failures) {
        std::printf("%d failed expectations\n", failures.load());
        return 1;
    }
    std::printf("all checks passed\n");
    return 0;
}
