// Continuous-batching server over batched forests (host code only).
//
// Clients submit jobs (a system plus a step count) from any thread.  A
// worker thread owns each forest (lane): before every step it admits waiting jobs
// (plan_admission: arrival order with backfill, each job bypassed a bounded
// number of times) while the forest's body and system capacity (sized from
// an HBM byte budget) allow, steps every admitted system once, retires the
// systems whose state became non-finite (that job alone fails) and those
// that have reached their step count (final state copied to host memory,
// slot freed).  A job therefore runs alongside whatever else is on the
// device, and its result is the same bits it gets alone.
//
// A job's state can be copied out between two steps into a snapshot buffer
// (checkpoint: the job runs on; suspend: it leaves the server) and queued
// again from one (resume).  Only a lane's worker touches its forest, so a
// snapshot of a resident job is taken by that worker at its next step
// boundary while the client waits.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/nbg.h"
#include "admission.hpp"
#include "host_util.hpp"
#include "launch.hpp"

using namespace nbg;

struct ServeLane;

namespace {

using Clock = std::chrono::steady_clock;

struct Job {
    int64_t id = 0;
    nbg_params p{};
    int64_t n = 0;
    int32_t steps = 0, remaining = 0;
    int32_t steps_before = 0;   // steps run before the snapshot this job resumed from
    ServeLane *lane = nullptr;  // the forest it is queued for
    std::vector<float> st[7];   // x y z m vx vy vz in; x y z vx vy vz out
    int32_t sys = -1;           // forest handle while admitted
    int32_t bypassed = 0;       // later jobs admitted while this one waited
    Clock::time_point t_submit, t_admit, t_done;
    bool done = false;
    int status = NBG_OK;
    std::string error;
};

// A snapshot asked of a job that may be resident in a forest.  The lane's
// worker fills the arrays section of the client's buffer between two steps.
struct SnapRequest {
    std::shared_ptr<Job> job;
    bool leave = false;       // suspend: the job leaves the server
    char *arrays = nullptr;   // the buffer's x y z m vx vy vz section
    int32_t steps_done = 0;
    bool done = false;
    int status = NBG_OK;
    std::string error;
};

// Job snapshot, one contiguous little-endian buffer:
//   [0, 64)    header (SnapHeader)
//   [64, 112)  nbg_params, zero padded
//   [112, ...) float32 arrays x, y, z, m, vx, vy, vz of n values each
// The checksum is FNV-1a 64 over every byte from offset 64 to the end.
struct SnapHeader {
    char magic[8];          // "NBGSNAP1"
    uint32_t version;       // 1
    uint32_t params_bytes;  // sizeof(nbg_params)
    int64_t n;
    int32_t steps_total;    // steps of the job as first submitted
    int32_t steps_done;     // steps run when the snapshot was taken
    uint64_t checksum;
    char reserved[24];
};
static_assert(sizeof(SnapHeader) == 64, "snapshot header layout");
constexpr char kSnapMagic[8] = {'N', 'B', 'G', 'S', 'N', 'A', 'P', '1'};
constexpr int64_t kSnapArrays = 64 + (int64_t)((sizeof(nbg_params) + 15) / 16 * 16);

int64_t snap_bytes(int64_t n) { return kSnapArrays + 7 * (int64_t)sizeof(float) * n; }

uint64_t fnv1a64(const unsigned char *p, int64_t len) {
    uint64_t h = 14695981039346656037ull;
    for (int64_t i = 0; i < len; ++i) h = (h ^ p[i]) * 1099511628211ull;
    return h;
}

// The host copy of a job's state (submitted state while queued, final state
// once finished) into a snapshot's arrays section.
void snap_copy_host(char *arrays, const Job &j) {
    for (int k = 0; k < 7; ++k) std::memcpy(arrays + (size_t)k * j.n * sizeof(float), j.st[k].data(), j.n * sizeof(float));
}

// Header, parameters and checksum around arrays already in place.
void snap_finish(char *buf, const Job &j, int32_t steps_done) {
    std::memset(buf + 64, 0, (size_t)(kSnapArrays - 64));
    std::memcpy(buf + 64, &j.p, sizeof(nbg_params));
    SnapHeader h{};
    std::memcpy(h.magic, kSnapMagic, 8);
    h.version = 1;
    h.params_bytes = (uint32_t)sizeof(nbg_params);
    h.n = j.n;
    h.steps_total = j.steps_before + j.steps;
    h.steps_done = steps_done;
    h.checksum = fnv1a64((const unsigned char *)buf + 64, snap_bytes(j.n) - 64);
    std::memcpy(buf, &h, sizeof h);
}

double secs(Clock::time_point a, Clock::time_point b) { return std::chrono::duration<double>(b - a).count(); }

// Admissions of later jobs a waiting job lets pass before it holds the queue.
constexpr int32_t kMaxBypass = 32;

}  // namespace

// One forest with its own worker thread, HIP stream and queue.  A server has
// a bulk lane and optionally an express lane for small jobs: the lanes step
// concurrently, so a small job advances at the pace of a forest of small
// systems instead of one step per step of the largest resident job.
struct ServeLane {
    nbg_batch *batch = nullptr;
    int64_t cap_n = 0;    // forest body capacity
    int32_t cap_s = 0;    // forest system slots
    int64_t max_job = 0;  // largest job routed to this lane
    std::deque<std::shared_ptr<Job>> pending;
    std::vector<std::shared_ptr<Job>> active;
    std::vector<std::shared_ptr<SnapRequest>> requests;  // snapshots of jobs not queued here
    int64_t used_n = 0;
    std::thread worker;
};

struct nbg_server {
    nbg_ctx *ctx = nullptr;
    std::vector<std::unique_ptr<ServeLane>> lanes;  // [0] bulk, [1] express (optional)
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::map<int64_t, std::shared_ptr<Job>> jobs;
    int64_t next_id = 1;
    bool stop = false;
    int waiters = 0;  // clients inside nbg_server_wait (destroy waits for them to leave)
    // statistics
    Clock::time_point t_start, t_busy_from;
    int busy_lanes = 0;    // lanes inside a step; device-busy time is their union
    double busy_s = 0.0;
    int64_t steps_run = 0, jobs_done = 0;
    double body_steps = 0.0;
    int64_t resident_n = 0, resident_s = 0, peak_bodies = 0, peak_systems = 0;
    std::string worker_error;
};

namespace {

// A suspended job ends here: clients waiting on it get NBG_ERR_STATE and its
// id is forgotten.  Called with S->mu held.
void end_suspended(nbg_server *S, Job &j) {
    j.status = NBG_ERR_STATE;
    j.error = "the job was suspended into a snapshot";
    j.t_done = Clock::now();
    j.done = true;
    S->jobs.erase(j.id);
}

// Fulfils snapshot requests at a step boundary (worker thread).  resident[i]:
// request i's job is in L->active, so its state is in the forest.
void serve_snapshots(nbg_server *S, ServeLane *L, const std::vector<std::shared_ptr<SnapRequest>> &reqs,
                     const std::vector<char> &resident) {
    for (size_t i = 0; i < reqs.size(); ++i) {
        if (!resident[i]) continue;
        SnapRequest &q = *reqs[i];
        const Job &j = *q.job;
        float *a = reinterpret_cast<float *>(q.arrays);
        const int64_t n = j.n;
        const int r = nbg_batch_get(L->batch, j.sys, a, a + n, a + 2 * n, a + 4 * n, a + 5 * n, a + 6 * n, nullptr,
                                    nullptr, nullptr);
        std::memcpy(a + 3 * n, j.st[3].data(), (size_t)n * sizeof(float));  // masses never change
        if (r != NBG_OK) {
            q.status = r;
            q.error = nbg_last_error();
        } else if (q.leave) {
            (void)nbg_batch_remove(L->batch, j.sys);
        }
    }
    std::lock_guard<std::mutex> lk(S->mu);
    for (size_t i = 0; i < reqs.size(); ++i) {
        SnapRequest &q = *reqs[i];
        Job &j = *q.job;
        if (resident[i]) {
            if (q.status == NBG_OK) {
                q.steps_done = j.steps_before + j.steps - j.remaining;
                if (q.leave) {
                    L->active.erase(std::find(L->active.begin(), L->active.end(), q.job));
                    L->used_n -= j.n;
                    S->resident_n -= j.n;
                    S->resident_s -= 1;
                    end_suspended(S, j);
                }
            }
        } else if (j.done && j.status == NBG_OK && !q.leave) {
            // finished before this step boundary: its final state
            snap_copy_host(q.arrays, j);
            q.steps_done = j.steps_before + j.steps;
        } else {
            q.status = NBG_ERR_STATE;
            q.error = !j.done ? "the job is not resident in its forest"
                              : j.status == NBG_OK ? "the job has finished: collect it with nbg_server_wait"
                                                   : "job failed: " + j.error;
        }
        q.done = true;
    }
    S->cv_done.notify_all();
}

void worker_loop(nbg_server *S, ServeLane *L) {
    DeviceGuard g(S->ctx->device);
    for (;;) {
        std::vector<std::shared_ptr<Job>> admit;
        std::vector<std::shared_ptr<SnapRequest>> reqs;
        std::vector<char> resident;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_work.wait(lk, [&] {
                return S->stop || !L->pending.empty() || !L->active.empty() || !L->requests.empty();
            });
            if (S->stop) return;
            std::vector<Waiting> q(L->pending.size());
            for (size_t i = 0; i < q.size(); ++i) q[i] = Waiting{L->pending[i]->n, L->pending[i]->bypassed};
            const std::vector<size_t> pick =
                plan_admission(q, L->cap_n - L->used_n, (int64_t)L->cap_s - (int64_t)L->active.size(), kMaxBypass);
            std::deque<std::shared_ptr<Job>> rest;
            size_t k = 0;
            for (size_t i = 0; i < L->pending.size(); ++i) {
                auto &j = L->pending[i];
                j->bypassed = q[i].bypassed;
                if (k < pick.size() && pick[k] == i) {
                    L->used_n += j->n;
                    admit.push_back(j);
                    ++k;
                } else {
                    rest.push_back(j);
                }
            }
            L->pending.swap(rest);
            reqs.swap(L->requests);
            for (auto &q : reqs)
                resident.push_back(std::find(L->active.begin(), L->active.end(), q->job) != L->active.end());
        }
        if (!reqs.empty()) serve_snapshots(S, L, reqs, resident);
        for (auto &j : admit) {
            int32_t sid = -1;
            const int r = nbg_batch_add(L->batch, &j->p, j->n, j->st[0].data(), j->st[1].data(), j->st[2].data(),
                                        j->st[3].data(), j->st[4].data(), j->st[5].data(), j->st[6].data(), &sid);
            std::lock_guard<std::mutex> lk(S->mu);
            j->t_admit = Clock::now();
            if (r != NBG_OK) {
                j->status = r;
                j->error = nbg_last_error();
                j->done = true;
                j->t_done = j->t_admit;
                L->used_n -= j->n;
                S->cv_done.notify_all();
                continue;
            }
            j->sys = sid;
            L->active.push_back(j);
            S->resident_n += j->n;
            S->resident_s += 1;
        }
        int64_t nb = 0;
        {
            std::lock_guard<std::mutex> lk(S->mu);
            if (L->active.empty()) continue;
            for (auto &j : L->active) nb += j->n;
            S->peak_bodies = std::max(S->peak_bodies, S->resident_n);
            S->peak_systems = std::max(S->peak_systems, S->resident_s);
            if (S->busy_lanes++ == 0) S->t_busy_from = Clock::now();
        }
        int r = nbg_batch_step(L->batch, 1);
        if (r == NBG_OK) r = nbg_batch_synchronize(L->batch);
        // systems whose state became NaN or infinite fail alone and leave
        // before the next build
        std::vector<int32_t> bad(L->active.size() + 1);
        int32_t nbad = 0;
        if (r == NBG_OK) r = nbg_batch_nonfinite(L->batch, bad.data(), (int32_t)bad.size(), &nbad);
        std::vector<std::shared_ptr<Job>> finished, keep;
        {
            std::lock_guard<std::mutex> lk(S->mu);
            if (--S->busy_lanes == 0) S->busy_s += secs(S->t_busy_from, Clock::now());
            S->steps_run += 1;
            S->body_steps += (double)nb;
            if (r != NBG_OK) S->worker_error = nbg_last_error();
            for (auto &j : L->active) {
                if (r != NBG_OK) {
                    j->status = r;
                    j->error = S->worker_error;
                    finished.push_back(j);
                    continue;
                }
                --j->remaining;
                if (std::find(bad.begin(), bad.begin() + std::min<int32_t>(nbad, (int32_t)bad.size()), j->sys) !=
                    bad.begin() + std::min<int32_t>(nbad, (int32_t)bad.size())) {
                    j->status = NBG_ERR_DEVICE;
                    j->error = "the system's state became non-finite after step " +
                               std::to_string(j->steps - j->remaining);
                    finished.push_back(j);
                } else if (j->remaining <= 0) {
                    finished.push_back(j);
                } else {
                    keep.push_back(j);
                }
            }
            L->active = keep;
        }
        for (auto &j : finished) {
            if (j->status == NBG_OK) {
                // out: x y z into st[0..2], vx vy vz into st[4..6]
                const int rr = nbg_batch_get(L->batch, j->sys, j->st[0].data(), j->st[1].data(), j->st[2].data(),
                                             j->st[4].data(), j->st[5].data(), j->st[6].data(), nullptr, nullptr,
                                             nullptr);
                if (rr != NBG_OK) { j->status = rr; j->error = nbg_last_error(); }
            }
            (void)nbg_batch_remove(L->batch, j->sys);
            std::lock_guard<std::mutex> lk(S->mu);
            j->t_done = Clock::now();
            j->done = true;
            L->used_n -= j->n;
            S->resident_n -= j->n;
            S->resident_s -= 1;
            S->jobs_done += 1;
            S->cv_done.notify_all();
        }
    }
}

// Largest body capacity whose forest of max_systems slots fits the budget.
int64_t forest_capacity(int64_t budget, int32_t max_systems) {
    const int64_t fixed = nbg_batch_bytes(1, max_systems);
    const int64_t per = (nbg_batch_bytes(1 << 20, max_systems) - fixed) / ((1 << 20) - 1) + 1;
    int64_t cap = budget > fixed ? (budget - fixed) / per : 0;
    if (cap > kMaxBodies) cap = kMaxBodies;
    while (cap > 0 && nbg_batch_bytes(cap, max_systems) > budget) cap -= 1 + cap / 1000;
    return cap;
}

void free_server(nbg_server *S) {
    for (auto &L : S->lanes)
        if (L->batch) nbg_batch_destroy(L->batch);
    delete S;
}

}  // namespace

extern "C" {

int nbg_server_create_express(nbg_ctx *ctx, int64_t hbm_budget_bytes, int32_t max_systems, int32_t sum_mode,
                              int64_t express_max_bodies, int64_t express_budget_bytes, int32_t express_max_systems,
                              nbg_server **out) {
    if (!ctx || !out) return fail(NBG_ERR_ARG, "ctx or out is NULL");
    *out = nullptr;
    if (max_systems < 1) return fail(NBG_ERR_ARG, "max_systems must be >= 1");
    const bool express = express_max_bodies > 0;
    if (express && (express_max_systems < 1 || express_budget_bytes <= 0 || express_budget_bytes >= hbm_budget_bytes))
        return fail(NBG_ERR_ARG, "the express lane needs system slots and a budget inside the server's");
    const int64_t bulk_budget = express ? hbm_budget_bytes - express_budget_bytes : hbm_budget_bytes;
    const int64_t cap = forest_capacity(bulk_budget, max_systems);
    if (cap < 1) return fail(NBG_ERR_BUDGET, "the HBM budget does not hold a single body");
    int64_t xcap = 0;
    if (express) {
        xcap = forest_capacity(express_budget_bytes, express_max_systems);
        if (xcap < express_max_bodies)
            return fail(NBG_ERR_BUDGET, "the express budget does not hold one job of express_max_bodies");
    }
    nbg_server *S = new nbg_server;
    S->ctx = ctx;
    auto add_lane = [&](int64_t c, int32_t slots, int64_t max_job) {
        auto L = std::make_unique<ServeLane>();
        L->cap_n = c;
        L->cap_s = slots;
        L->max_job = max_job;
        if (int r = nbg_batch_create(ctx, c, slots, sum_mode, &L->batch)) return r;
        nbg_batch_enable_counts(L->batch, 0);
        S->lanes.push_back(std::move(L));
        return (int)NBG_OK;
    };
    int r = add_lane(cap, max_systems, cap);
    if (r == NBG_OK && express) r = add_lane(xcap, express_max_systems, express_max_bodies);
    if (r != NBG_OK) {
        free_server(S);
        return r;
    }
    S->t_start = Clock::now();
    for (auto &L : S->lanes) L->worker = std::thread(worker_loop, S, L.get());
    *out = S;
    return NBG_OK;
}

int nbg_server_create(nbg_ctx *ctx, int64_t hbm_budget_bytes, int32_t max_systems, int32_t sum_mode,
                      nbg_server **out) {
    return nbg_server_create_express(ctx, hbm_budget_bytes, max_systems, sum_mode, 0, 0, 0, out);
}

int nbg_server_destroy(nbg_server *S) {
    if (!S) return NBG_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    for (auto &L : S->lanes)
        if (L->worker.joinable()) L->worker.join();
    {
        // jobs still queued or running end with an error, so that clients
        // blocked in nbg_server_wait return before the server is freed
        std::unique_lock<std::mutex> lk(S->mu);
        const auto now = Clock::now();
        auto abandon = [&](const std::shared_ptr<Job> &j) {
            if (j->done) return;
            j->status = NBG_ERR_STATE;
            j->error = "the server was destroyed before the job finished";
            j->t_done = now;
            j->done = true;
        };
        for (auto &L : S->lanes) {
            for (auto &j : L->pending) abandon(j);
            for (auto &j : L->active) abandon(j);
            for (auto &q : L->requests) {
                q->status = NBG_ERR_STATE;
                q->error = "the server was destroyed before the snapshot was taken";
                q->done = true;
            }
            L->pending.clear();
            L->active.clear();
            L->requests.clear();
        }
        S->cv_done.notify_all();
        S->cv_done.wait(lk, [&] { return S->waiters == 0; });
    }
    free_server(S);
    return NBG_OK;
}

int nbg_server_capacity(nbg_server *S, int64_t *max_bodies, int32_t *max_systems) {
    if (!S) return fail(NBG_ERR_ARG, "server is NULL");
    int64_t nb = 0;
    int32_t ns = 0;
    for (auto &L : S->lanes) {
        nb = std::max(nb, L->max_job);
        ns += L->cap_s;
    }
    if (max_bodies) *max_bodies = nb;
    if (max_systems) *max_systems = ns;
    return NBG_OK;
}

}  // extern "C"

namespace {

int submit_job(nbg_server *S, const nbg_params *p, int64_t n, const float *x, const float *y, const float *z,
               const float *m, const float *vx, const float *vy, const float *vz, int32_t nsteps,
               int32_t steps_before, int64_t *job) {
    if (!S || !p || !job) return fail(NBG_ERR_ARG, "server, params or job is NULL");
    if (n < 1 || !x || !y || !z || !m) return fail(NBG_ERR_ARG, "a job needs n >= 1 positions and masses");
    if (nsteps < 1) return fail(NBG_ERR_ARG, "a job needs at least one step");
    if (p->solver != NBG_LINEAR_BVH || p->morton_bits != 30)
        return fail(NBG_ERR_ARG, "served jobs run the linear-BVH solver with 30-bit keys");
    // the express lane (if any) takes the jobs it is sized for; the bulk lane
    // takes the rest
    ServeLane *L = S->lanes[0].get();
    if (S->lanes.size() > 1 && n <= S->lanes[1]->max_job) L = S->lanes[1].get();
    if (n > L->cap_n) return fail(NBG_ERR_BUDGET, "the job is larger than the server's HBM budget holds");
    // a non-finite input would enter the shared build: refuse it here
    const float *chk[7] = {x, y, z, m, vx, vy, vz};
    for (const float *a : chk)
        if (a)
            for (int64_t i = 0; i < n; ++i)
                if (!std::isfinite(a[i])) return fail(NBG_ERR_ARG, "a job's positions, masses and velocities must be finite");
    auto j = std::make_shared<Job>();
    j->p = *p;
    j->n = n;
    j->steps = j->remaining = nsteps;
    j->steps_before = steps_before;
    j->lane = L;
    const float *src[7] = {x, y, z, m, vx, vy, vz};
    for (int k = 0; k < 7; ++k) {
        if (src[k]) j->st[k].assign(src[k], src[k] + n);
        else j->st[k].assign((size_t)n, 0.0f);
    }
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (!S->worker_error.empty()) return fail(NBG_ERR_DEVICE, "server stopped: " + S->worker_error);
        j->id = S->next_id++;
        j->t_submit = Clock::now();
        S->jobs[j->id] = j;
        L->pending.push_back(j);
        *job = j->id;
    }
    S->cv_work.notify_all();
    return NBG_OK;
}

int snapshot_job(nbg_server *S, int64_t job, bool leave, void *buf, int64_t cap, int64_t *size) {
    if (!S || !size) return fail(NBG_ERR_ARG, "server or size is NULL");
    std::shared_ptr<Job> j;
    int32_t steps_done = 0;
    {
        std::unique_lock<std::mutex> lk(S->mu);
        auto it = S->jobs.find(job);
        if (it == S->jobs.end()) return fail(NBG_ERR_ARG, "unknown job id");
        j = it->second;
        *size = snap_bytes(j->n);
        if (!buf) return NBG_OK;
        if (cap < *size) return fail(NBG_ERR_ARG, "the snapshot buffer is smaller than the job's snapshot");
        if (reinterpret_cast<uintptr_t>(buf) % alignof(float))
            return fail(NBG_ERR_ARG, "the snapshot buffer must be 4-byte aligned");
        char *arrays = static_cast<char *>(buf) + kSnapArrays;
        ServeLane *L = j->lane;
        auto pit = std::find(L->pending.begin(), L->pending.end(), j);
        if (j->done) {
            if (j->status != NBG_OK) return fail(j->status, "job failed: " + j->error);
            if (leave) return fail(NBG_ERR_STATE, "the job has finished: collect it with nbg_server_wait");
            snap_copy_host(arrays, *j);  // its final state
            steps_done = j->steps_before + j->steps;
        } else if (pit != L->pending.end()) {
            // still queued: the submitted state (not admitted while the lock is held)
            snap_copy_host(arrays, *j);
            steps_done = j->steps_before;
            if (leave) {
                L->pending.erase(pit);
                end_suspended(S, *j);
                S->cv_done.notify_all();
            }
        } else {
            // resident (or being admitted): the lane's worker copies it at its
            // next step boundary
            auto q = std::make_shared<SnapRequest>();
            q->job = j;
            q->leave = leave;
            q->arrays = arrays;
            L->requests.push_back(q);
            S->cv_work.notify_all();
            ++S->waiters;
            S->cv_done.wait(lk, [&] { return q->done; });
            --S->waiters;
            // as in nbg_server_wait: S is not touched after this block
            if (S->waiters == 0) S->cv_done.notify_all();
            if (q->status != NBG_OK) return fail(q->status, q->error);
            steps_done = q->steps_done;
        }
    }
    snap_finish(static_cast<char *>(buf), *j, steps_done);
    return NBG_OK;
}

}  // namespace

extern "C" {

int nbg_server_submit(nbg_server *S, const nbg_params *p, int64_t n, const float *x, const float *y, const float *z,
                      const float *m, const float *vx, const float *vy, const float *vz, int32_t nsteps,
                      int64_t *job) {
    return submit_job(S, p, n, x, y, z, m, vx, vy, vz, nsteps, 0, job);
}

int nbg_server_checkpoint(nbg_server *S, int64_t job, void *buf, int64_t cap, int64_t *size) {
    return snapshot_job(S, job, false, buf, cap, size);
}

int nbg_server_suspend(nbg_server *S, int64_t job, void *buf, int64_t cap, int64_t *size) {
    return snapshot_job(S, job, true, buf, cap, size);
}

int nbg_server_resume(nbg_server *S, const void *buf, int64_t size, int64_t *job) {
    if (!S || !buf || !job) return fail(NBG_ERR_ARG, "server, snapshot or job is NULL");
    if (reinterpret_cast<uintptr_t>(buf) % alignof(float))
        return fail(NBG_ERR_ARG, "the snapshot buffer must be 4-byte aligned");
    if (size < kSnapArrays) return fail(NBG_ERR_ARG, "the buffer is too short to hold a job snapshot");
    const char *b = static_cast<const char *>(buf);
    SnapHeader h;
    std::memcpy(&h, b, sizeof h);
    if (std::memcmp(h.magic, kSnapMagic, 8) != 0 || h.version != 1 || h.params_bytes != sizeof(nbg_params))
        return fail(NBG_ERR_ARG, "not a version-1 job snapshot with this library's parameter layout");
    if (h.n < 1 || h.n > (size - kSnapArrays) / (7 * (int64_t)sizeof(float)) || size != snap_bytes(h.n))
        return fail(NBG_ERR_ARG, "the snapshot's size does not match its body count");
    if (h.steps_done < 0 || h.steps_total < 1 || h.steps_done >= h.steps_total)
        return fail(NBG_ERR_ARG, "the snapshot's job has no steps left to run");
    if (fnv1a64(reinterpret_cast<const unsigned char *>(b) + 64, size - 64) != h.checksum)
        return fail(NBG_ERR_ARG, "the snapshot's checksum does not match its contents");
    nbg_params p;
    std::memcpy(&p, b + 64, sizeof p);
    const float *a = reinterpret_cast<const float *>(b + kSnapArrays);
    const int64_t n = h.n;
    return submit_job(S, &p, n, a, a + n, a + 2 * n, a + 3 * n, a + 4 * n, a + 5 * n, a + 6 * n,
                      h.steps_total - h.steps_done, h.steps_done, job);
}

int nbg_server_wait(nbg_server *S, int64_t job, double timeout_s, float *x, float *y, float *z, float *vx, float *vy,
                    float *vz, double *times3) {
    if (!S) return fail(NBG_ERR_ARG, "server is NULL");
    std::shared_ptr<Job> j;
    {
        std::unique_lock<std::mutex> lk(S->mu);
        auto it = S->jobs.find(job);
        if (it == S->jobs.end()) return fail(NBG_ERR_ARG, "unknown job id");
        j = it->second;
        auto ready = [&] { return j->done; };
        ++S->waiters;
        bool ok = true;
        if (timeout_s < 0) S->cv_done.wait(lk, ready);
        else ok = S->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready);
        --S->waiters;
        // the last waiter to leave lets a pending nbg_server_destroy go on; S is
        // not touched after this block
        if (S->waiters == 0) S->cv_done.notify_all();
        if (!ok) return fail(NBG_ERR_STATE, "job not finished yet");
        S->jobs.erase(job);
    }
    if (j->status != NBG_OK) return fail(j->status, "job failed: " + j->error);
    float *dst[6] = {x, y, z, vx, vy, vz};
    const int map[6] = {0, 1, 2, 4, 5, 6};
    for (int k = 0; k < 6; ++k)
        if (dst[k]) std::copy(j->st[map[k]].begin(), j->st[map[k]].end(), dst[k]);
    if (times3) {
        times3[0] = secs(j->t_submit, j->t_admit);  // queueing
        times3[1] = secs(j->t_admit, j->t_done);    // on the device
        times3[2] = secs(j->t_submit, j->t_done);   // latency
    }
    return NBG_OK;
}

// out8: jobs finished, forest steps run, body-steps, device-busy seconds,
// seconds since creation, peak bodies, peak systems, jobs waiting.
int nbg_server_stats(nbg_server *S, double *out8) {
    if (!S || !out8) return fail(NBG_ERR_ARG, "server or out is NULL");
    std::lock_guard<std::mutex> lk(S->mu);
    out8[0] = (double)S->jobs_done;
    out8[1] = (double)S->steps_run;
    out8[2] = S->body_steps;
    out8[3] = S->busy_s;
    out8[4] = secs(S->t_start, Clock::now());
    out8[5] = (double)S->peak_bodies;
    out8[6] = (double)S->peak_systems;
    size_t waiting = 0;
    for (auto &L : S->lanes) waiting += L->pending.size();
    out8[7] = (double)waiting;
    return NBG_OK;
}

}  // extern "C"
