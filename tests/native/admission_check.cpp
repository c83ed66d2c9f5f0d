// CPU check of the serving loop's admission policy (n-body_amd/csrc/admission.hpp),
// built with AddressSanitizer and UndefinedBehaviorSanitizer by
// tests/test_admission.py.  Exits nonzero on the first failed expectation.
#include <cstdio>
#include <cstdlib>
#include <deque>
#include <vector>

#include "admission.hpp"

using nbg::plan_admission;
using nbg::Waiting;

static int failures = 0;
#define EXPECT(c)                                                        \
    do {                                                                 \
        if (!(c)) {                                                      \
            std::fprintf(stderr, "%s:%d: expected %s\n", __FILE__, __LINE__, #c); \
            ++failures;                                                  \
        }                                                                \
    } while (0)

static std::vector<Waiting> queue_of(std::vector<long> ns) {
    std::vector<Waiting> q;
    for (long n : ns) q.push_back(Waiting{n, 0});
    return q;
}

int main() {
    {  // everything fits: arrival order
        auto q = queue_of({10, 20, 30});
        auto a = plan_admission(q, 100, 10, 4);
        EXPECT((a == std::vector<size_t>{0, 1, 2}));
    }
    {  // backfill: the 80-body head waits, smaller jobs behind it use the room
        auto q = queue_of({80, 50, 10, 20});
        auto a = plan_admission(q, 60, 10, 4);
        EXPECT((a == std::vector<size_t>{1, 2}));
        EXPECT(q[0].bypassed == 2);  // passed by two admissions
        EXPECT(q[3].bypassed == 0);  // nothing was admitted behind it
    }
    {  // a job bypassed max_bypass times holds the queue
        auto q = queue_of({80, 10});
        q[0].bypassed = 4;
        auto a = plan_admission(q, 60, 10, 4);
        EXPECT(a.empty());
        a = plan_admission(q, 80, 10, 4);  // once it fits it goes first
        EXPECT((a == std::vector<size_t>{0}));
    }
    {  // system slots bound the admissions
        auto q = queue_of({5, 5, 5});
        auto a = plan_admission(q, 100, 2, 4);
        EXPECT((a == std::vector<size_t>{0, 1}));
        a = plan_admission(q, 100, 0, 4);
        EXPECT(a.empty());
    }
    {  // empty queue, zero room
        std::vector<Waiting> q;
        EXPECT(plan_admission(q, 100, 10, 4).empty());
        auto q2 = queue_of({1});
        EXPECT(plan_admission(q2, 0, 10, 4).empty());
    }
    {  // bounded wait: a 90-body job in a 100-body server fed a steady stream
       // of 20-body jobs (each runs one step) is admitted after at most
       // max_bypass bypasses plus the steps its blockers need to finish
        const int max_bypass = 8;
        struct J { long n; int id; };
        std::deque<J> pending;
        std::vector<Waiting> w;
        std::vector<int> bypass_of(1000, 0);
        pending.push_back({90, 0});
        int next = 1;
        long running = 0;
        int admitted_at = -1;
        for (int step = 0; step < 200 && admitted_at < 0; ++step) {
            pending.push_back({20, next++});
            pending.push_back({20, next++});
            std::vector<Waiting> q;
            for (auto &j : pending) q.push_back(Waiting{j.n, bypass_of[j.id]});
            auto a = plan_admission(q, 100 - running, 64, max_bypass);
            std::deque<J> rest;
            size_t k = 0;
            long now = 0;
            for (size_t i = 0; i < pending.size(); ++i) {
                bypass_of[pending[i].id] = q[i].bypassed;
                if (k < a.size() && a[k] == i) {
                    if (pending[i].id == 0) admitted_at = step;
                    now += pending[i].n;
                    ++k;
                } else {
                    rest.push_back(pending[i]);
                }
            }
            pending.swap(rest);
            running = now;  // the jobs admitted this step run it; older ones finished
        }
        EXPECT(admitted_at >= 0);
        EXPECT(admitted_at <= max_bypass + 2);
        EXPECT(bypass_of[0] <= max_bypass);
    }
    if (failures) return 1;
    std::printf("admission policy: all checks passed\n");
    return 0;
}
