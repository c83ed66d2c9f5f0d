// Admission policy of the serving loop (host code, no HIP): which waiting
// jobs join the forest before a step.
//
// Jobs are considered in arrival order.  A job that fits the free bodies and
// system slots is admitted.  A job that does not fit is passed over, so a
// smaller job behind it can use the room (backfill), and every admission
// behind it counts as one bypass.  A job that has been bypassed max_bypass
// times holds the queue: nothing behind it is admitted until it fits.  Since
// running jobs always finish, that bounds how long any job waits.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

namespace nbg {

struct Waiting {
    int64_t n = 0;         // bodies
    int32_t bypassed = 0;  // admissions of later jobs while this one waited
};

// Returns the positions in q (arrival order) of the jobs to admit now and
// updates the bypass counts of the jobs passed over.
inline std::vector<size_t> plan_admission(std::vector<Waiting> &q, int64_t free_bodies, int64_t free_slots,
                                          int32_t max_bypass) {
    std::vector<size_t> admit;
    std::vector<size_t> passed;  // waiting jobs ahead of the current position
    for (size_t i = 0; i < q.size() && free_slots > 0; ++i) {
        if (q[i].n <= free_bodies) {
            admit.push_back(i);
            free_bodies -= q[i].n;
            --free_slots;
            for (size_t k : passed) ++q[k].bypassed;
            continue;
        }
        if (q[i].bypassed >= max_bypass) break;  // this job holds the queue
        passed.push_back(i);
    }
    return admit;
}

}  // namespace nbg
