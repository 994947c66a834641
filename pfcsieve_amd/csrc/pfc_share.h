// pfc_share.h -- how a batcher round sized by work divides its updates among
// the resident segments (host only, no HIP; shared with the CPU tests).
#ifndef PFC_SHARE_H
#define PFC_SHARE_H

#include <stddef.h>
#include <stdint.h>

#include <algorithm>
#include <numeric>
#include <vector>

namespace pfc {

// Steps of a round of about `budget` p*n updates over n segments, segment i
// costing cost[i] updates per multiplier step and taking at most cap[i] steps
// (what it has left, bounded by the steps one wavefront walks in a round).
// Every segment gets the same share of the updates, and a segment that needs
// less than its share takes all it needs, the rest going to the others
// (water-filling):
//   steps[i] = min(cap[i], L / cost[i]),  sum_i min(cost[i] * cap[i], L) = budget,
// and at least min(min_steps, cap[i]).  A segment of few candidates thereby
// finishes in a round or a few whatever else is resident; dividing the round
// by candidates instead (the same step count for every segment) made such a
// segment wait for as many rounds as the largest one.  cost[i] == 0 gives
// cap[i].
inline void share_steps(const uint64_t* cost, const uint64_t* cap, size_t n, uint64_t budget,
                        uint64_t min_steps, uint64_t* steps) {
    std::vector<size_t> order(n);
    std::iota(order.begin(), order.end(), (size_t)0);
    auto demand = [&](size_t i) -> unsigned __int128 { return (unsigned __int128)cost[i] * cap[i]; };
    std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return demand(a) < demand(b); });
    unsigned __int128 left = budget;
    size_t k = 0;
    for (; k < n; ++k) {                     // segments whose whole demand is within the level
        const size_t i = order[k];
        const unsigned __int128 d = demand(i);
        if (d * (n - k) > left) break;       // d > left / (n - k): the level binds from here on
        steps[i] = cap[i];
        left -= d;
    }
    const uint64_t level = k < n ? (uint64_t)(left / (n - k)) : 0;
    for (; k < n; ++k) {
        const size_t i = order[k];
        uint64_t s = cost[i] ? level / cost[i] : cap[i];
        s = std::max<uint64_t>(s, std::min<uint64_t>(min_steps, cap[i]));
        steps[i] = std::min<uint64_t>(s, cap[i]);
    }
}

}  // namespace pfc

#endif
