// Order in which the request server hands out photon ids to the A-scans in flight.
//
// A slice's kernel claims photon ids in the order of its segment table and, in ~20 ms,
// claims only a fraction of the ids the table lists, so the table's order is the order of
// service. Host-only and free of HIP, so the CPU tests compile it on its own.
#pragma once

#include <algorithm>
#include <cstdint>
#include <map>
#include <utility>
#include <vector>

namespace octmps {

// One A-scan with photons still to hand out.
struct OpenScan {
    uint64_t req_id;     // request id (submission order)
    double t_submit_ms;  // request submission time
    uint64_t from, rem;  // first unclaimed photon, photons left
    uint32_t slot;       // device scan slot
};

// One segment: ids [start, next start) map to photons poff.. of `slot`; `open` indexes the
// OpenScan it came from.
struct Segment {
    uint64_t start, poff;
    uint32_t slot;
    size_t open;
};

// Waiting time over which a request's remaining photons count half as much.
constexpr double kAgingMs = 10000.0;

// Least remaining photons first, aged: requests go by their unclaimed photons divided by
// (1 + wait / kAgingMs), ties by submission order; each request's A-scans follow in the
// order given. For a burst this is shortest-remaining-first (least mean latency); a large
// request that has waited overtakes newer small ones, so none starves. Lists whole
// remaining ranges until `budget` ids are listed; returns the ids listed.
inline uint64_t order_least_remaining(const std::vector<OpenScan> &open, double now_ms, uint64_t budget,
                                      std::vector<Segment> &out) {
    std::map<uint64_t, uint64_t> left;
    for (const OpenScan &o : open) left[o.req_id] += o.rem;
    std::vector<std::pair<double, size_t>> key(open.size());
    for (size_t k = 0; k < open.size(); ++k) {
        const double wait = std::max(0.0, now_ms - open[k].t_submit_ms);
        key[k] = {(double)left[open[k].req_id] / (1.0 + wait / kAgingMs), k};
    }
    std::stable_sort(key.begin(), key.end(), [&](const std::pair<double, size_t> &a, const std::pair<double, size_t> &b) {
        if (a.first != b.first) return a.first < b.first;
        return open[a.second].req_id < open[b.second].req_id;
    });
    uint64_t total = 0;
    for (size_t j = 0; j < key.size() && total < budget; ++j) {
        const OpenScan &o = open[key[j].second];
        const uint64_t len = std::min<uint64_t>(o.rem, budget - total);
        out.push_back({total, o.from, o.slot, key[j].second});
        total += len;
    }
    return total;
}

// Equal shares: round-robin chunks of max(256, round / open A-scans) ids, so that one round
// covers about `round` ids (the caller passes what the last slice claimed) and every open
// A-scan is reached in each slice. Lists until `budget` ids; returns the ids listed.
inline uint64_t order_equal_shares(const std::vector<OpenScan> &open, uint64_t round, uint64_t budget,
                                   std::vector<Segment> &out) {
    uint64_t total = 0;
    if (open.empty()) return 0;
    const uint64_t chunk = std::max<uint64_t>(256, round / open.size());
    std::vector<uint64_t> cur(open.size());
    for (size_t k = 0; k < open.size(); ++k) cur[k] = open[k].from;
    bool any = true;
    while (any && total < budget) {
        any = false;
        for (size_t k = 0; k < open.size() && total < budget; ++k) {
            const uint64_t lim = open[k].from + open[k].rem;
            if (cur[k] >= lim) continue;
            const uint64_t len = std::min<uint64_t>(std::min<uint64_t>(chunk, lim - cur[k]), budget - total);
            out.push_back({total, cur[k], open[k].slot, k});
            cur[k] += len;
            total += len;
            any = true;
        }
    }
    return total;
}

}  // namespace octmps
