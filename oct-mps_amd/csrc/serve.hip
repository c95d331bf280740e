// Serving: continuous batching of A-scan / B-scan requests on one GPU.
//
// The reference runs one A-scan at a time from a host loop (src/main.cu:948-1051). Here a
// scheduler thread runs the transport kernel in slices: before each slice it admits queued
// requests, first come first served, while their HBM footprint fits the server's budget -
// the request's tallies, scan slots and parameter set, plus its mesh unless the mesh already
// counts for an admitted request (photon streams are derived per photon, so no RNG state is
// stored per request; the per-lane state of live photons is a fixed cost of the server). A
// mesh that is not resident is uploaded on admission; meshes no admitted request uses are
// kept resident while the budget allows and evicted, least recently used first, when an
// admission needs the room (registered meshes only: pinned ones stay). It then lists the
// photon ids of all admitted requests in the order of its hand-out
// policy (least remaining photons first with aging, or equal shares; serve_order.hpp);
// during the slice, lanes whose photon dies pick up the next listed id; at the end
// of a slice live photons are saved per lane and resumed by the next slice
// (SaveThreadState/RestoreThreadState, src/octmps_kernel.cuh:1154-1327). A request whose
// photons are all done is handed to a completion thread, which copies its tallies to the
// caller on its own stream while the next slice runs and then ends the request. Every
// photon's streams depend only on (seed, salt, photon index), so a request's tallies are
// bitwise the same alone or in any mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "device_types.hpp"
#include "host_common.hpp"
#include "runtime_internal.hpp"
#include "serve_order.hpp"

namespace octmps {
namespace {

constexpr int kMaxSets = kMaxServeSets;
// Hand-out policies (octmps_serve_opts.policy), serve_order.hpp.
constexpr int kLeastRemainingFirst = 0, kEqualShares = 1;

struct ScanState {
    int slot = -1;
    uint64_t n = 0;       // photons to run
    uint64_t claimed = 0; // photons handed out
    uint64_t limit = 0;   // n, or claimed once cancelled
};

struct Req {
    uint64_t id = 0;
    int index = 0;
    int32_t mesh_id = 0;
    MeshEntry *me = nullptr;  // the context's mesh entry (entries live as long as the context)
    uint64_t charge = 0;      // bytes of the budget this request holds while admitted
    bool counted = false;     // counted in me->refs (from submit to done)
    bool mesh_active = false; // counted in me->active (from staging to done)
    octmps_params p;
    std::vector<octmps_ascan> scans;
    uint64_t seed = 0;
    uint64_t *out = nullptr;
    uint64_t words = 0;
    int status = OCTMPS_OK;
    bool done = false, admitted = false, cancelled = false, fail_once = false;
    double t_submit = 0, t_admit = 0, t_done = 0;
    uint64_t photons = 0;
    int set = -1;
    int64_t slot_base = -1;  // first of the request's contiguous scan slots (while admitted)
    uint64_t tally_base = 0;
    std::vector<ScanState> ss;
};

struct Range {
    uint64_t off, len;
};

// A mesh counted in a server's budget: resident on its behalf, used by `users` admitted
// requests (an idle one may be evicted when an admission needs the room).
struct MeshUse {
    uint64_t bytes = 0;
    int users = 0;
    double last_use = 0;
};

// Budget bytes of one request's device state besides its tallies: scan slots (the slot
// record and its completion counter) and a parameter set with its region table.
uint64_t request_state_bytes(size_t n_scans) {
    return (uint64_t)n_scans * (sizeof(DevScan) + 8) + sizeof(DevSet) + kMaxRegions * sizeof(DevRegion);
}

}  // namespace
}  // namespace octmps

using namespace octmps;

struct octmps_server {
    octmps_ctx *ctx = nullptr;
    octmps_serve_opts o;
    std::thread th;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    bool stop = false;
    std::deque<Req *> pending;
    std::vector<Req *> active;
    std::map<uint64_t, Req *> reqs;
    uint64_t next_id = 1;
    int submits = 0;
    double t0 = 0;
    // device state (owned by the scheduler thread)
    hipStream_t stream = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    DevBuf regions, sets, slots, seg_start, seg_slot, seg_poff, tallies, ctrl, slot_done, save, cont, set_map;
    // Host copy of set_map: the LDS slot of each set id, then the set id of each LDS slot.
    uint8_t set_slot_h[kMaxServeSets] = {};
    int32_t lds_ids_h[kLdsSets] = {};
    int grid = 0, block = 0, q = 0;
    uint64_t lanes = 0;
    int n_slots = 0;
    std::vector<Range> free_slots;  // contiguous free scan-slot ranges
    std::vector<int> free_sets;
    std::vector<Range> free_tally;
    uint64_t tally_cap_words = 0;
    bool have_saved = false;
    uint64_t last_claims = 0;  // photon ids claimed by the last slice
    // HBM footprint admission
    uint64_t budget = 0, used = 0;  // request footprints admitted (bytes) against the budget
    uint64_t fixed_bytes = 0;       // the server's own device state, outside the budget
    std::map<int32_t, MeshUse> mesh_use;
    std::vector<int32_t> to_evict;  // meshes dropped from the budget, evicted before the next slice
    // Completion: a finished request's tallies are copied to the caller on their own thread and
    // stream while the next slice runs, and the request ends (its resources freed, waiters
    // woken) once its copy is done. tally_mu keeps the tally pool from being replaced under a
    // copy (lock order: mu, then tally_mu; the copy thread never holds tally_mu while taking mu).
    std::thread copy_th;
    std::mutex copy_mu, tally_mu;
    std::condition_variable cv_copy;
    std::deque<Req *> copy_q;
    bool copy_stop = false;
    int copies_out = 0;  // requests handed to the copy thread and not yet ended (under mu)
    hipStream_t copy_stream = nullptr;
    // stats
    std::atomic<uint64_t> st_uploads{0}, st_evictions{0};
    uint64_t st_photons = 0, st_steps = 0, st_slices = 0;
    double st_kernel_ms = 0;
    int fatal = OCTMPS_OK;
};

namespace octmps {
namespace {

// First-fit allocation from a list of free ranges sorted by offset.
bool range_alloc(std::vector<Range> &free, uint64_t n, uint64_t &off) {
    for (size_t i = 0; i < free.size(); ++i) {
        Range &r = free[i];
        if (r.len >= n) {
            off = r.off;
            r.off += n;
            r.len -= n;
            if (r.len == 0) free.erase(free.begin() + i);
            return true;
        }
    }
    return false;
}

// Returns a range to the list, merging it with its neighbours.
void range_free(std::vector<Range> &free, uint64_t off, uint64_t n) {
    auto it = std::lower_bound(free.begin(), free.end(), off, [](const Range &r, uint64_t o) { return r.off < o; });
    it = free.insert(it, Range{off, n});
    if (it + 1 != free.end() && it->off + it->len == (it + 1)->off) {
        it->len += (it + 1)->len;
        free.erase(it + 1);
    }
    if (it != free.begin() && (it - 1)->off + (it - 1)->len == it->off) {
        (it - 1)->len += it->len;
        free.erase(it);
    }
}

// Grows the device tally pool to hold at least `words` more (a new allocation; the live
// tallies are copied over). Called between slices, from the scheduler thread.
bool grow_tallies(octmps_server *s, uint64_t words) {
    std::lock_guard<std::mutex> tk(s->tally_mu);  // no completion copy reads the old pool
    const uint64_t old = s->tally_cap_words;
    const uint64_t ncap = std::max<uint64_t>(old * 2, old + words);
    void *np = nullptr;
    if (hipMalloc(&np, ncap * 8) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    if (hipMemcpyAsync(np, s->tallies.p, old * 8, hipMemcpyDeviceToDevice, s->stream) != hipSuccess ||
        hipStreamSynchronize(s->stream) != hipSuccess) {
        (void)hipFree(np);
        return false;
    }
    (void)hipFree(s->tallies.p);
    s->tallies.p = np;
    s->tallies.cap = ncap * 8;
    s->tally_cap_words = ncap;
    range_free(s->free_tally, old, ncap - old);
    return true;
}

bool alloc_tally(octmps_server *s, uint64_t words, uint64_t &off) {
    return range_alloc(s->free_tally, words, off) ||
           (grow_tallies(s, words) && range_alloc(s->free_tally, words, off));
}

void release(octmps_server *s, Req *r) {
    if (r->slot_base >= 0) range_free(s->free_slots, (uint64_t)r->slot_base, r->scans.size());
    r->slot_base = -1;
    if (r->set >= 0) {
        // Smallest id first: ids stay dense (the slice maps the sets in flight to LDS
        // slots whatever their ids).
        s->free_sets.push_back(r->set);
        std::sort(s->free_sets.begin(), s->free_sets.end(), std::greater<int>());
    }
    if (r->admitted) {
        range_free(s->free_tally, r->tally_base, r->words * r->scans.size());
        MeshUse &u = s->mesh_use[r->mesh_id];
        u.users--;
        u.last_use = now_ms();
        s->used -= r->charge;
        r->charge = 0;
    }
    r->set = -1;
    r->admitted = false;
    for (ScanState &x : r->ss) x.slot = -1;
}

// Uploads a request's parameter set and scan slots and zeroes its tallies and per-slot done
// counters. The hand-out counters (claimed, limit) belong to the caller, under the lock: a
// cancel may already have lowered `limit`.
octmps_status stage(octmps_server *s, Req *r) {
    octmps_ctx *ctx = s->ctx;
    const MeshEntry &me = *r->me;
    DevSet S;
    std::vector<DevRegion> R(kMaxRegions);
    octmps_status st = make_set(&r->p, me.max_region, r->set * kMaxRegions, S, R.data());
    if (st != OCTMPS_OK) return st;
    hipError_t e;
    if ((e = hipMemcpyAsync((DevRegion *)s->regions.p + (size_t)r->set * kMaxRegions, R.data(),
                            R.size() * sizeof(DevRegion), hipMemcpyHostToDevice, s->stream)) ||
        (e = hipMemcpyAsync((DevSet *)s->sets.p + r->set, &S, sizeof S, hipMemcpyHostToDevice, s->stream)))
        return hip_fail(e, "serve stage");
    // The request's scan slots are one contiguous range: one copy of their records and one
    // fill of their completion counters.
    const size_t n = r->scans.size();
    std::vector<DevScan> d(n);
    for (size_t i = 0; i < n; ++i) {
        st = make_scan(ctx, me, &r->p, r->scans[i], r->seed, r->set, r->tally_base + i * r->words, d[i]);
        if (st != OCTMPS_OK) return st;
    }
    if ((e = hipMemcpyAsync((DevScan *)s->slots.p + r->slot_base, d.data(), n * sizeof(DevScan),
                            hipMemcpyHostToDevice, s->stream)) ||
        (e = launch_fill_u64((uint64_t *)s->slot_done.p + r->slot_base, n, 0, s->stream)))
        return hip_fail(e, "serve stage");
    if ((e = launch_fill_u64((uint64_t *)s->tallies.p + r->tally_base, r->words * r->scans.size(), 0, s->stream)))
        return hip_fail(e, "serve stage");
    return OCTMPS_OK;
}

// Ends a request that holds no server resources (cancelled or failed before admission, or
// already released): drops its count on the mesh and wakes waiters. Under the lock.
void end_unadmitted(octmps_server *s, Req *r, int status) {
    r->status = status;
    r->done = true;
    r->t_done = now_ms() - s->t0;
    if (r->counted) r->me->refs--;
    r->counted = false;
    s->cv_done.notify_all();
}

// Admission under the lock, first come first served: the head request is admitted when scan
// slots and a parameter set are free and its footprint fits the budget, after dropping idle
// meshes (no admitted user) from the budget, least recently used first, when it needs the
// room; they are evicted from the device before the next slice. A request whose footprint
// cannot fit even alone was refused at submit.
void admit(octmps_server *s, std::vector<Req *> &staged) {
    while (!s->pending.empty()) {
        Req *r = s->pending.front();
        if (r->cancelled) {
            s->pending.pop_front();
            end_unadmitted(s, r, OCTMPS_E_CANCELLED);
            continue;
        }
        const size_t need_slots = r->scans.size();
        if (s->free_sets.empty()) break;
        const uint64_t charge = r->words * need_slots * 8 + request_state_bytes(need_slots);
        const bool counted_mesh = s->mesh_use.count(r->mesh_id) != 0;
        const uint64_t mesh_bytes = counted_mesh ? 0 : r->me->device_bytes();
        while (s->used + charge + mesh_bytes > s->budget) {
            auto lru = s->mesh_use.end();
            for (auto it = s->mesh_use.begin(); it != s->mesh_use.end(); ++it)
                if (it->second.users == 0 && it->first != r->mesh_id &&
                    (lru == s->mesh_use.end() || it->second.last_use < lru->second.last_use))
                    lru = it;
            if (lru == s->mesh_use.end()) break;
            s->used -= lru->second.bytes;
            s->to_evict.push_back(lru->first);
            s->mesh_use.erase(lru);
        }
        if (s->used + charge + mesh_bytes > s->budget) break;  // waits for admitted requests to end
        uint64_t slot_off;
        if (!range_alloc(s->free_slots, need_slots, slot_off)) break;  // waits for slots to free up
        uint64_t off;
        if (!alloc_tally(s, r->words * need_slots, off)) {
            range_free(s->free_slots, slot_off, need_slots);
            if (s->active.empty() && s->copies_out == 0) {  // the device cannot hold its tallies at all
                s->pending.pop_front();
                end_unadmitted(s, r, OCTMPS_E_NOMEM);
                continue;
            }
            break;
        }
        MeshUse &u = s->mesh_use[r->mesh_id];
        if (!counted_mesh) u.bytes = mesh_bytes;
        u.users++;
        u.last_use = now_ms();
        s->used += charge + mesh_bytes;
        r->charge = charge;
        r->tally_base = off;
        r->admitted = true;
        r->set = s->free_sets.back();
        s->free_sets.pop_back();
        r->slot_base = (int64_t)slot_off;
        for (size_t i = 0; i < need_slots; ++i) r->ss[i].slot = (int)(slot_off + i);
        r->t_admit = now_ms() - s->t0;
        s->pending.pop_front();
        s->active.push_back(r);
        staged.push_back(r);
    }
}

void finish(octmps_server *s, Req *r, int status) {
    r->status = status;
    r->done = true;
    r->t_done = now_ms() - s->t0;
    release(s, r);
    if (r->mesh_active) r->me->active--;
    r->mesh_active = false;
    if (r->counted) r->me->refs--;
    r->counted = false;
    auto it = std::find(s->active.begin(), s->active.end(), r);
    if (it != s->active.end()) s->active.erase(it);  // a request being copied has left it
    s->cv_done.notify_all();
}

// Copies finished requests' tallies to their callers and ends them, in hand-off order, while
// the scheduler runs the next slices. Leaves once the server closes and the queue is empty.
void copier(octmps_server *s) {
    (void)hipSetDevice(s->ctx->device);
    for (;;) {
        Req *r = nullptr;
        {
            std::unique_lock<std::mutex> ck(s->copy_mu);
            s->cv_copy.wait(ck, [&] { return s->copy_stop || !s->copy_q.empty(); });
            if (s->copy_q.empty()) return;
            r = s->copy_q.front();
            s->copy_q.pop_front();
        }
        hipError_t e;
        {
            std::lock_guard<std::mutex> tk(s->tally_mu);
            e = hipMemcpyAsync(r->out, (uint64_t *)s->tallies.p + r->tally_base, r->words * r->scans.size() * 8,
                               hipMemcpyDeviceToHost, s->copy_stream);
            if (e == hipSuccess) e = hipStreamSynchronize(s->copy_stream);
        }
        std::lock_guard<std::mutex> lk(s->mu);
        finish(s, r, e == hipSuccess ? OCTMPS_OK : OCTMPS_E_DEVICE);
        s->copies_out--;
        s->cv_work.notify_all();  // its slots, set and tallies are free for admission
    }
}

// Before a slice, with the pool held exclusively: evicts the meshes admission dropped from
// the budget (unless another user holds them or they are pinned) and makes the meshes of
// newly admitted requests resident. Returns the requests whose mesh could not be placed.
std::vector<std::pair<Req *, octmps_status>> place_meshes(octmps_server *s, const std::vector<Req *> &staged,
                                                          std::vector<int32_t> &evict) {
    octmps_ctx *ctx = s->ctx;
    std::vector<std::pair<Req *, octmps_status>> failed;
    std::unique_lock<std::shared_mutex> pk(ctx->pool_mu);
    for (int32_t id : evict) {
        MeshEntry *me = mesh_at(ctx, id);
        if (me && me->resident && !me->pinned && me->active.load() == 0) {
            mesh_evict(ctx, *me);
            s->st_evictions++;
        }
    }
    for (Req *r : staged) {
        const bool was = r->me->resident;
        const octmps_status st = mesh_make_resident(ctx, *r->me);
        if (st != OCTMPS_OK) {
            failed.push_back({r, st});
            continue;
        }
        if (!was) s->st_uploads++;
        r->me->active++;
        r->me->last_use = now_ms();
        r->mesh_active = true;
    }
    return failed;
}

void scheduler(octmps_server *s) {
    octmps_ctx *ctx = s->ctx;
    (void)hipSetDevice(ctx->device);
    std::vector<uint64_t> done_h((size_t)s->n_slots);
    std::vector<int32_t> evict;
    for (;;) {
        std::vector<Req *> staged;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            while (!s->stop && s->pending.empty() && s->active.empty()) s->cv_work.wait(lk);
            if (s->stop && s->pending.empty() && s->active.empty()) break;
            admit(s, staged);
            if (staged.empty() && s->active.empty() && s->copies_out > 0) {
                // Nothing to run: the queued requests wait for resources that requests still
                // being copied hold. Their copies wake the scheduler (meshes to evict stay
                // listed for the next pass).
                s->cv_work.wait(lk);
                continue;
            }
            evict.swap(s->to_evict);
        }
        if (!staged.empty() || !evict.empty()) {
            auto failed = place_meshes(s, staged, evict);
            evict.clear();
            if (!failed.empty()) {
                std::lock_guard<std::mutex> lk(s->mu);
                for (auto &f : failed) {
                    finish(s, f.first, f.second);
                    staged.erase(std::find(staged.begin(), staged.end(), f.first));
                }
            }
        }
        std::shared_lock<std::shared_mutex> pk(ctx->pool_mu);
        for (Req *r : staged) {
            octmps_status st = stage(s, r);
            if (st != OCTMPS_OK) {
                std::lock_guard<std::mutex> lk(s->mu);
                finish(s, r, st);
            }
        }
        // Segments: the photon-id ranges of the A-scans that still have photons to hand
        // out, in the order the policy gives them.
        std::vector<uint64_t> seg_start, seg_poff;
        std::vector<uint32_t> seg_slot;
        std::vector<std::pair<Req *, int>> seg_owner;
        std::vector<std::pair<Req *, int>> open;
        // Every live photon belongs to an active request (a request ends only once all
        // photons it handed out are done), so the sets of the active requests are the
        // only ones the slice can touch: when there are at most kLdsSets of them and their
        // meshes' regions fit, the slice keeps them in LDS through the set_map.
        // The open A-scans' first unclaimed photon and photons left are read under the lock
        // (a cancel lowers `limit` from another thread).
        bool sets_in_lds = true;
        int n_lds = 0, region_slots = 1;  // region slots per set: the largest region index + 1
        std::vector<OpenScan> os;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            for (Req *r : s->active) {
                if (r->me->max_region >= kLdsSetRegions || n_lds == kLdsSets) sets_in_lds = false;
                region_slots = std::max(region_slots, r->me->max_region + 1);
                if (sets_in_lds) {
                    s->set_slot_h[r->set] = (uint8_t)n_lds;
                    s->lds_ids_h[n_lds++] = r->set;
                }
                for (size_t i = 0; i < r->ss.size(); ++i)
                    if (r->ss[i].claimed < r->ss[i].limit) {
                        open.push_back({r, (int)i});
                        os.push_back({r->id, r->t_submit, r->ss[i].claimed, r->ss[i].limit - r->ss[i].claimed,
                                      (uint32_t)r->ss[i].slot});
                    }
            }
        }
        const uint64_t budget = s->o.slice_photons ? s->o.slice_photons : s->lanes * 4;
        std::vector<Segment> segs;
        const uint64_t total =
            s->o.policy == kLeastRemainingFirst
                ? order_least_remaining(os, now_ms() - s->t0, budget, segs)
                : order_equal_shares(os, std::max<uint64_t>(s->last_claims, s->lanes / 8), budget, segs);
        for (const Segment &g : segs) {
            seg_start.push_back(g.start);
            seg_poff.push_back(g.poff);
            seg_slot.push_back(g.slot);
            seg_owner.push_back(open[g.open]);
        }
        seg_start.push_back(total);
        if (seg_slot.empty()) {  // keep the table non-empty for the binary search
            seg_slot.push_back(0);
            seg_poff.push_back(0);
            seg_start.insert(seg_start.begin(), 0);
        }
        // Kernel variant of the slice. The cooperative record fetch runs whenever the blocks
        // fit its stage (at most 4 waves) and the options allow it. Beside its stage the LDS
        // set table has kCoopRegionSlots slots: 16 sets of up to 8 regions or 8 sets of up to
        // 16. Sets that fit the per-lane variant's larger table (16 sets of 16) but not this
        // one run the per-lane variant with the sets in LDS (region reads from global memory
        // cost more there than the cooperative fetch saves); sets that fit no LDS table run
        // the cooperative variant reading them from global memory.
        bool coop = s->block <= 256 && s->o.record_fetch != 1;
        uint32_t set_shift = 4;
        if (coop && sets_in_lds) {
            if (region_slots <= 8 && n_lds <= 16) set_shift = 3;
            else if (region_slots <= 16 && n_lds <= 8) set_shift = 4;
            else coop = false;
        }
        for (int j = n_lds; j < kLdsSets; ++j) s->lds_ids_h[j] = s->lds_ids_h[0];  // unused slots: any valid id
        hipError_t e;
        bool ok = (e = s->seg_start.ensure(seg_start.size() * 8)) == hipSuccess &&
                  (e = s->seg_slot.ensure(seg_slot.size() * 4)) == hipSuccess &&
                  (e = s->seg_poff.ensure(seg_poff.size() * 8)) == hipSuccess &&
                  (e = hipMemcpyAsync(s->seg_start.p, seg_start.data(), seg_start.size() * 8, hipMemcpyHostToDevice,
                                      s->stream)) == hipSuccess &&
                  (e = hipMemcpyAsync(s->seg_slot.p, seg_slot.data(), seg_slot.size() * 4, hipMemcpyHostToDevice,
                                      s->stream)) == hipSuccess &&
                  (e = hipMemcpyAsync(s->seg_poff.p, seg_poff.data(), seg_poff.size() * 8, hipMemcpyHostToDevice,
                                      s->stream)) == hipSuccess &&
                  (e = launch_fill_u64((uint64_t *)s->ctrl.p, CTRL_WORDS, 0, s->stream)) == hipSuccess;
        if (ok && sets_in_lds)
            ok = (e = hipMemcpyAsync(s->set_map.p, s->set_slot_h, sizeof s->set_slot_h, hipMemcpyHostToDevice,
                                     s->stream)) == hipSuccess &&
                 (e = hipMemcpyAsync((uint8_t *)s->set_map.p + sizeof s->set_slot_h, s->lds_ids_h,
                                     sizeof s->lds_ids_h, hipMemcpyHostToDevice, s->stream)) == hipSuccess;
        KArgs A;
        std::memset(&A, 0, sizeof A);
        A.tets = ctx->pool_tets;
        A.tets_bytes = (uint32_t)(ctx->pool_cap * (int64_t)sizeof(TetRec));
        A.regions = (const DevRegion *)s->regions.p;
        A.sets = (const DevSet *)s->sets.p;
        A.slots = (const DevScan *)s->slots.p;
        A.seg_start = (const uint64_t *)s->seg_start.p;
        A.seg_slot = (const uint32_t *)s->seg_slot.p;
        A.seg_poff = (const uint64_t *)s->seg_poff.p;
        A.mult = ctx->d_mult;
        A.tallies = (uint64_t *)s->tallies.p;
        A.ctrl = (uint64_t *)s->ctrl.p;
        A.slot_done = (uint64_t *)s->slot_done.p;
        A.save = (SavedLane *)s->save.p;
        A.cont = (ContRec *)s->cont.p;
        A.id_end = total;
        // A slice lasts ~20 ms of device wall clock (iteration cap as a backstop).
        A.iter_budget = 1u << 16;
        A.slice_ticks = ctx->wall_khz > 0 ? (uint64_t)ctx->wall_khz * 20 : 0;
        if (!A.slice_ticks) A.iter_budget = 8192;
        A.max_steps = 1ull << 22;
        A.n_seg = (int32_t)seg_slot.size();
        A.n_triplets = ctx->table->n_triplets;
        A.refill_q = (uint32_t)s->q;
        A.scatter_q = 32;
        A.slice_mode = 1;
        A.restore = s->have_saved ? 1u : 0u;
        // A slice whose table lists every unclaimed id runs its full length (waves leave once
        // their photons are done and the ids are gone); ending it as soon as the ids are
        // handed out would save and restore every live photon each time a request arrives,
        // which under a steady stream of small requests left most slices a few iterations
        // long. Only a table the budget cut short ends on drain, so that the next slice
        // lists the rest.
        uint64_t unclaimed = 0;
        for (const OpenScan &o : os) unclaimed += o.rem;
        A.exit_on_drain = (total > 0 && total < unclaimed) ? 1u : 0u;
        A.sets_in_lds = sets_in_lds ? 1u : 0u;
        A.coop_fetch = coop ? 1u : 0u;
        A.set_region_shift = set_shift;
        A.set_slot = (const uint8_t *)s->set_map.p;
        A.lds_set_ids = (const int32_t *)((const uint8_t *)s->set_map.p + sizeof s->set_slot_h);
        uint64_t ctrl[CTRL_WORDS] = {0};
        if (ok) ok = (e = hipEventRecord(s->e0, s->stream)) == hipSuccess &&
                     (e = launch_transport(A, s->grid, s->block, true, false, s->stream)) == hipSuccess &&
                     (e = hipEventRecord(s->e1, s->stream)) == hipSuccess &&
                     (e = hipMemcpyAsync(ctrl, s->ctrl.p, sizeof ctrl, hipMemcpyDeviceToHost, s->stream)) == hipSuccess &&
                     (e = hipMemcpyAsync(done_h.data(), s->slot_done.p, done_h.size() * 8, hipMemcpyDeviceToHost,
                                         s->stream)) == hipSuccess &&
                     (e = hipStreamSynchronize(s->stream)) == hipSuccess;
        if (!ok) {
            hip_fail(e, "serve slice");
            std::lock_guard<std::mutex> lk(s->mu);
            s->fatal = OCTMPS_E_DEVICE;
            while (!s->active.empty()) finish(s, s->active.front(), OCTMPS_E_DEVICE);
            while (!s->pending.empty()) {
                Req *r = s->pending.front();
                s->pending.pop_front();
                end_unadmitted(s, r, OCTMPS_E_DEVICE);
            }
            break;
        }
        s->have_saved = true;
        float kms = 0;
        (void)hipEventElapsedTime(&kms, s->e0, s->e1);
        // Counters, claims and completions change under the lock: stats and cancel read them.
        std::lock_guard<std::mutex> lk(s->mu);
        s->st_kernel_ms += kms;
        s->st_steps += ctrl[CTRL_STEPS];
        s->st_slices++;
        // Ids are claimed in increasing order: every segment below `claimed` was handed out.
        uint64_t claimed = std::min<uint64_t>(ctrl[CTRL_NEXT_ID], total);
        s->last_claims = claimed;
        for (size_t k = 0; k < seg_owner.size(); ++k) {
            uint64_t a = seg_start[k], b = seg_start[k + 1];
            if (claimed <= a) break;
            uint64_t used = std::min(claimed, b) - a;
            seg_owner[k].first->ss[seg_owner[k].second].claimed += used;
        }
        std::vector<Req *> act = s->active;
        for (Req *r : act) {
            bool complete = true;
            uint64_t ph = 0;
            for (ScanState &x : r->ss) {
                uint64_t d = done_h[(size_t)x.slot];
                ph += d;
                if (!(x.claimed >= x.limit && d >= x.claimed)) complete = false;
            }
            if (!complete) continue;
            if (r->cancelled) {
                finish(s, r, OCTMPS_E_CANCELLED);
                continue;
            }
            if (r->fail_once) {
                // Injected device fault on this work item: discard and rerun it from its
                // deterministic streams in the same slots.
                r->fail_once = false;
                for (ScanState &x : r->ss) {
                    x.claimed = 0;
                    x.limit = x.n;
                }
                octmps_status st = stage(s, r);
                if (st != OCTMPS_OK) finish(s, r, st);
                continue;
            }
            // Complete: its tallies are final (none of its photons is left in a lane). The copy
            // to the caller runs beside the next slice; the request keeps its slots, set and
            // tally range until the copy is done.
            r->photons = ph;
            s->st_photons += ph;
            s->active.erase(std::find(s->active.begin(), s->active.end(), r));
            s->copies_out++;
            {
                std::lock_guard<std::mutex> ck(s->copy_mu);
                s->copy_q.push_back(r);
            }
            s->cv_copy.notify_one();
        }
    }
}

}  // namespace
}  // namespace octmps

extern "C" octmps_status octmps_serve_open(octmps_ctx *ctx, const octmps_serve_opts *opts, octmps_server **out) {
    if (!ctx || !out) return fail(OCTMPS_E_INVALID, "null argument");
    octmps_server *s = new (std::nothrow) octmps_server();
    if (!s) return fail(OCTMPS_E_NOMEM, "host allocation failed");
    std::memset(&s->o, 0, sizeof s->o);
    s->o.fail_request_index = -1;
    if (opts) s->o = *opts;
    s->ctx = ctx;
    s->t0 = now_ms();
    s->n_slots = s->o.max_inflight_scans > 0 ? s->o.max_inflight_scans : 32768;
    s->block = s->o.block_size > 0 ? s->o.block_size : 256;
    s->q = s->o.refill_quantum > 0 ? s->o.refill_quantum : 8;
    if (s->block % 64 != 0 || s->block > 256) {
        delete s;
        return fail(OCTMPS_E_INVALID, "block size must be 64, 128, 192 or 256");
    }
    if (s->o.policy != kLeastRemainingFirst && s->o.policy != kEqualShares) {
        delete s;
        return fail(OCTMPS_E_INVALID, "policy must be 0 (least remaining first) or 1 (equal shares)");
    }
    if (s->o.record_fetch < 0 || s->o.record_fetch > 2) {
        delete s;
        return fail(OCTMPS_E_INVALID, "record_fetch must be 0 (auto), 1 (per-lane) or 2 (cooperative where it fits)");
    }
    if (hipError_t se = hipSetDevice(ctx->device); se != hipSuccess) {
        delete s;
        return hip_fail(se, "set device");
    }
    grid_for(ctx, s->block, 0, true, s->grid);
    s->lanes = (uint64_t)s->grid * s->block;
    s->last_claims = s->lanes;
    // Default budget: three quarters of the device's memory. The budget covers the requests'
    // footprints (meshes, tallies, per-request state); the server's fixed state (per-lane
    // live photons and continuation photons, slot and set tables) comes on top of it.
    s->budget = s->o.hbm_budget_bytes ? s->o.hbm_budget_bytes : ctx->hbm_total / 4 * 3;
    s->fixed_bytes = s->lanes * (sizeof(SavedLane) + sizeof(ContRec)) +
                     (uint64_t)s->n_slots * (sizeof(DevScan) + 8) +
                     (uint64_t)kMaxSets * (sizeof(DevSet) + kMaxRegions * sizeof(DevRegion));
    s->used = 0;
    // The tally pool starts small and grows with the admitted requests.
    s->tally_cap_words = (8ull << 20) / 8;
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipStreamCreateWithFlags(&s->copy_stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreate(&s->e0)) != hipSuccess || (e = hipEventCreate(&s->e1)) != hipSuccess ||
        (e = s->regions.ensure((size_t)kMaxSets * kMaxRegions * sizeof(DevRegion))) != hipSuccess ||
        (e = s->sets.ensure((size_t)kMaxSets * sizeof(DevSet))) != hipSuccess ||
        (e = s->set_map.ensure(sizeof s->set_slot_h + sizeof s->lds_ids_h)) != hipSuccess ||
        (e = s->slots.ensure((size_t)s->n_slots * sizeof(DevScan))) != hipSuccess ||
        (e = s->slot_done.ensure((size_t)s->n_slots * 8)) != hipSuccess ||
        (e = s->ctrl.ensure(CTRL_WORDS * 8)) != hipSuccess ||
        (e = s->tallies.ensure(s->tally_cap_words * 8)) != hipSuccess ||
        (e = s->save.ensure(s->lanes * sizeof(SavedLane))) != hipSuccess ||
        (e = s->cont.ensure(s->lanes * sizeof(ContRec))) != hipSuccess ||
        (e = launch_clear_saved((SavedLane *)s->save.p, s->lanes, s->stream)) != hipSuccess ||
        (e = launch_fill_u64((uint64_t *)s->slot_done.p, (uint64_t)s->n_slots, 0, s->stream)) != hipSuccess ||
        (e = hipStreamSynchronize(s->stream)) != hipSuccess) {
        octmps_serve_close(s);
        return fail(OCTMPS_E_NOMEM, std::string("server allocation: ") + hipGetErrorString(e));
    }
    s->free_slots.push_back({0, (uint64_t)s->n_slots});
    for (int i = kMaxSets - 1; i >= 0; --i) s->free_sets.push_back(i);
    s->free_tally.push_back({0, s->tally_cap_words});
    s->copy_th = std::thread(copier, s);
    s->th = std::thread(scheduler, s);
    *out = s;
    return OCTMPS_OK;
}

extern "C" octmps_status octmps_serve_submit(octmps_server *s, const octmps_request *rq, uint64_t *req_id) {
    if (!s || !rq || !rq->params || !rq->scans || rq->n_scans <= 0 || !rq->tallies_out)
        return fail(OCTMPS_E_INVALID, "bad request");
    octmps_ctx *ctx = s->ctx;
    // The request counts as a user of its mesh from here until it is done, so the mesh cannot
    // be released under it (octmps_ctx_release_mesh answers OCTMPS_E_BUSY meanwhile).
    MeshEntry *me = nullptr;
    const uint64_t words = tally_words(rq->params->num_depth_bins);
    {
        std::shared_lock<std::shared_mutex> pk(ctx->pool_mu);
        me = mesh_at(ctx, rq->mesh_id);
        if (!me) return fail(OCTMPS_E_NOT_FOUND, "unknown mesh id");
        DevSet S;
        std::vector<DevRegion> R(kMaxRegions);
        octmps_status st = make_set(rq->params, me->max_region, 0, S, R.data());
        if (st != OCTMPS_OK) return st;
        for (int i = 0; i < rq->n_scans; ++i)
            if (rq->scans[i].root_tet < 0 || rq->scans[i].root_tet >= me->n_tets)
                return fail(OCTMPS_E_PROBE, "root tet out of range");
        if (!seed_valid(rq->seed, ctx->table->mult[0])) return fail(OCTMPS_E_SEED, "invalid seed");
        const uint64_t footprint = words * (uint64_t)rq->n_scans * 8 + request_state_bytes((size_t)rq->n_scans) +
                                   me->device_bytes();
        if ((uint64_t)rq->n_scans > (uint64_t)s->n_slots)
            return fail(OCTMPS_E_NOMEM, "request has more A-scans than the server has scan slots");
        if (footprint > s->budget)
            return fail(OCTMPS_E_NOMEM, "request's footprint (mesh " + std::to_string(me->device_bytes()) +
                                            " B + tallies and state " + std::to_string(footprint - me->device_bytes()) +
                                            " B) exceeds the server's HBM budget");
        me->refs++;
    }
    Req *r = new (std::nothrow) Req();
    if (!r) {
        me->refs--;
        return fail(OCTMPS_E_NOMEM, "host allocation failed");
    }
    r->mesh_id = rq->mesh_id;
    r->me = me;
    r->counted = true;
    r->p = *rq->params;
    r->scans.assign(rq->scans, rq->scans + rq->n_scans);
    r->seed = rq->seed;
    r->out = rq->tallies_out;
    r->words = words;
    r->ss.resize((size_t)rq->n_scans);
    for (int i = 0; i < rq->n_scans; ++i) r->ss[i].n = r->ss[i].limit = rq->scans[i].n_photons;
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop || s->fatal != OCTMPS_OK) {
        const int fatal = s->fatal;
        me->refs--;
        delete r;
        if (fatal != OCTMPS_OK) return fail((octmps_status)fatal, "server stopped after a device error");
        return fail(OCTMPS_E_BUSY, "server is closing");
    }
    r->id = s->next_id++;
    r->index = s->submits++;
    r->fail_once = (r->index == s->o.fail_request_index);
    r->t_submit = now_ms() - s->t0;
    s->reqs[r->id] = r;
    s->pending.push_back(r);
    if (req_id) *req_id = r->id;
    s->cv_work.notify_all();
    return OCTMPS_OK;
}

static void fill_info(const Req *r, octmps_request_info *info) {
    if (!info) return;
    info->status = r->status;
    info->done = r->done ? 1 : 0;
    info->t_submit_ms = r->t_submit;
    info->t_admit_ms = r->t_admit;
    info->t_done_ms = r->t_done;
    info->photons = r->photons;
}

extern "C" octmps_status octmps_serve_poll(octmps_server *s, uint64_t id, octmps_request_info *info) {
    if (!s) return fail(OCTMPS_E_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return fail(OCTMPS_E_NOT_FOUND, "unknown request");
    fill_info(it->second, info);
    return OCTMPS_OK;
}

extern "C" octmps_status octmps_serve_wait(octmps_server *s, uint64_t id, double timeout_s, octmps_request_info *info) {
    if (!s) return fail(OCTMPS_E_INVALID, "null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return fail(OCTMPS_E_NOT_FOUND, "unknown request");
    Req *r = it->second;
    auto pred = [&] { return r->done; };
    if (timeout_s < 0) s->cv_done.wait(lk, pred);
    else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), pred)) {
        fill_info(r, info);
        return fail(OCTMPS_E_TIMEOUT, "request still running");
    }
    fill_info(r, info);
    return OCTMPS_OK;
}

extern "C" octmps_status octmps_serve_cancel(octmps_server *s, uint64_t id) {
    if (!s) return fail(OCTMPS_E_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return fail(OCTMPS_E_NOT_FOUND, "unknown request");
    Req *r = it->second;
    if (r->done) return OCTMPS_OK;
    r->cancelled = true;
    for (ScanState &x : r->ss) x.limit = x.claimed;  // hand out nothing more
    s->cv_work.notify_all();
    return OCTMPS_OK;
}

extern "C" octmps_status octmps_serve_release(octmps_server *s, uint64_t id) {
    if (!s) return fail(OCTMPS_E_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return fail(OCTMPS_E_NOT_FOUND, "unknown request");
    if (!it->second->done) return fail(OCTMPS_E_BUSY, "request still running");
    delete it->second;
    s->reqs.erase(it);
    return OCTMPS_OK;
}

extern "C" octmps_status octmps_serve_stats(octmps_server *s, uint64_t *photons, uint64_t *steps, double *kernel_ms,
                                            uint64_t *slices) {
    if (!s) return fail(OCTMPS_E_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    if (photons) *photons = s->st_photons;
    if (steps) *steps = s->st_steps;
    if (kernel_ms) *kernel_ms = s->st_kernel_ms;
    if (slices) *slices = s->st_slices;
    return OCTMPS_OK;
}

extern "C" octmps_status octmps_serve_residency(octmps_server *s, uint64_t *uploads, uint64_t *evictions,
                                                uint64_t *used_bytes, uint64_t *budget_bytes) {
    if (!s) return fail(OCTMPS_E_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    if (uploads) *uploads = s->st_uploads.load();
    if (evictions) *evictions = s->st_evictions.load();
    if (used_bytes) *used_bytes = s->used;
    if (budget_bytes) *budget_bytes = s->budget;
    return OCTMPS_OK;
}

extern "C" octmps_status octmps_serve_close(octmps_server *s) {
    if (!s) return fail(OCTMPS_E_INVALID, "null argument");
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv_work.notify_all();
    }
    if (s->th.joinable()) s->th.join();
    {
        std::lock_guard<std::mutex> ck(s->copy_mu);
        s->copy_stop = true;
        s->cv_copy.notify_all();
    }
    if (s->copy_th.joinable()) s->copy_th.join();
    (void)hipSetDevice(s->ctx->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (DevBuf *b : {&s->regions, &s->sets, &s->slots, &s->seg_start, &s->seg_slot, &s->seg_poff, &s->tallies,
                      &s->ctrl, &s->slot_done, &s->save, &s->cont, &s->set_map})
        b->release();
    if (s->e0) (void)hipEventDestroy(s->e0);
    if (s->e1) (void)hipEventDestroy(s->e1);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    if (s->copy_stream) (void)hipStreamDestroy(s->copy_stream);
    for (auto &kv : s->reqs) delete kv.second;
    delete s;
    return OCTMPS_OK;
}
