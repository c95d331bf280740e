// Internal state of the device runtime (context, buffers, graph cache).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <atomic>
#include <map>
#include <memory>
#include <mutex>
#include <shared_mutex>
#include <string>
#include <vector>

#include "device_types.hpp"
#include "host_common.hpp"

namespace octmps {

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t n);
    void release();
};

struct WorkBuffers {
    DevBuf regions, sets, slots, seg_start, seg_slot, seg_poff, tallies, ctrl, slot_done, plog, cont;
    void release_all() {
        for (DevBuf *b : {&regions, &sets, &slots, &seg_start, &seg_slot, &seg_poff, &tallies, &ctrl, &slot_done, &plog, &cont})
            b->release();
    }
    uint64_t bytes() const {
        uint64_t s = 0;
        for (const DevBuf *b : {&regions, &sets, &slots, &seg_start, &seg_slot, &seg_poff, &tallies, &ctrl, &slot_done, &plog, &cont})
            s += b->cap;
        return s;
    }
};

// A mesh known to the context. Pinned meshes (octmps_ctx_upload_mesh) stay resident in the
// pool until released; registered meshes (octmps_ctx_register_mesh) keep their records on the
// host and are made resident on demand (a run, a server's admission) and evicted when idle.
struct MeshEntry {
    int64_t offset = 0;   // first global tet id in the pool (while resident)
    int64_t n_tets = 0;
    int32_t max_region = 0;
    bool live = false;      // uploaded or registered, not released
    bool resident = false;  // records in the pool
    bool pinned = false;
    std::atomic<int> refs{0};    // submitted requests not finished + runs in progress: no release
    std::atomic<int> active{0};  // admitted requests + runs in progress: no eviction
    double last_use = 0;
    std::vector<int32_t> region;  // host copy of tet regions (root region, Rspecular)
    std::vector<TetRec> host;     // registered meshes: records with mesh-relative neighbour ids
    uint64_t device_bytes() const { return (uint64_t)n_tets * sizeof(TetRec); }
};

struct GraphCache {
    hipGraphExec_t exec = nullptr;
    int grid = 0, block = 0;
    KArgs args;
};

octmps_status hip_fail(hipError_t e, const char *what);
double now_ms();
octmps_status make_set(const octmps_params *p, int max_region, int region_base, DevSet &S, DevRegion *R);

}  // namespace octmps

struct octmps_ctx {
    int device = 0;
    int n_cu = 0;
    int wall_khz = 0;  // rate of the device wall clock read by s_memrealtime
    uint64_t hbm_total = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    const octmps::StreamTable *table = nullptr;
    uint32_t *d_mult = nullptr;
    // Resident mesh pool (global tet ids).
    std::shared_mutex pool_mu;  // shared: launches and submits; exclusive: mesh upload/release
    octmps::TetRec *pool_tets = nullptr;
    int64_t pool_cap = 0;                  // tets the pool allocation holds
    std::map<int64_t, int64_t> pool_free;  // free ranges below pool_cap: first tet -> count
    std::vector<std::unique_ptr<octmps::MeshEntry>> meshes;
    // Plain/batched runs.
    std::mutex run_mu;
    octmps::WorkBuffers work;
    octmps::GraphCache graph;
    int tuned_block = 0, tuned_q = 0, tuned_sq = 0, tuned_copies = 0;
    // Device address and size of the allocation-free path's scratch array (noalloc.hip).
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
};

namespace octmps {
octmps_status make_scan(octmps_ctx *ctx, const MeshEntry &me, const octmps_params *p, const octmps_ascan &s,
                        uint64_t seed, int set, uint64_t tally_off, DevScan &d);
void grid_for(octmps_ctx *ctx, int block, int blocks_per_cu, bool multi, int &grid, uint32_t coop_regions = 0);
// Runs A-scans to completion in one persistent launch; device buffers come from the context's
// work buffers, or from `arena` (arena_bytes) when it is set (runtime.hip).
octmps_status run_scans(octmps_ctx *ctx, int32_t mesh_id, const octmps_params *p, int32_t n,
                        const octmps_ascan *scans, uint64_t seed, const octmps_batch_opts *opts,
                        uint64_t *tallies_out, octmps_run_stats *stats, PhotonLogRec *plog_out,
                        uint64_t plog_n, void *arena, size_t arena_bytes);
// Mesh pool (mesh_pool.hip). The caller holds ctx->pool_mu exclusively.
octmps_status mesh_make_resident(octmps_ctx *ctx, MeshEntry &me);
void mesh_evict(octmps_ctx *ctx, MeshEntry &me);
// Live mesh of an id, or nullptr (ctx->pool_mu held, shared or exclusive).
MeshEntry *mesh_at(octmps_ctx *ctx, int32_t id);
// For a plain or batched run: makes the mesh resident and counts the run as a user (refs and
// active); mesh_release_run undoes the counts. Takes ctx->pool_mu exclusively itself.
octmps_status mesh_acquire_run(octmps_ctx *ctx, int32_t id);
void mesh_release_run(octmps_ctx *ctx, int32_t id);
// Looks up the device address of the allocation-free path's scratch array (noalloc.hip);
// called by octmps_ctx_create with the context's device current.
octmps_status scratch_init(octmps_ctx *ctx);
}  // namespace octmps

// The entry-face skip pays when tet records come from cache: it is on by default for meshes
// whose records fit the 256 MB Infinity Cache (slab4, 101 MB: +2 %). On the 805 MB bump
// mesh the kernel without it is 30 % faster: there the skip raised the L1's misses by 13 %
// and its cycles stalled on L2 data fourfold.
inline bool entry_skip_fits(int64_t n_tets) { return n_tets * (int64_t)sizeof(octmps::TetRec) <= (256ll << 20); }
