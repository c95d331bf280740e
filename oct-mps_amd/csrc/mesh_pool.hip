// Resident mesh pool: one device array of 128-byte tet records (TetRec) for every resident
// mesh, with global tet ids so that one launch can walk several meshes. Meshes take
// first-fit ranges of the pool; a freed range is reused. The pool grows (a new allocation and
// a device copy) when no free range fits, up to 2^25 - 1 tets (the id bits of a packed
// neighbour word).
//
// Pinned meshes (octmps_ctx_upload_mesh) are resident until released. Registered meshes
// (octmps_ctx_register_mesh) keep their records on the host, with neighbour ids relative to
// the mesh; they are made resident on demand and may be evicted while nothing uses them. A
// record's neighbour ids are relocated on the device after each upload (a kernel adds the
// mesh's pool offset to every link), so the host copy never changes.
//
// Replaces the reference's per-A-scan cudaMalloc and copy of the mesh (RunGPUi,
// src/main.cu:100-465) and its cudaFree (src/main.cu:638).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "device_types.hpp"
#include "host_common.hpp"
#include "runtime_internal.hpp"

namespace octmps {
namespace {

// Adds `off` to the tet id of every link that is not the mesh boundary. Ids stay below
// 2^25 - 1 (placement never ends past kMaxPoolTets), so the add cannot reach the face and
// region fields above bit 24.
__global__ void relocate_links_kernel(TetRec *recs, int64_t n, uint32_t off) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        uint4 l = *reinterpret_cast<const uint4 *>(recs[i].link);
        l.x = l.x == 0xffffffffu ? l.x : l.x + off;
        l.y = l.y == 0xffffffffu ? l.y : l.y + off;
        l.z = l.z == 0xffffffffu ? l.z : l.z + off;
        l.w = l.w == 0xffffffffu ? l.w : l.w + off;
        *reinterpret_cast<uint4 *>(recs[i].link) = l;
    }
}

// Pool records of a host mesh with mesh-relative neighbour ids: planes rounded to FP32,
// packed neighbour words (neighbour id, the shared face's index in the neighbour, the
// neighbour's region), the tet's own region.
octmps_status build_records(const HostMesh &m, std::vector<TetRec> &recs) {
    const int64_t T = m.ntets();
    if (m.max_region >= kMaxRegions) return fail(OCTMPS_E_RANGE, "regions must be < 100");
    if (T > kMaxPoolTets) return fail(OCTMPS_E_RANGE, "a mesh may have at most 2^25 - 1 tets (packed neighbour ids)");
    recs.assign((size_t)T, TetRec());
    for (int64_t t = 0; t < T; ++t) {
        TetRec &r = recs[t];
        std::memset(&r, 0, sizeof r);
        for (int j = 0; j < 4; ++j) {
            const double *p = &m.planes[16 * t + 4 * j];
            r.plane[j] = make_float4((float)p[0], (float)p[1], (float)p[2], (float)p[3]);
        }
        r.region = m.region[t];
    }
    for (int64_t t = 0; t < T; ++t) {
        TetRec &r = recs[t];
        for (int j = 0; j < 4; ++j) {
            const int32_t l = m.links[4 * t + j];
            if (l < 0) {
                r.link[j] = 0xffffffffu;
                continue;
            }
            // The face of l that links back to t through the exactly negated plane: a photon
            // that crosses into l through it moves away from it (its cosine there is the
            // negation of a negative one), so its next exit is one of the other three.
            uint32_t back = 4;
            for (int k = 0; k < 4 && back == 4; ++k) {
                const float4 a = r.plane[j], b = recs[l].plane[k];
                if (m.links[4 * (int64_t)l + k] == t && a.x == -b.x && a.y == -b.y && a.z == -b.z && a.w == -b.w)
                    back = (uint32_t)k;
            }
            const uint32_t reg =
                (back < 4 && m.region[l] < (int32_t)kRegionEscape) ? (uint32_t)m.region[l] : kRegionEscape;
            r.link[j] = (reg << kRegionShift) | ((back & 3u) << kBackShift) | (uint32_t)l;
        }
    }
    return OCTMPS_OK;
}

// End of the last allocated range (tets below it may hold records).
int64_t pool_extent(const octmps_ctx *ctx) {
    int64_t e = 0;
    for (const auto &m : ctx->meshes)
        if (m && m->resident) e = std::max(e, m->offset + m->n_tets);
    return e;
}

void pool_release(octmps_ctx *ctx, int64_t off, int64_t n) {
    auto &F = ctx->pool_free;
    auto it = F.emplace(off, n).first;
    if (it != F.begin()) {
        auto prev = std::prev(it);
        if (prev->first + prev->second == it->first) {
            prev->second += it->second;
            F.erase(it);
            it = prev;
        }
    }
    auto next = std::next(it);
    if (next != F.end() && it->first + it->second == next->first) {
        it->second += next->second;
        F.erase(next);
    }
}

octmps_status pool_grow(octmps_ctx *ctx, int64_t need_end) {
    if (need_end > kMaxPoolTets)
        return fail(OCTMPS_E_RANGE, "resident meshes would exceed 2^25 - 1 tets (packed neighbour ids)");
    int64_t ncap = std::min<int64_t>(kMaxPoolTets, std::max<int64_t>(need_end, ctx->pool_cap * 2));
    TetRec *np = nullptr;
    if (hipMalloc(&np, (size_t)ncap * sizeof(TetRec)) != hipSuccess) {
        (void)hipGetLastError();
        ncap = need_end;
        if (hipMalloc(&np, (size_t)ncap * sizeof(TetRec)) != hipSuccess) {
            (void)hipGetLastError();
            return fail(OCTMPS_E_NOMEM, "mesh pool allocation failed");
        }
    }
    const int64_t used = pool_extent(ctx);
    hipError_t e;
    if (used > 0 &&
        (e = hipMemcpy(np, ctx->pool_tets, (size_t)used * sizeof(TetRec), hipMemcpyDeviceToDevice)) != hipSuccess) {
        (void)hipFree(np);
        return hip_fail(e, "mesh pool growth");
    }
    if (ctx->pool_tets) (void)hipFree(ctx->pool_tets);
    const int64_t old = ctx->pool_cap;
    ctx->pool_tets = np;
    ctx->pool_cap = ncap;
    pool_release(ctx, old, ncap - old);
    if (ctx->graph.exec) {  // the captured launch holds the old pool address
        (void)hipGraphExecDestroy(ctx->graph.exec);
        ctx->graph.exec = nullptr;
    }
    return OCTMPS_OK;
}

// First-fit range of n tets, growing the pool when none fits.
octmps_status pool_place(octmps_ctx *ctx, int64_t n, int64_t &off) {
    for (int pass = 0; pass < 2; ++pass) {
        for (auto it = ctx->pool_free.begin(); it != ctx->pool_free.end(); ++it) {
            if (it->second < n) continue;
            off = it->first;
            const int64_t rest = it->second - n;
            ctx->pool_free.erase(it);
            if (rest > 0) ctx->pool_free.emplace(off + n, rest);
            return OCTMPS_OK;
        }
        if (pass == 1) break;
        // Grow from the free range that ends at the pool's end, if there is one.
        int64_t start = ctx->pool_cap;
        if (!ctx->pool_free.empty()) {
            const auto last = std::prev(ctx->pool_free.end());
            if (last->first + last->second == ctx->pool_cap) start = last->first;
        }
        const octmps_status st = pool_grow(ctx, start + n);
        if (st != OCTMPS_OK) return st;
    }
    return fail(OCTMPS_E_NOMEM, "mesh pool placement failed");
}

// Evicts idle registered meshes, least recently used first, until `bytes` are freed or none
// is left (ctx->pool_mu held exclusively). Returns the bytes freed.
uint64_t evict_idle(octmps_ctx *ctx, uint64_t bytes, const MeshEntry *keep) {
    uint64_t freed = 0;
    while (freed < bytes) {
        MeshEntry *lru = nullptr;
        for (auto &m : ctx->meshes)
            if (m && m.get() != keep && m->live && m->resident && !m->pinned && m->active.load() == 0 &&
                (!lru || m->last_use < lru->last_use))
                lru = m.get();
        if (!lru) break;
        freed += lru->device_bytes();
        mesh_evict(ctx, *lru);
    }
    return freed;
}

}  // namespace

MeshEntry *mesh_at(octmps_ctx *ctx, int32_t id) {
    if (id < 0 || id >= (int32_t)ctx->meshes.size() || !ctx->meshes[id] || !ctx->meshes[id]->live) return nullptr;
    return ctx->meshes[id].get();
}

octmps_status mesh_make_resident(octmps_ctx *ctx, MeshEntry &me) {
    if (me.resident) return OCTMPS_OK;
    if ((int64_t)me.host.size() != me.n_tets) return fail(OCTMPS_E_NOT_FOUND, "mesh records are not on the host");
    if (hipError_t se = hipSetDevice(ctx->device); se != hipSuccess) return hip_fail(se, "set device");
    int64_t off = 0;
    octmps_status st = pool_place(ctx, me.n_tets, off);
    if (st == OCTMPS_E_NOMEM && evict_idle(ctx, me.device_bytes(), &me) > 0) st = pool_place(ctx, me.n_tets, off);
    if (st != OCTMPS_OK) return st;
    hipError_t e = hipMemcpy(ctx->pool_tets + off, me.host.data(), me.host.size() * sizeof(TetRec),
                             hipMemcpyHostToDevice);
    if (e == hipSuccess && off > 0) {
        const int64_t blocks = std::min<int64_t>(4096, (me.n_tets + 255) / 256);
        hipLaunchKernelGGL(relocate_links_kernel, dim3((unsigned)blocks), dim3(256), 0, ctx->stream,
                           ctx->pool_tets + off, me.n_tets, (uint32_t)off);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    }
    if (e != hipSuccess) {
        pool_release(ctx, off, me.n_tets);
        return hip_fail(e, "mesh upload");
    }
    me.offset = off;
    me.resident = true;
    me.last_use = now_ms();
    return OCTMPS_OK;
}

void mesh_evict(octmps_ctx *ctx, MeshEntry &me) {
    if (!me.resident || me.pinned || me.active.load() > 0) return;
    pool_release(ctx, me.offset, me.n_tets);
    me.resident = false;
}

octmps_status mesh_acquire_run(octmps_ctx *ctx, int32_t id) {
    std::unique_lock<std::shared_mutex> pk(ctx->pool_mu);
    MeshEntry *me = mesh_at(ctx, id);
    if (!me) return fail(OCTMPS_E_NOT_FOUND, "unknown mesh id");
    const octmps_status st = mesh_make_resident(ctx, *me);
    if (st != OCTMPS_OK) return st;
    me->refs++;
    me->active++;
    me->last_use = now_ms();
    return OCTMPS_OK;
}

void mesh_release_run(octmps_ctx *ctx, int32_t id) {
    std::shared_lock<std::shared_mutex> pk(ctx->pool_mu);
    if (MeshEntry *me = mesh_at(ctx, id)) {
        me->active--;
        me->refs--;
    }
}

}  // namespace octmps

using namespace octmps;

namespace {
octmps_status add_mesh(octmps_ctx *c, const octmps_mesh *mh, bool pinned, int32_t *mesh_id) {
    if (!c || !mh || !mesh_id) return fail(OCTMPS_E_INVALID, "null argument");
    auto me = std::make_unique<MeshEntry>();
    octmps_status st = build_records(mh->m, me->host);
    if (st != OCTMPS_OK) return st;
    me->n_tets = mh->m.ntets();
    me->max_region = mh->m.max_region;
    me->region = mh->m.region;
    me->pinned = pinned;
    me->live = true;
    std::unique_lock<std::shared_mutex> pk(c->pool_mu);
    if (pinned) {
        st = mesh_make_resident(c, *me);
        if (st != OCTMPS_OK) return st;
        me->host.clear();
        me->host.shrink_to_fit();
    }
    c->meshes.push_back(std::move(me));
    *mesh_id = (int32_t)c->meshes.size() - 1;
    return OCTMPS_OK;
}
}  // namespace

extern "C" octmps_status octmps_ctx_upload_mesh(octmps_ctx *c, const octmps_mesh *mh, int32_t *mesh_id) {
    return add_mesh(c, mh, true, mesh_id);
}

extern "C" octmps_status octmps_ctx_register_mesh(octmps_ctx *c, const octmps_mesh *mh, int32_t *mesh_id) {
    return add_mesh(c, mh, false, mesh_id);
}

extern "C" octmps_status octmps_ctx_release_mesh(octmps_ctx *c, int32_t id) {
    if (!c) return fail(OCTMPS_E_INVALID, "null argument");
    std::unique_lock<std::shared_mutex> pk(c->pool_mu);
    MeshEntry *me = mesh_at(c, id);
    if (!me) return fail(OCTMPS_E_NOT_FOUND, "unknown mesh id");
    if (me->refs.load() > 0) return fail(OCTMPS_E_BUSY, "requests or runs still use the mesh");
    if (me->resident) pool_release(c, me->offset, me->n_tets);
    me->resident = false;
    me->live = false;
    me->region.clear();
    me->region.shrink_to_fit();
    me->host.clear();
    me->host.shrink_to_fit();
    return OCTMPS_OK;
}

extern "C" octmps_status octmps_ctx_mesh_info(octmps_ctx *c, int32_t id, int32_t *resident, uint64_t *device_bytes) {
    if (!c) return fail(OCTMPS_E_INVALID, "null argument");
    std::shared_lock<std::shared_mutex> pk(c->pool_mu);
    MeshEntry *me = mesh_at(c, id);
    if (!me) return fail(OCTMPS_E_NOT_FOUND, "unknown mesh id");
    if (resident) *resident = me->resident ? 1 : 0;
    if (device_bytes) *device_bytes = me->device_bytes();
    return OCTMPS_OK;
}

extern "C" octmps_status octmps_ctx_memory(octmps_ctx *c, uint64_t *used, uint64_t *total) {
    if (!c) return fail(OCTMPS_E_INVALID, "null argument");
    std::shared_lock<std::shared_mutex> pk(c->pool_mu);
    uint64_t u = (uint64_t)c->pool_cap * sizeof(TetRec) + c->work.bytes();
    if (used) *used = u;
    if (total) *total = c->hbm_total;
    return OCTMPS_OK;
}
