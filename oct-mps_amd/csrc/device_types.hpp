// Layouts shared by the HIP transport kernel (transport.hip) and the host runtime.
//
// Mesh in HBM: one 128-byte record per tet (TetRec), global tet ids across all resident
// meshes so one launch can walk several meshes:
//   plane[j] : float4 (nx, ny, nz, d), face j pre-signed so n.x + d >= 0 inside (64 B)
//   link[j]  : packed neighbour word of face j (below; 16 B)
//   region   : the tet's own region, read only through escaped links (4 B)
//   pad      : 44 B, so that a record is exactly one 128-byte cache line
// One tet visit is then 5 loads of 16 B (80 B) that all fall in one cache line. With the
// planes and the neighbour words in two arrays, a visit touched two lines, and the L1
// cache (which holds far less than the records of the photons in flight on a CU) missed
// on both.
// Reference layout being replaced: TetrahedronStructGPU + TriangleFaces (src/octmps.h:52-57,
// 193-196), read as 4 strided face gathers per step (src/octmps_kernel.cuh:517-545).
#pragma once

#include <stdint.h>

#include <hip/hip_runtime.h>

namespace octmps {

constexpr int kMaxRegions = 100;
// Parameter sets (and regions of each) a multi-set launch can keep in LDS, and the number
// of parameter set ids a request server hands out.
constexpr int kLdsSets = 16, kLdsSetRegions = 16, kMaxServeSets = 256;
// Region slots of the cooperative-fetch serving variant's LDS set table (its record stage
// leaves room for fewer than kLdsSets * kLdsSetRegions): 16 sets of up to 8 regions, or 8
// sets of up to 16, chosen per slice.
constexpr int kCoopRegionSlots = 128;
constexpr uint32_t kLinkBits = 25;                       // pool tet id bits of a packed link
constexpr uint32_t kLinkMask = (1u << kLinkBits) - 1u;   // also the boundary id
constexpr int64_t kMaxPoolTets = (int64_t)kLinkMask;     // ids 0 .. 2^25 - 2
// Packed neighbour word: bits 0-24 the neighbour's pool tet id, bits 25-26 the index of the
// shared face within the neighbour, bits 27-31 the neighbour's region, or kRegionEscape
// when that region is >= 31, when the neighbour does not link back through that face, or
// when the two copies of the shared plane are not exact negations (the kernel then reads
// the region from the neighbour's record and does not use the face index). 0xFFFFFFFF at
// the mesh boundary.
constexpr uint32_t kBackShift = 25;
constexpr uint32_t kRegionShift = 27;
constexpr uint32_t kRegionEscape = 31;

struct __align__(128) TetRec {
    float4 plane[4];
    uint32_t link[4];
    int32_t region;
    uint32_t pad[11];
};
static_assert(sizeof(TetRec) == 128, "TetRec must be one 128-byte line");

// Per-region constants of one parameter set (RegionStructGPU, src/octmps_kernel.h:132-140).
struct DevRegion {
    float n;         // refractive index
    float muas;      // mua + mus (0 for glass)
    float rmuas;     // 1/(mua+mus) (FLT_MAX for glass, like DBL_MAX in src/octmps_io.cpp:625)
    float mua_muas;  // mua/(mua+mus)
    float g;         // anisotropy
    float hg_a;      // 1 - g^2
    float hg_b;      // 1/(2g) (0 when g == 0)
    float lr8;       // (1-g^2) * (sqrt(1+a^2)-1+a) / (2a(1-a) sqrt(1+a^2))
};
static_assert(sizeof(DevRegion) == 32, "DevRegion must be 32 bytes");

// Per-parameter-set constants (SimParamGPU, src/octmps_kernel.h:112-130).
struct DevSet {
    float n_ambient;
    float bias_a;        // BackwardBiasCoefficient
    float td_min, td_max;
    float p_add;         // ProbabilityAdditionalBias
    float cos_angle;     // cos(MaxCollectingAngleDeg * pi / 180)
    float radius2;       // MaxCollectingRadius^2
    float next_step_num; // -log(rndStepSizeInTissue) (inf for 0)
    float shift;         // OpticalDepthShift
    float bin_w;         // CoherenceLengthSource / 6, rounded down
    float half_lc;       // CoherenceLengthSource / 2
    int32_t num_bins;
    int32_t type_bias;
    float inv_1ma;       // 1/(1-a)
    float inv_sq;        // 1/sqrt(1+a^2)
    float inv_2a;        // 1/(2a)
    float opa2;          // 1 + a^2
    float pdfb_k;        // sqrt(1+a^2) a (1-a) / (sqrt(1+a^2)-1+a)
    int32_t region_base; // index of region 0 of this set in the region table: always
                         // set * kMaxRegions (the kernel computes it from the set index)
    int32_t pad;
};
static_assert(sizeof(DevSet) == 80, "DevSet must be 80 bytes");

// One A-scan in flight.
struct DevScan {
    float px, py, pz;
    float w0;            // 1 - Rspecular of the root region
    int32_t root_tet;    // global tet id
    int32_t root_region;
    int32_t set;
    int32_t pad;
    uint64_t photon_offset;
    uint64_t salt;
    uint64_t seed;
    uint64_t tally_off;  // word offset of this scan's tally in the tally pool
};
static_assert(sizeof(DevScan) == 64, "DevScan must be 64 bytes");

// Live photon of one lane saved between slices (the reference's SaveThreadState,
// src/octmps_kernel.cuh:1154-1224, as one record per lane). Its continuation photon stays
// in the lane's ContRec, which persists across slices like this record.
struct SavedLane {
    float x, y, z, ux, uy, uz, w, tau, op, maxz, lr, lr_fst, loc_fst;
    int32_t tet, region;
    uint32_t flags, steps;
    uint32_t aG, aR, aS;
    uint64_t sG, sR, sS;
    uint32_t slot;
    uint64_t id;
    uint32_t ndet;
    float csum;
};

// Continuation photon (photon_cont, src/octmps_kernel.h:219-255) of one lane: the copy taken
// at the first backward bias and restored by roulette. Written and read rarely, so it lives
// in memory (one record per lane) rather than in eleven VGPRs.
struct __align__(16) ContRec {
    float x, y, z, w;
    float ux, uy, uz, op;
    float maxz;
    int32_t tet, region;
    uint32_t pad;
};
static_assert(sizeof(ContRec) == 48, "ContRec must be 48 bytes");

struct PhotonLogRec {
    uint32_t steps;
    uint32_t detections;
    float contrib;
    uint32_t killed;
};

// Control words.
// CTRL_BAD counts detector contributions the tally could not hold (see tally_detection).
enum { CTRL_NEXT_ID = 0, CTRL_STEPS = 1, CTRL_KILLED = 2, CTRL_BAD = 3, CTRL_ITERS = 4, CTRL_WORDS = 8 };

struct KArgs {
    const TetRec *tets;
    uint32_t tets_bytes;        // bytes of resident records (<= (2^25 - 1) * 128 < 2^32)
    uint32_t skip_entry;        // 1: the plane of a tet's entry face is not fetched
    uint32_t coop_fetch;        // 1: tet records fetched cooperatively into LDS (single-set;
                                // multi-set with the sets in LDS or in global memory)
    uint32_t n_regions;         // coop_fetch: regions of the set (its LDS region table)
    uint32_t sets_in_lds;       // multi-set launch: every lane's set is in the LDS map, regions < kLdsSetRegions
    uint32_t set_region_shift;  // sets_in_lds with coop_fetch: log2 of the region slots per set in
                                // the cooperative variant's LDS table of kCoopRegionSlots regions
                                // (3: 16 sets of up to 8 regions, 4: 8 sets of up to 16)
    const uint8_t *set_slot;    // sets_in_lds: LDS slot of each parameter set id [kMaxServeSets]
    const int32_t *lds_set_ids; // sets_in_lds: parameter set id held in each LDS slot [kLdsSets]
    const DevRegion *regions;
    const DevSet *sets;
    const DevScan *slots;
    const uint64_t *seg_start;  // [n_seg + 1] first global photon id of each segment
    const uint32_t *seg_slot;   // [n_seg] scan slot of each segment
    const uint64_t *seg_poff;   // [n_seg] photon index (within the scan) at the segment start
    const uint32_t *mult;       // MWC multiplier table: 3 seeding + 3 per triplet
    uint64_t *tallies;
    uint64_t *ctrl;
    uint64_t *slot_done;        // photons completed per slot
    SavedLane *save;            // lane states between slices (slice mode)
    PhotonLogRec *plog;         // optional per-photon summaries, indexed by global id
    ContRec *cont;              // continuation photon per lane (grid * block records)
    uint64_t id_end;            // ids [ctrl[NEXT], id_end) may be launched
    uint64_t iter_budget;       // slice mode: loop iterations per wave before saving
    uint64_t max_steps;         // per-photon step guard
    int32_t n_seg;
    uint32_t n_triplets;
    uint32_t refill_q;
    uint32_t slice_mode;        // 1: save live photons and exit when ids run out
    uint32_t restore;           // 1: start from the saved lane states
    uint32_t exit_on_drain;     // slice mode: leave as soon as the ids run out
    uint32_t scatter_q;         // parked lanes per wave before the scattering code runs
    uint32_t tally_copies;      // private tally copies; workgroup w adds into copy w % copies
    uint64_t copy_stride;       // words between two copies of the same tally
    uint64_t slice_ticks;       // slice mode: wall-clock length of a slice (device wall clock
                                // ticks, 0 = none); every wave leaves at about the same time
};

// Launchers implemented in transport.hip.
hipError_t launch_transport(const KArgs &a, int grid, int block, bool multi, bool log,
                            hipStream_t stream);
hipError_t launch_fill_u64(uint64_t *p, uint64_t n, uint64_t v, hipStream_t stream);
// Folds copies 1..copies-1 of n_scans tallies (each octmps_tally_words(bins) words, copy c at
// p + c*stride) into copy 0: sums, except the Max words, which take the maximum.
hipError_t launch_fold_tallies(uint64_t *p, uint64_t n_scans, int bins, uint32_t copies, uint64_t stride,
                               hipStream_t stream);
hipError_t launch_clear_saved(SavedLane *p, uint64_t n, hipStream_t stream);
// coop_regions > 0: the cooperative-fetch variant with that many regions in its LDS table.
int transport_blocks_per_cu(int block, bool multi, bool log, uint32_t coop_regions);
hipError_t launch_probe_rng(const uint32_t *mult, uint32_t n_triplets, uint64_t seed, uint64_t salt,
                            uint64_t photon0, int n_photons, int n_draws, float *out,
                            hipStream_t stream);
hipError_t launch_probe_fresnel(int n, const float *ni, const float *nt, const float *ca1, float *r,
                                hipStream_t stream);
hipError_t launch_probe_sampler(int kind, float param, int n, const float *u, float *out,
                                hipStream_t stream);
hipError_t launch_copy(const void *src, void *dst, uint64_t bytes, hipStream_t stream);
// n detections (op[i], maxz[i], c[i]) into the tally of a.slots[0] with parameter set S.
hipError_t launch_probe_tally(const KArgs &a, const DevSet &S, int n, const float *op, const float *maxz,
                              const float *c, hipStream_t stream);
// blocks x 256 lanes each read iters records (5 x 16 B) of recs[0..n).
hipError_t launch_gather_probe(const TetRec *recs, uint32_t n, int iters, uint32_t *out, int blocks,
                               hipStream_t stream);
// blocks x 256 lanes each issue iters x 8 loads of `pattern` from a 16 KiB buffer (buf).
hipError_t launch_l1_probe(int pattern, const uint8_t *buf, int iters, float *out, int blocks, hipStream_t stream);
// blocks x 256 lanes each run iters x 8 independent VALU operations of class `kind`.
hipError_t launch_valu_probe(int kind, int iters, float *out, int blocks, hipStream_t stream);

}  // namespace octmps
