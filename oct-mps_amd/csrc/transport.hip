// Photon transport kernel for gfx950 (CDNA4, wave64).
//
// One persistent kernel runs the reference's per-photon loop (OCTMPSKernel,
// src/octmps_kernel.cuh:1329-1402) for any mix of A-scans: each lane holds one photon in
// VGPRs; when enough lanes of a wave have lost their photon (refill quantum), the wave claims
// that many photon ids with one atomic and relaunches those lanes (ballot + mbcnt ranks).
// Photon ids map to A-scans through a sorted segment table, and every photon derives its
// three MWC streams from (seed, salt, photon index), so results do not depend on lane,
// launch or slice. Transport is FP32, compiled without FP contraction: every fused
// multiply-add is written as fmaf, so each kernel variant rounds identically. Tallies are
// exact fixed-point integers (32-bit digit pieces in 64-bit words, carry-free,
// non-returning atomics; the reference's EMULATED_ATOMIC, src/octmps_kernel.cuh:42-46,
// needed a returning add to propagate its carry).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>

#include "device_types.hpp"
#include "octmps.h"
#include "stream_init.hpp"

namespace octmps {
namespace {

// FP32 thresholds. The reference's COSZERO = 1 - 1e-14 (src/octmps.h:29) rounds to 1.0f,
// which would send exact normal incidence down the 0/0 Fresnel branch, so FP32 uses
// 1 - 1e-6; a direction is "normal" for the rotation when ux^2 + uy^2 <= 1e-12.
constexpr float kCosZeroF = 1.0f - 1e-6f;
constexpr float kCos90F = 1e-6f;        // COSNINETYDEG, src/octmps.h:28
constexpr float kWeightF = 1e-4f;       // WEIGHT, src/octmps.h:22
constexpr float kChanceF = 0.1f;        // CHANCE, src/octmps.h:23
constexpr float kNormalSin2 = 1e-12f;

// Lane flags: photon alive, first backward bias done, a backward specular reflection seen,
// the continuation photon's copy of that bit, scatter pending, interface pending (the exit
// face index 0-3 in bits kFaceShift..kFaceShift+1).
constexpr uint32_t F_ALIVE = 1u, F_FST = 2u, F_NBS = 4u, F_CNBS = 8u, F_PEND = 16u, F_IFACE = 32u;
constexpr uint32_t kFaceShift = 6, F_FACE = 3u << kFaceShift;
// Entry face of the current tet (0: none, k + 1: the photon crossed into it through face k).
constexpr uint32_t kEntryShift = 8, F_ENTRY = 7u << kEntryShift;
// Cooperative fetch: the lane's LDS stage holds the planes and links of its current tet.
constexpr uint32_t F_STAGED = 1u << 11;

typedef uint32_t u4v __attribute__((ext_vector_type(4)));
// Buffer-resource word 3 for raw (untyped) loads on CDNA: 32-bit data format.
constexpr int kBufferWord3 = 0x00020000;
// A record offset past every resident record (tets_bytes <= 2^32 - 128): a raw buffer load
// there is out of range, so it makes no memory request and returns zeros.
constexpr uint32_t kOutOfRange = 0xFFFFFF80u;
// Cooperative fetch: a wave's LDS stage holds four 64-lane plane loads at float4 offsets 0,
// 65, 130 and 195 (the one-float4 gaps spread consecutive owners over the LDS banks), then
// the links of lane l at kStageLinks + l.
constexpr uint32_t kStageLinks = 260, kStageF4 = kStageLinks + 64;

// __uint2float_rz(u) / 2^32: keep the top 24 significant bits, then an exact scale.
__device__ __forceinline__ float unit_rz24(uint32_t u) {
    int shift = 8 - __clz((int)u);
    shift = shift < 0 ? 0 : shift;
    uint32_t t = u & (0xffffffffu << shift);
    return (float)t * 2.3283064365386963e-10f;
}

// One multiply-with-carry step; returns the low word that the draw converts.
__device__ __forceinline__ uint32_t mwc_next(uint64_t &x, uint32_t a) {
    x = (x & 0xffffffffull) * (uint64_t)a + (x >> 32);
    return (uint32_t)x;
}

// rand_MWC_co / rand_MWC_oc, src/octmps_kernel.cuh:10-29.
__device__ __forceinline__ float rand_co(uint64_t &x, uint32_t a) { return unit_rz24(mwc_next(x, a)); }
__device__ __forceinline__ float rand_oc(uint64_t &x, uint32_t a) { return 1.0f - rand_co(x, a); }

// Hardware reciprocal (v_rcp_f32, 1 ulp); the reference divides with round-down FP64
// division (__ddiv_rd), so no FP32 rounding mode is "the" reference one.
__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// Hardware square root (v_sqrt_f32, 1 ulp) instead of the correctly rounded sequence that
// sqrtf compiles to; arguments here are clamped to >= 0.
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// -ln(u) for u in (0, 1] with the hardware base-2 logarithm (v_log_f32); u >= 2^-24 here,
// so no denormal scaling is needed.
__device__ __forceinline__ float neg_ln(float u) { return -0.69314718055994531f * __builtin_amdgcn_logf(u); }

// Fresnel reflectance of FastReflectTransmit, src/octmps_kernel.cuh:608-645.
__device__ __forceinline__ float fresnel_f(float ni, float nt, float ca1) {
    float ni_nt = ni * rcp(nt);
    bool normal = ca1 > kCosZeroF;
    float sa1 = normal ? 0.0f : fsqrt(fmaxf(fmaf(-ca1, ca1, 1.0f), 0.0f));
    float sa2 = fminf(ni_nt * sa1, 1.0f);
    float ca2 = fsqrt(fmaxf(fmaf(-sa2, sa2, 1.0f), 0.0f));
    float ca1ca2 = ca1 * ca2, sa1sa2 = sa1 * sa2;
    float sa1ca2 = normal ? 1.0f : sa1 * ca2;
    float ca1sa2 = normal ? ni_nt : ca1 * sa2;
    float cam = ca1ca2 + sa1sa2, sap = sa1ca2 + ca1sa2, sam = sa1ca2 - ca1sa2;
    float r = sam * rcp(sap * cam);
    r *= r;
    r *= fmaf(ca1ca2, ca1ca2, sa1sa2 * sa1sa2);
    if (ca1 < kCos90F || sa2 == 1.0f) r = 1.0f;
    return r;
}

// The samplers and the rotation below compute every case and select the result, so a wave
// whose lanes take different cases runs no divergent branch (the loop is bound by
// instruction issue, and a divergent branch costs several scalar instructions).

// SpinTheta, src/octmps_kernel.cuh:807-820 (hg_a = 1-g^2, hg_b = 1/(2g); g == 0 gives
// hg_a = 1, hg_b = 0, so the discarded general case stays finite).
__device__ __forceinline__ float hg_cos(float g, float hg_a, float hg_b, float u_oc) {
    const float t = hg_a * rcp(fmaf(2.0f * g, u_oc, 1.0f - g));
    const float c = fminf(fmaxf(fmaf(-t, t, fmaf(g, g, 1.0f)) * hg_b, -1.0f), 1.0f);
    return g == 0.0f ? 2.0f * u_oc - 1.0f : c;
}

// SpinThetaForwardFstBias, src/octmps_kernel.cuh:881-896.
__device__ __forceinline__ float bias_cos(float a, float inv_1ma, float inv_sq, float inv_2a,
                                          float opa2, float u_oc) {
    const float t = fmaf(u_oc, inv_1ma, (1.0f - u_oc) * inv_sq);
    const float c = fminf(fmaxf((opa2 - rcp(t * t)) * inv_2a, -1.0f), 1.0f);
    return a == 0.0f ? u_oc : c;
}

// Rotation of Spin / SpinBias (src/octmps_kernel.cuh:850-866, 1009-1018) around axis
// (ax, ay, az), written with sqrt(ax^2+ay^2) in place of sqrt(1-az^2) and renormalised
// like the reference's SINGLE_PRECISION variant (src/octmps_kernel.cuh:870-877).
// copysign_mode: 0 -> uz = cost*sign(az) (Spin), 1 -> copysign(cost, az*cost) (SpinBias).
// The reciprocal square roots are v_rsq_f32 on normal arguments (sp2 > 1e-12 where the
// general case is kept, a squared norm near 1), where it equals rsqrtf bit for bit.
__device__ __forceinline__ void rotate(float ax, float ay, float az, float cost, float r_psi,
                                       int copysign_mode, float &ux, float &uy, float &uz) {
    const float sint = fsqrt(fmaxf(fmaf(-cost, cost, 1.0f), 0.0f));
    // psi = 2 pi r: v_sin_f32 / v_cos_f32 take their argument in revolutions.
    const float sinp = __builtin_amdgcn_sinf(r_psi), cosp = __builtin_amdgcn_cosf(r_psi);
    const float stcp = sint * cosp, stsp = sint * sinp;
    const float sp2 = fmaf(ax, ax, ay * ay);
    const bool normal = sp2 <= kNormalSin2;
    const float rs = __builtin_amdgcn_rsqf(sp2);
    const float gx = fmaf(fmaf(stcp * ax, az, -stsp * ay), rs, ax * cost);
    const float gy = fmaf(fmaf(stcp * ay, az, stsp * ax), rs, ay * cost);
    const float gz = fmaf(-stcp, sp2 * rs, az * cost);
    const float zn = copysign_mode ? copysignf(cost, az * cost) : cost * (az >= 0.0f ? 1.0f : -1.0f);
    const float nx = normal ? stcp : gx, ny = normal ? stsp : gy, nz = normal ? zn : gz;
    const float rn = __builtin_amdgcn_rsqf(fmaf(nx, nx, fmaf(ny, ny, nz * nz)));
    ux = nx * rn;
    uy = ny * rn;
    uz = nz * rn;
}

// Exact fixed-point pieces of a detector contribution c (FP32, c = m * 2^(e-150) with the
// 24-bit significand m): Sum gets m << (e + 10) in units of 2^-160 and SumSq gets
// m*m << (2e + 20) in units of 2^-320, cut at 32-bit boundaries into pieces that go to
// consecutive digit words (include/octmps.h). Every positive finite FP32 value and its
// square are then held exactly: Sum spans 2^-149 .. 2^128 in 9 digits and SumSq 2^-298 ..
// 2^256 in 18. A piece is < 2^32 and no carry crosses words, so each add is a
// non-returning atomic and a word cannot wrap before 2^32 adds.
struct ExactPieces {
    int ks, kq;            // first Sum digit, first SumSq digit
    uint32_t s0, s1;       // Sum pieces for digits ks, ks + 1
    uint32_t q0, q1, q2;   // SumSq pieces for digits kq, kq + 1, kq + 2
};

__device__ __forceinline__ ExactPieces exact_pieces(float c) {
    ExactPieces P;
    const uint32_t b = __float_as_uint(c);
    int e = (int)((b >> 23) & 0xffu);
    uint32_t m = b & 0x7fffffu;
    if (e == 0) e = 1;  // subnormal: no implicit bit, exponent of the smallest normal
    else m |= 0x800000u;
    const int sh = e + 10;        // 11 .. 264
    P.ks = sh >> 5;               // 0 .. 8; s1 != 0 needs (sh & 31) >= 9, so ks + 1 <= 8
    const uint64_t v = (uint64_t)m << (sh & 31);
    P.s0 = (uint32_t)v;
    P.s1 = (uint32_t)(v >> 32);
    const uint64_t M = (uint64_t)m * (uint64_t)m;  // < 2^48
    const int sh2 = 2 * e + 20;   // 22 .. 528
    const int r2 = sh2 & 31;
    P.kq = sh2 >> 5;              // 0 .. 16; q2 != 0 needs r2 >= 17, so kq + 2 <= 17
    const uint64_t lo = M << r2;
    P.q0 = (uint32_t)lo;
    P.q1 = (uint32_t)(lo >> 32);
    P.q2 = r2 ? (uint32_t)(M >> (64 - r2)) : 0u;
    return P;
}

__device__ __forceinline__ void add_piece(uint64_t *p, uint32_t v) {
    if (v) atomicAdd((unsigned long long *)p, (unsigned long long)v);
}

// Detector tally, src/octmps_kernel.cuh:684-741: six bins around the optical-depth bin,
// Class I when MaxDepth > OP/2 - Lc/2. A contribution that is not finite (an overflowed
// likelihood ratio) cannot be held: it is counted in word 3 of the tally and in
// ctrl[CTRL_BAD], and the bins still count the detection.
__device__ void tally_detection(const KArgs &A, const DevSet &S, uint32_t slot, float op, float maxz,
                                float c) {
    // Tally privatisation: workgroup w adds into copy w % copies (folded after the launch).
    const uint64_t copy = A.tally_copies > 1 ? (uint64_t)(blockIdx.x % A.tally_copies) : 0;
    uint64_t *T = A.tallies + A.slots[slot].tally_off + copy * A.copy_stride;
    atomicAdd((unsigned long long *)&T[0], 1ull);
    float depth = op * 0.5f;
    float q = (depth - S.shift) / S.bin_w;
    int vp = q <= 0.0f ? 0 : (q >= 1.0e9f ? 1000000000 : (int)q);
    bool c1 = maxz > depth - S.half_lc;
    int cls = c1 ? 0 : 1;
    const uint64_t B = (uint64_t)S.num_bins;
    const bool held = c >= 0.0f && c <= FLT_MAX;  // false for NaN and infinity
    if (!held) {
        atomicAdd((unsigned long long *)&T[3], 1ull);
        atomicAdd((unsigned long long *)&A.ctrl[CTRL_BAD], 1ull);
    }
    ExactPieces P = exact_pieces(held ? c : 0.0f);
    if (!(c > 0.0f) || !held) P.s0 = P.s1 = P.q0 = P.q1 = P.q2 = 0u;
    const uint64_t maxbits = (uint64_t)__double_as_longlong((double)(held ? c : 0.0f));
    // Per class, B words each: 9 Sum digits, 18 SumSq digits, count, Max.
    uint64_t *base = T + 4 + (uint64_t)cls * OCTMPS_TALLY_CLASS_WORDS * B;
    uint64_t *sum = base + (uint64_t)P.ks * B;
    uint64_t *sq = base + (uint64_t)(OCTMPS_TALLY_SUM_DIGITS + P.kq) * B;
    for (int ii = vp - 3; ii < vp + 3; ++ii) {
        if (ii >= 0 && ii < (int)B) {
            if (ii == vp) atomicAdd((unsigned long long *)&T[1 + cls], 1ull);
            add_piece(&sum[ii], P.s0);
            add_piece(&sum[B + ii], P.s1);
            add_piece(&sq[ii], P.q0);
            add_piece(&sq[B + ii], P.q1);
            add_piece(&sq[2 * B + ii], P.q2);
            atomicAdd((unsigned long long *)&base[(OCTMPS_TALLY_CLASS_WORDS - 2) * B + ii], 1ull);
            atomicMax((unsigned long long *)&base[(OCTMPS_TALLY_CLASS_WORDS - 1) * B + ii],
                      (unsigned long long)maxbits);
        }
    }
}

// Value of lane `src` (wave-uniform) as a scalar: the refill decisions made from it stay
// uniform branches, and the loop's exit stays a uniform branch.
__device__ __forceinline__ uint64_t readlane_u64(uint64_t v, int src) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, src);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), src);
    return ((uint64_t)hi << 32) | lo;
}

// Six waves per SIMD (<= 80 VGPRs, no spills). The loop is bound by the per-lane cost of its
// tet-record gathers rather than by their latency: five waves per SIMD run as fast, and
// seven are slower (more photons in flight per CU thrash the L1).
// SKIP: the entry-face plane is not fetched (its load goes out of range). Without SKIP all
// four plane loads share one offset register, which the gather path serves markedly faster
// when many records miss the caches (an 805 MB mesh ran 30 % faster than with SKIP, and 50 %
// faster than with the same offsets computed per plane); with SKIP the smaller gather wins
// when records come from cache (+2 % on a 101 MB mesh). The launch picks one per mesh.
// SETS_LDS (MULTI only): up to kLdsSets parameter sets (A.lds_set_ids) and their regions
// 0..kLdsSetRegions-1 are copied to LDS at launch, and every lane's set must be among them;
// a lane then carries its set's LDS slot (A.set_slot) in place of the set id. Otherwise
// each lane reads its set's records from memory, a per-lane gather in the parked phase
// that made serving 25 % slower than a single-set launch.
//
// COOP (single-set only): tet records are fetched cooperatively into a per-wave LDS stage.
// A 16-byte vector load costs the L1 one lookup cycle per lane when the 64 lanes of an
// instruction hit 64 different lines, while four lanes reading the four 16-byte quarters of
// one 64-byte piece share one cycle. So four DMA loads (buffer_load ... lds) give quad q of
// instruction j the four planes of the tet of its lane j (a DPP broadcast of that lane's
// record offset), and one more gives each lane its tet's links: two L1 cycles per record
// instead of five. A lane's stage keeps its tet (F_STAGED) until the photon moves to
// another tet, so a photon that scattered, or was reflected, in its tet reads it again from
// LDS, and a photon that moved fetches its new tet at the end of the iteration, ahead of
// the crossing phase that needs it. Parked lanes read their exit face from the stage.
template <bool MULTI, bool LOG, bool SKIP, bool SETS_LDS, bool COOP = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MULTI ? 5 : 6))) transport_kernel(KArgs A) {
    static_assert(!(COOP && SKIP), "the cooperative fetch does not skip the entry plane");
    const uint32_t lane = __lane_id();
    const uint64_t gl = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;

    // Photon (PhotonStructGPU, src/octmps_kernel.h:219-255) with sleft kept as tau.
    float x = 0, y = 0, z = 0, ux = 0, uy = 0, uz = 1, w = 0, tau = 0, op = 0, maxz = 0;
    float lr = 1, lr_fst = 1, loc_fst = -1;
    int tet = 0, region = 1;
    uint32_t flags = 0, steps = 0;
    // The continuation photon (photon_cont) is this lane's ContRec.
    ContRec &C = A.cont[gl];
    // Streams: General, ReflectT, Spin.
    uint64_t sG = 1, sR = 1, sS = 1;
    uint32_t aG = 2, aR = 2, aS = 2;
    // A-scan of the photon.
    uint32_t slot = 0;
    uint64_t id = 0;
    int set = 0;
    // Current region constants kept in registers.
    float rn = 1, rmuas = 1, muas = 1;
    uint32_t ndet = 0;
    float csum = 0;

    // One parameter set per launch unless MULTI: read it once (uniform, scalar registers),
    // and keep its region table in LDS, since a photon changes region on a large share of
    // its steps in layered tissue. MULTI with SETS_LDS keeps the first kLdsSets sets and
    // their region tables in LDS; MULTI without it reads each lane's set from memory.
    const DevSet S0 = A.sets[0];
    // The LDS set table: set slot s holds its regions at s << set_shift (fewer slots beside
    // the cooperative fetch's stage, split per launch between sets and regions per set).
    constexpr int kRegionSlots =
        MULTI ? (SETS_LDS ? (COOP ? kCoopRegionSlots : kLdsSets * kLdsSetRegions) : 1) : (COOP ? 1 : kMaxRegions);
    const uint32_t set_shift = COOP ? A.set_region_shift : 4u;
    static_assert(kLdsSetRegions == 16, "per-lane set table: 16 regions per set (shift 4)");
    __shared__ DevRegion lds_regions[kRegionSlots];
    __shared__ DevSet lds_sets[MULTI && SETS_LDS ? kLdsSets : 1];
    // COOP: the stages of the (at most four) waves in their own LDS object, so that the
    // compiler does not make reads of the other LDS data wait for the stage loads; a probe
    // position per lane and the set's region table in dynamic LDS sized by the launch.
    __shared__ float4 lds_stage[COOP ? 4 * kStageF4 : 1];
    extern __shared__ float4 lds_dyn[];
    DevRegion *const coop_regions = reinterpret_cast<DevRegion *>(lds_dyn + blockDim.x);
    if (!MULTI) {
        if (COOP)
            for (int i = threadIdx.x; i < (int)A.n_regions; i += blockDim.x) coop_regions[i] = A.regions[S0.region_base + i];
        else
            for (int i = threadIdx.x; i < kMaxRegions; i += blockDim.x) lds_regions[i] = A.regions[S0.region_base + i];
        __syncthreads();
    } else if (SETS_LDS) {
        for (int i = threadIdx.x; i < kRegionSlots; i += blockDim.x)
            lds_regions[i] = A.regions[A.lds_set_ids[i >> set_shift] * kMaxRegions + (i & ((1 << set_shift) - 1))];
        for (int i = threadIdx.x; i < kLdsSets; i += blockDim.x) lds_sets[i] = A.sets[A.lds_set_ids[i]];
        __syncthreads();
    }
    // Per-lane LDS stash (blocks have at most 256 lanes): the probe position of the lane's
    // A-scan, written at launch, and the exit face of a lane parked at an interface (normal
    // and packed neighbour word), written when it parks, so that the parked phase reads them
    // from LDS instead of making a dependent global round trip.
    // COOP reads the face from the lane's stage; the single-set COOP variant keeps the probe
    // in dynamic LDS.
    constexpr bool kProbeDyn = COOP && !MULTI;
    __shared__ float4 lds_probe_s[kProbeDyn ? 1 : 256], lds_face[COOP ? 1 : 256];
    auto lds_probe_at = [&]() -> float4 & {
        return kProbeDyn ? lds_dyn[threadIdx.x] : lds_probe_s[threadIdx.x];
    };
    // This lane's four planes in the stage: owner l was fetched by instruction l & 3, quad
    // l >> 2, at 65 (l & 3) + 4 (l >> 2).
    const float4 *const own_planes = lds_stage + (threadIdx.x >> 6) * kStageF4 + 65u * (lane & 3u) + (lane & ~3u);
    const uint4 *const own_links = reinterpret_cast<const uint4 *>(lds_stage + (threadIdx.x >> 6) * kStageF4 + kStageLinks + lane);
    auto region_rec = [&](int r) -> const DevRegion & {
        // The region table of set s starts at s * kMaxRegions (DevSet::region_base).
        return MULTI ? (SETS_LDS ? lds_regions[(set << set_shift) + r] : A.regions[set * kMaxRegions + r])
                     : (COOP ? coop_regions[r] : lds_regions[r]);
    };
    auto load_region = [&](int r) {
        const DevRegion &R = region_rec(r);
        rn = R.n;
        rmuas = R.rmuas;
        muas = R.muas;
    };

    // Slices (save/restore of live photons between launches) serve the request server, which
    // always launches the MULTI variant; the single-set variants run in drain mode only.
    constexpr bool SLICES = MULTI;
    if (SLICES && A.restore) {
        const SavedLane &s = A.save[gl];
        if (s.flags & F_ALIVE) {
            x = s.x; y = s.y; z = s.z; ux = s.ux; uy = s.uy; uz = s.uz; w = s.w; tau = s.tau;
            op = s.op; maxz = s.maxz; lr = s.lr; lr_fst = s.lr_fst; loc_fst = s.loc_fst;
            tet = s.tet; region = s.region; flags = s.flags; steps = s.steps;
            aG = s.aG; aR = s.aR; aS = s.aS; sG = s.sG; sR = s.sR; sS = s.sS;
            slot = s.slot; id = s.id; ndet = s.ndet; csum = s.csum;
            set = MULTI ? A.slots[slot].set : 0;
            if (MULTI && SETS_LDS) set = A.set_slot[set];
            load_region(region);
            const DevScan &sc = A.slots[slot];
            lds_probe_at() = make_float4(sc.px, sc.py, sc.pz, 0.0f);
            if (COOP) {
                // The stage of this launch holds nothing yet: every restored photon is fetched
                // below, before the loop (a parked one reads its exit face from the stage).
                flags &= ~F_STAGED;
            } else if (flags & F_IFACE) {
                const uint32_t f = (flags & F_FACE) >> kFaceShift;
                const float4 pl = A.tets[tet].plane[f];
                lds_face[threadIdx.x] = make_float4(pl.x, pl.y, pl.z, __uint_as_float(A.tets[tet].link[f]));
            }
        }
    }

    bool drained = false;
    uint64_t wave_steps = 0, iters = 0;
    const uint64_t t_start = (SLICES && A.slice_ticks) ? __builtin_amdgcn_s_memrealtime() : 0;
    // Step guard as a 32-bit count (the per-photon counter is 32-bit).
    const uint32_t max_steps = A.max_steps > 0xffffffffull ? 0xffffffffu : (uint32_t)A.max_steps;
    const uint32_t q = A.refill_q < 1u ? 1u : A.refill_q;
    const __amdgpu_buffer_rsrc_t tets_rsrc =
        __builtin_amdgcn_make_buffer_rsrc((void *)A.tets, (short)0, (int)A.tets_bytes, kBufferWord3);

    // Cooperative fetch (COOP; call with the whole wave): stages the planes and links of the
    // tets of the lanes with `want` set. Instruction j: quad q fetches the planes of the tet of
    // its lane j (owner 4q + j, a DPP broadcast within the quad), lane l the plane l & 3; then
    // every wanting lane fetches its own links. The loads complete into LDS (vmcnt).
    auto coop_fetch = [&](bool want) {
        if (__ballot(want) == 0) return;
        const int own = want ? (int)((uint32_t)tet << 7) : -1;
        const int r0 = __builtin_amdgcn_mov_dpp(own, 0x00, 0xF, 0xF, false);
        const int r1 = __builtin_amdgcn_mov_dpp(own, 0x55, 0xF, 0xF, false);
        const int r2 = __builtin_amdgcn_mov_dpp(own, 0xAA, 0xF, 0xF, false);
        const int r3 = __builtin_amdgcn_mov_dpp(own, 0xFF, 0xF, 0xF, false);
        const uint32_t quarter = (lane & 3u) << 4;
        // The five loads run under their own lane masks, set in one asm block: written as five
        // masked builtins, each mask costs a save, a branch and a restore of the exec mask
        // (scalar instructions, which the loop is short of). M0 holds the LDS destination of
        // each load (the wave's stage, wave-uniform); the LDS writes complete under vmcnt,
        // which the crossing phase waits for before it reads the stage. The exec mask is
        // restored before the block ends. The buffer resource is the one of tets_rsrc: base
        // address, stride 0, tets_bytes records (out-of-range offsets read nothing), raw dwords.
        typedef __attribute__((address_space(3))) void *lds_ptr;
        const uint32_t sb = (uint32_t)(uintptr_t)(lds_ptr)lds_stage +
                            __builtin_amdgcn_readfirstlane(threadIdx.x >> 6) * (kStageF4 * 16u);
        const uint64_t k0 = __ballot(r0 != -1), k1 = __ballot(r1 != -1), k2 = __ballot(r2 != -1),
                       k3 = __ballot(r3 != -1), k4 = __ballot(want);
        const uint64_t rbase = (uint64_t)(uintptr_t)A.tets;
        u4v rs;
        rs.x = (uint32_t)rbase;
        rs.y = (uint32_t)(rbase >> 32) & 0xFFFFu;
        rs.z = A.tets_bytes;
        rs.w = (uint32_t)kBufferWord3;
        const uint32_t a0 = (uint32_t)r0 + quarter, a1 = (uint32_t)r1 + quarter, a2 = (uint32_t)r2 + quarter,
                       a3 = (uint32_t)r3 + quarter, a4 = (uint32_t)own + 64u;
        static_assert(65 * 16 == 0x410 && 130 * 16 == 0x820 && 195 * 16 == 0xc30 && kStageLinks * 16 == 0x1040,
                      "stage offsets of the fetch block");
        uint64_t saved;
        // M0 is a reserved register: the compiler also sets it for LDS DMA loads it emits
        // itself, and this kernel variant emits none (a CPU test compiles the kernel and
        // checks that M0 is written only by these blocks).
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winline-asm"
        asm volatile(
            "s_mov_b64 %[sv], exec\n\t"
            "s_mov_b32 m0, %[sb]\n\t"
            "s_mov_b64 exec, %[k0]\n\t"
            "s_nop 0\n\t"
            "buffer_load_dwordx4 %[a0], %[rs], 0 offen lds\n\t"
            "s_add_u32 m0, %[sb], 0x410\n\t"
            "s_mov_b64 exec, %[k1]\n\t"
            "s_nop 0\n\t"
            "buffer_load_dwordx4 %[a1], %[rs], 0 offen lds\n\t"
            "s_add_u32 m0, %[sb], 0x820\n\t"
            "s_mov_b64 exec, %[k2]\n\t"
            "s_nop 0\n\t"
            "buffer_load_dwordx4 %[a2], %[rs], 0 offen lds\n\t"
            "s_add_u32 m0, %[sb], 0xc30\n\t"
            "s_mov_b64 exec, %[k3]\n\t"
            "s_nop 0\n\t"
            "buffer_load_dwordx4 %[a3], %[rs], 0 offen lds\n\t"
            "s_add_u32 m0, %[sb], 0x1040\n\t"
            "s_mov_b64 exec, %[k4]\n\t"
            "s_nop 0\n\t"
            "buffer_load_dwordx4 %[a4], %[rs], 0 offen lds\n\t"
            "s_mov_b64 exec, %[sv]"
            : [sv] "=&s"(saved)
            : [k0] "s"(k0), [k1] "s"(k1), [k2] "s"(k2), [k3] "s"(k3), [k4] "s"(k4), [a0] "v"(a0), [a1] "v"(a1),
              [a2] "v"(a2), [a3] "v"(a3), [a4] "v"(a4), [sb] "s"(sb), [rs] "s"(rs)
            : "memory", "m0", "scc");
#pragma clang diagnostic pop
        if (want) flags |= F_STAGED;
    };

    if (COOP && SLICES && A.restore) {
        // Restored photons, parked ones included, get their tets staged, and the stage is
        // complete before the loop reads it.
        coop_fetch((flags & F_ALIVE) != 0);
        __builtin_amdgcn_s_waitcnt(0x0F70);
        asm volatile("" ::: "memory");
    }

    for (;;) {
        // ---- refill dead lanes (wave-uniform decision)
        {
            const bool alive = (flags & F_ALIVE) != 0;
            const uint64_t dmask = __ballot(!alive);
            const uint32_t ndead = (uint32_t)__popcll(dmask);
            if (!drained && ndead != 0 && (ndead >= q || ndead == 64u)) {
                const int leader = __ffsll((unsigned long long)dmask) - 1;
                uint64_t base = 0;
                if ((int)lane == leader)
                    base = atomicAdd((unsigned long long *)&A.ctrl[CTRL_NEXT_ID], (unsigned long long)ndead);
                base = readlane_u64(base, leader);
                if (!alive) {
                    const uint32_t rank = (uint32_t)__popcll(dmask & ((1ull << lane) - 1ull));
                    const uint64_t nid = base + rank;
                    if (nid < A.id_end) {
                        // LaunchPhoton, src/octmps_kernel.cuh:475-503.
                        int lo = 0, hi = A.n_seg - 1;
                        while (lo < hi) {
                            int mid = (lo + hi + 1) >> 1;
                            if (A.seg_start[mid] <= nid) lo = mid;
                            else hi = mid - 1;
                        }
                        slot = A.seg_slot[lo];
                        const DevScan &sc = A.slots[slot];
                        const uint64_t pidx = sc.photon_offset + A.seg_poff[lo] + (nid - A.seg_start[lo]);
                        const uint64_t h = photon_hash(sc.seed, sc.salt, pidx);
                        const uint32_t k = (uint32_t)pidx & (A.n_triplets - 1u);
                        aG = A.mult[3 + 3 * k];
                        aR = A.mult[4 + 3 * k];
                        aS = A.mult[5 + 3 * k];
                        const uint32_t b0 = A.mult[0], b1 = A.mult[1], b2 = A.mult[2];
                        sG = derive_state(seeding_state(h, 0, b0), b0, aG);
                        sR = derive_state(seeding_state(h, 1, b1), b1, aR);
                        sS = derive_state(seeding_state(h, 2, b2), b2, aS);
                        x = sc.px; y = sc.py; z = sc.pz;
                        lds_probe_at() = make_float4(sc.px, sc.py, sc.pz, 0.0f);
                        ux = 0.0f; uy = 0.0f; uz = 1.0f;
                        w = sc.w0;
                        // The first step length (General stream) is drawn at launch; the
                        // reference draws it at the top of the first step, with no other
                        // General draw in between.
                        tau = neg_ln(rand_oc(sG, aG));
                        op = 0.0f; maxz = 0.0f;
                        lr = 1.0f; lr_fst = 1.0f; loc_fst = -1.0f;
                        tet = sc.root_tet;
                        region = sc.root_region;
                        set = MULTI ? sc.set : 0;
                        if (MULTI && SETS_LDS) set = A.set_slot[set];
                        flags = F_ALIVE;
                        steps = 0;
                        id = nid;
                        ndet = 0;
                        csum = 0.0f;
                        load_region(region);
                    }
                }
                if (base + ndead >= A.id_end) drained = true;
            }
        }
        const uint64_t amask = __ballot((flags & F_ALIVE) != 0);
        if (amask == 0) {
            if (drained) break;
            continue;
        }
        if (SLICES && A.slice_mode &&
            ((drained && A.exit_on_drain) || (A.iter_budget && iters >= A.iter_budget)))
            break;
        // Slices end on the device wall clock (checked every 32 iterations), so all waves of
        // the grid leave together instead of the slice lasting as long as its slowest wave.
        if (SLICES && A.slice_mode && A.slice_ticks && (iters & 31u) == 0 &&
            __builtin_amdgcn_s_memrealtime() - t_start > A.slice_ticks)
            break;
        ++iters;
        // The crossing phase handles the common step: a hop that ends inside the tet or on
        // a face into the same region. A lane whose step ended inside the tet (absorb +
        // scatter) parks with F_PEND, and one that reached a face into another region, the
        // mesh boundary, or a face at grazing incidence parks with F_IFACE (Fresnel,
        // reflection, detection), until enough lanes of the wave need the parked phase
        // (scatter quantum), while the other lanes keep crossing faces. Each photon still
        // runs its operations in the reference's order, so results do not depend on the
        // quantum.
        const bool ready = (flags & (F_ALIVE | F_PEND | F_IFACE)) == F_ALIVE;
        wave_steps += (uint64_t)__popcll(__ballot(ready));

        const DevSet &S = MULTI ? (SETS_LDS ? lds_sets[set] : A.sets[set]) : S0;
        bool finish = false, killed = false;

        // Roulette, src/octmps_kernel.cuh:1376-1392; returns true when the photon is done.
        auto roulette = [&]() -> bool {
            if (!(w < kWeightF)) return false;
            const float r = rand_co(sG, aG);
            if (w != 0.0f && r < kChanceF) {
                w *= 10.0f;
                return false;
            }
            if ((flags & F_FST) && S.type_bias != 3) {
                const float lrt = lr_fst;
                const float4 c0 = *reinterpret_cast<const float4 *>(&C.x);
                const float4 c1 = *reinterpret_cast<const float4 *>(&C.ux);
                const int4 c2 = *reinterpret_cast<const int4 *>(&C.maxz);
                x = c0.x; y = c0.y; z = c0.z; w = c0.w;
                ux = c1.x; uy = c1.y; uz = c1.z; op = c1.w;
                maxz = __int_as_float(c2.x);
                tet = c2.y;
                const int cregion = c2.z;
                tau = 0.0f;
                flags = F_ALIVE | ((flags & F_CNBS) ? F_NBS : 0u);
                loc_fst = -1.0f;
                if (cregion != region) {
                    region = cregion;
                    load_region(region);
                }
                const DevRegion &RC = region_rec(region);
                const float cost = hg_cos(RC.g, RC.hg_a, RC.hg_b, rand_oc(sS, aS));
                rotate(ux, uy, uz, cost, rand_co(sS, aS), 0, ux, uy, uz);
                lr = (lrt < 1.0f) ? 1.0f - lrt : 1.0f;
                return false;
            }
            return true;
        };

        if (COOP) {
            // Lanes about to cross whose stage does not hold their tet yet (most were
            // fetched ahead, at the end of the previous iteration).
            coop_fetch(ready && !(steps + 1u > max_steps) && !(flags & F_STAGED));
        }

        // Crossing phase: one step of every ready lane. The two record paths get different
        // control flow, each the faster one measured for it: with the cooperative fetch the
        // records were loaded ahead and the phase is bound by instruction issue, so it runs
        // as one divergent branch of selects; with per-lane gathers the phase issues its own
        // loads and waits for them, and the branching form below runs faster.
        if constexpr (COOP) {
            // Step guard: a photon past max_steps is stopped and counted.
            steps = ready ? steps + 1u : steps;
            killed = ready && steps > max_steps;
            finish = killed;
            // Lanes whose step ended in this phase (a face into the same region), for roulette.
            bool crossed = false;
            // One divergent branch: every outcome inside is a select, since a divergent branch
            // costs several scalar instructions and the loop is bound by instruction issue.
            if (ready && !killed) {
                // ---- ComputeStepSize, src/octmps_kernel.cuh:505-515. A fresh step length
                // (sleft == 0) is normally drawn ahead, where the previous step ended (parked
                // phase, launch), so this phase rarely runs the logarithm. +0 marks "draw now";
                // a drawn length of zero is -0 (-ln(1)) and is used as is.
                float tau_d = tau;
                if (__float_as_uint(tau_d) == 0u) tau_d = neg_ln(rand_oc(sG, aG));
                float s = tau_d * rmuas;
                // ---- HitBoundary, src/octmps_kernel.cuh:517-572. The tet's four planes and
                // links are in the lane's stage (fetched ahead); the stage loads write LDS when
                // they complete: wait for all of the wave's vector memory operations (s_waitcnt
                // vmcnt(0)) before reading it. The entry face is not needed here.
                flags &= ~F_ENTRY;
                __builtin_amdgcn_s_waitcnt(0x0F70);
                asm volatile("" ::: "memory");
                // From here until its next fetch is issued the wave has the stage's data, and
                // runs ahead of waves that wait for theirs (issue priority 1, dropped to 0 after
                // the end-of-iteration fetch): the next fetches go out sooner.
                __builtin_amdgcn_s_setprio(1);
                const uint4 L = *own_links;
                const float4 p0 = own_planes[0], p1 = own_planes[1], p2 = own_planes[2], p3 = own_planes[3];
                __builtin_amdgcn_sched_barrier(0);
                const float c0 = fmaf(p0.z, uz, fmaf(p0.y, uy, p0.x * ux));
                const float c1 = fmaf(p1.z, uz, fmaf(p1.y, uy, p1.x * ux));
                const float c2 = fmaf(p2.z, uz, fmaf(p2.y, uy, p2.x * ux));
                const float c3 = fmaf(p3.z, uz, fmaf(p3.y, uy, p3.x * ux));
                float dmin = 1e10f, mincos = 0.0f;
                int imin = -1;
                // Faces in order with a strict '<', as the reference; written with selects so
                // every plane is loaded whole and the wave does not branch per face.
                auto face = [&](int I, const float4 &PL, float C) {
                    const float perp = fmaf(PL.z, z, fmaf(PL.y, y, fmaf(PL.x, x, PL.w)));
                    const float d = -perp * rcp(C);
                    const bool take = (C < 0.0f) & (d < dmin);
                    dmin = take ? d : dmin;
                    imin = take ? I : imin;
                    mincos = take ? C : mincos;
                };
                face(0, p0, c0);
                face(1, p1, c1);
                face(2, p2, c2);
                face(3, p3, c3);
                const bool hit = imin >= 0 && s > dmin;
                // The leftover of a step that reaches a face, in optical depth. Glass (mus = 0)
                // keeps the leftover unchanged, as (s - d) / DBL_MAX does in the reference; an
                // overflowed step stays infinite, and a zero leftover is stored as +0 so that the
                // next step draws.
                const float left_m = (s - dmin) * muas;
                const float left_g = __builtin_isinf(s) ? INFINITY : tau_d + 0.0f;
                tau = hit ? (muas == 0.0f ? left_g : left_m) : 0.0f;
                s = hit ? dmin : s;
                // ---- Hop, src/octmps_kernel.cuh:574-582
                x = fmaf(s, ux, x);
                y = fmaf(s, uy, y);
                z = fmaf(s, uz, z);
                op += s;
                maxz = (maxz < z) ? z : maxz;
                // Packed neighbour word of the exit face: pool tet id (all ones at the mesh
                // boundary), the shared face's index within the neighbour, and the neighbour's
                // region. Selected on the bits of imin (-1 gives L.w, unused: no hit).
                const bool i0 = (imin & 1) != 0, i1 = (imin & 2) != 0;
                const uint32_t pk_lo = i0 ? L.y : L.x, pk_hi = i0 ? L.w : L.z;
                const uint32_t pk = i1 ? pk_hi : pk_lo;
                // A face into the same region (the region field equals the photon's and is not
                // the escape, which the mesh boundary's all-ones word also carries) away from
                // grazing incidence: FastReflectTransmit (src/octmps_kernel.cuh:590-805) with
                // equal indices gives R = 0 (sam == 0), so the photon transmits unless the
                // ReflectT draw is 0, which it is exactly when the new MWC word is 0.
                const uint32_t field = pk >> kRegionShift;
                const bool same = hit & (field == (uint32_t)region) & (field != kRegionEscape) & (-mincos >= kCos90F);
                const uint64_t sRn = (sR & 0xffffffffull) * (uint64_t)aR + (sR >> 32);
                sR = same ? sRn : sR;
                const bool xmit = same & ((uint32_t)sRn != 0u);
                tet = xmit ? (int)(pk & kLinkMask) : tet;
                flags = xmit ? ((flags & ~F_STAGED) | ((((pk >> kBackShift) & 3u) + 1u) << kEntryShift)) : flags;
                if (same & ((uint32_t)sRn == 0u)) {
                    const float4 pf_lo = i0 ? p1 : p0, pf_hi = i0 ? p3 : p2;
                    const float4 pf = i1 ? pf_hi : pf_lo;
                    const float two_ca1 = -2.0f * mincos;
                    if (uz > 0.0f) flags |= F_NBS;
                    ux = fmaf(two_ca1, pf.x, ux);
                    uy = fmaf(two_ca1, pf.y, uy);
                    uz = fmaf(two_ca1, pf.z, uz);
                }
                // Any other face (another region, the mesh boundary, grazing incidence) parks the
                // lane for the parked phase with its face index (the face's normal and neighbour
                // word stay in the lane's stage). A step that ends inside the tet parks the lane
                // to scatter.
                const bool iface = hit & !same;
                flags |= iface ? (F_IFACE | ((uint32_t)imin << kFaceShift)) : (hit ? 0u : F_PEND);
                crossed = same;
            }
            // Roulette after a step completed in the crossing phase.
            if (crossed && w < kWeightF) finish = roulette();
        } else {
            if (!ready) {
                // parked or dead lane: nothing to do in the crossing phase
            } else if (++steps > max_steps) {
                finish = true;
                killed = true;
            } else {
                // ---- ComputeStepSize, src/octmps_kernel.cuh:505-515. A fresh step length
                // (sleft == 0) is normally drawn ahead, where the previous step ended (parked
                // phase, launch), so this phase rarely runs the logarithm. +0 marks "draw now";
                // a drawn length of zero is -0 (-ln(1)) and is used as is.
                float tau_d = tau;
                if (__float_as_uint(tau_d) == 0u) tau_d = neg_ln(rand_oc(sG, aG));
                float s = tau_d * rmuas;
                tau = 0.0f;
                // ---- HitBoundary, src/octmps_kernel.cuh:517-572
                // The tet record is read through a buffer resource (32-bit offsets). With
                // SKIP, a photon that crossed into this tet through face k does not fetch
                // plane k: that plane is the exact negation of the one it crossed, so its cosine
                // is > 0 and the face is never a candidate. Its load is sent out of range instead,
                // which costs no memory request and returns a zero plane (cosine 0, also never a
                // candidate).
                const uint32_t entry = (flags & F_ENTRY) >> kEntryShift;
                flags &= ~F_ENTRY;
                const uint32_t rec = (uint32_t)tet << 7;
                auto plane_at = [&](uint32_t k) -> float4 {
                    const u4v v = __builtin_amdgcn_raw_buffer_load_b128(
                        tets_rsrc, (SKIP && entry == k + 1u) ? kOutOfRange : rec, (int)(16u * k), 0);
                    return make_float4(__uint_as_float(v.x), __uint_as_float(v.y), __uint_as_float(v.z),
                                       __uint_as_float(v.w));
                };
                float4 p0, p1, p2, p3;
                uint4 L;
                p0 = plane_at(0);
                p1 = plane_at(1);
                p2 = plane_at(2);
                p3 = plane_at(3);
                {
                    const u4v lv = __builtin_amdgcn_raw_buffer_load_b128(tets_rsrc, rec + 64u, 0, 0);
                    L = make_uint4(lv.x, lv.y, lv.z, lv.w);
                }
                // Keep the loads together ahead of the face tests (one memory round trip).
                __builtin_amdgcn_sched_barrier(0);
                const float c0 = fmaf(p0.z, uz, fmaf(p0.y, uy, p0.x * ux));
                const float c1 = fmaf(p1.z, uz, fmaf(p1.y, uy, p1.x * ux));
                const float c2 = fmaf(p2.z, uz, fmaf(p2.y, uy, p2.x * ux));
                const float c3 = fmaf(p3.z, uz, fmaf(p3.y, uy, p3.x * ux));
                float dmin = 1e10f, mincos = 0.0f;
                int imin = -1;
                // Faces in order with a strict '<', as the reference; written with selects so
                // every plane is loaded whole and the wave does not branch per face.
                auto face = [&](int I, const float4 &PL, float C) {
                    const float perp = fmaf(PL.z, z, fmaf(PL.y, y, fmaf(PL.x, x, PL.w)));
                    const float d = -perp * rcp(C);
                    const bool take = (C < 0.0f) & (d < dmin);
                    dmin = take ? d : dmin;
                    imin = take ? I : imin;
                    mincos = take ? C : mincos;
                };
                face(0, p0, c0);
                face(1, p1, c1);
                face(2, p2, c2);
                face(3, p3, c3);
                const bool hit = imin >= 0 && s > dmin;
                if (hit) {
                    // Glass (mus = 0) keeps the leftover unchanged, as (s - d) / DBL_MAX does
                    // in the reference; an overflowed step stays infinite, and a zero leftover
                    // is stored as +0 so that the next step draws.
                    tau = (muas == 0.0f) ? (isinf(s) ? INFINITY : tau_d + 0.0f) : (s - dmin) * muas;
                    s = dmin;
                }
                // ---- Hop, src/octmps_kernel.cuh:574-582
                x = fmaf(s, ux, x);
                y = fmaf(s, uy, y);
                z = fmaf(s, uz, z);
                op += s;
                if (maxz < z) maxz = z;

                if (hit) {
                    // Packed neighbour word: pool tet id (all ones at the mesh boundary), the
                    // shared face's index within the neighbour, and the neighbour's region.
                    const uint32_t pk = imin == 0 ? L.x : (imin == 1 ? L.y : (imin == 2 ? L.z : L.w));
                    // Same region: the region field equals the photon's and is not the escape
                    // (which the mesh boundary's all-ones word also carries).
                    const uint32_t field = pk >> kRegionShift;
                    if (field == (uint32_t)region && field != kRegionEscape && -mincos >= kCos90F) {
                        // FastReflectTransmit (src/octmps_kernel.cuh:590-805) on a face into the
                        // same region away from grazing incidence: equal indices give R = 0
                        // (sam == 0), so the photon transmits unless the ReflectT draw is 0,
                        // which it is exactly when the new MWC word is 0.
                        if (mwc_next(sR, aR) != 0u) {
                            tet = (int)(pk & kLinkMask);
                            flags = (flags & ~F_STAGED) | ((((pk >> kBackShift) & 3u) + 1u) << kEntryShift);
                        } else {
                            const float4 pf = imin == 0 ? p0 : (imin == 1 ? p1 : (imin == 2 ? p2 : p3));
                            const float two_ca1 = -2.0f * mincos;
                            if (uz > 0.0f) flags |= F_NBS;
                            ux = fmaf(two_ca1, pf.x, ux);
                            uy = fmaf(two_ca1, pf.y, uy);
                            uz = fmaf(two_ca1, pf.z, uz);
                        }
                    } else {
                        const float4 pf = imin == 0 ? p0 : (imin == 1 ? p1 : (imin == 2 ? p2 : p3));
                        lds_face[threadIdx.x] = make_float4(pf.x, pf.y, pf.z, __uint_as_float(pk));
                        flags |= F_IFACE | ((uint32_t)imin << kFaceShift);
                    }
                } else {
                    flags |= F_PEND;
                }
            }
            // Roulette after a step completed in the crossing phase.
            if (ready && !finish && !(flags & (F_PEND | F_IFACE))) finish = roulette();
        }

        // ---- Parked phase: interface events (Fresnel, reflection or transmission,
        // detection), or Drop + Spin/SpinBias, then roulette, for the parked lanes.
        const uint64_t pmask = __ballot((flags & (F_PEND | F_IFACE)) != 0);
        if (pmask != 0 && ((uint32_t)__popcll(pmask) >= A.scatter_q ||
                           __ballot((flags & (F_ALIVE | F_PEND | F_IFACE)) == F_ALIVE) == 0)) {
            const bool parked = (flags & (F_PEND | F_IFACE)) != 0;
            if (flags & F_IFACE) {
                // ---- FastReflectTransmit, src/octmps_kernel.cuh:590-805
                const uint32_t fidx = (flags & F_FACE) >> kFaceShift;
                flags &= ~(F_IFACE | F_FACE);
                // The exit face's normal and packed neighbour word (pool tet id, all ones at
                // the mesh boundary, and its region): in the stage (COOP), or stashed when the
                // lane parked.
                float3 pf;
                uint32_t pk;
                if (COOP) {
                    const float4 fr = own_planes[fidx];
                    const uint4 lk = *own_links;
                    pf = make_float3(fr.x, fr.y, fr.z);
                    pk = fidx == 0 ? lk.x : (fidx == 1 ? lk.y : (fidx == 2 ? lk.z : lk.w));
                } else {
                    const float4 fr = lds_face[threadIdx.x];
                    pf = make_float3(fr.x, fr.y, fr.z);
                    pk = __float_as_uint(fr.w);
                }
                const int nxt = (pk & kLinkMask) == kLinkMask ? -1 : (int)(pk & kLinkMask);
                const uint32_t field = pk >> kRegionShift;
                // An escaped link carries no region: read it from the neighbour's record.
                const int nr = (nxt < 0 || field != kRegionEscape) ? (int)field : A.tets[nxt].region;
                // The same operations as the crossing phase's cosine of this face.
                const float ca1 = -fmaf(pf.z, uz, fmaf(pf.y, uy, pf.x * ux));
                // The empty asm keeps rn a register value: otherwise the three candidate
                // loads are merged into one load through a phi of pointers, which puts rn
                // in scratch memory and turns the load into a flat load.
                float nt = rn;
                asm("" : "+v"(nt));
                if (nxt < 0) nt = S.n_ambient;
                else if (nr != region) nt = region_rec(nr).n;
                const float rnd = rand_co(sR, aR);
                float R;
                if (nxt >= 0 && nt == rn) R = (ca1 < kCos90F) ? 1.0f : 0.0f;  // sam == 0
                else R = fresnel_f(rn, nt, ca1);
                const float two_ca1 = 2.0f * ca1;
                const float rx = fmaf(two_ca1, pf.x, ux);
                const float ry = fmaf(two_ca1, pf.y, uy);
                const float rz = fmaf(two_ca1, pf.z, uz);
                if (nxt < 0 && R < 1.0f) {
                    if (uz < 0.0f) {
                        bool flag = false;
                        const float4 probe = lds_probe_at();
                        const float dx = x - probe.x, dy = y - probe.y;
                        if (fmaf(dx, dx, dy * dy) < S.radius2 && -uz > S.cos_angle) {
                            const int tb = S.type_bias;
                            if (tb == 0 || tb == 3) flag = true;
                            if (((flags & F_FST) && tb == 37) || (loc_fst == maxz && tb != 37) ||
                                (flags & F_NBS))
                                flag = true;
                        }
                        if (flag) {
                            const float c = w * (1.0f - R) * lr;
                            tally_detection(A, S, slot, op, maxz, c);
                            if (LOG) {
                                ++ndet;
                                csum += c;
                            }
                        }
                    } else {
                        flags |= F_NBS;
                    }
                    w *= R;
                    ux = rx;
                    uy = ry;
                    uz = rz;
                } else if (rnd > R) {
                    tet = nxt;
                    flags &= ~F_STAGED;
                    if (field != kRegionEscape) flags |= (((pk >> kBackShift) & 3u) + 1u) << kEntryShift;
                    if (nr != region) {
                        region = nr;
                        load_region(region);
                    }
                } else {
                    if (uz > 0.0f) flags |= F_NBS;
                    ux = rx;
                    uy = ry;
                    uz = rz;
                }
            } else if (flags & F_PEND) {
                flags &= ~F_PEND;
                // ---- Drop, src/octmps_kernel.cuh:584-588
                const DevRegion RG = region_rec(region);
                w = fmaf(-w, RG.mua_muas, w);
                const int tb = S.type_bias;
                if (tb == 0) {
                    // Spin, src/octmps_kernel.cuh:822-879
                    const float cost = hg_cos(RG.g, RG.hg_a, RG.hg_b, rand_oc(sS, aS));
                    rotate(ux, uy, uz, cost, rand_co(sS, aS), 0, ux, uy, uz);
                } else if (tb == 37) {
                    // SpinBias, src/octmps_kernel.cuh:898-1086, written with selects: the
                    // branch states (first bias, F = 1 or 2, selector) only choose values, and
                    // the Spin stream's draws are taken in the reference's order (the selector
                    // draw only when F != 0, then the cost draw, then psi).
                    const float ux0 = ux, uy0 = uy, uz0 = uz;
                    // First backward bias: the continuation photon is copied before the flag
                    // is set (a store, so this stays a branch; it happens once per photon).
                    const bool first = (z > S.td_min) & (z < S.td_max) & (uz > 0.0f) & ((flags & F_FST) == 0);
                    if (first) {
                        *reinterpret_cast<float4 *>(&C.x) = make_float4(x, y, z, w);
                        *reinterpret_cast<float4 *>(&C.ux) = make_float4(ux, uy, uz, op);
                        *reinterpret_cast<int4 *>(&C.maxz) = make_int4(__float_as_int(maxz), tet, region, 0);
                    }
                    flags = first ? ((flags & ~F_CNBS) | ((flags & F_NBS) ? F_CNBS : 0u) | F_FST) : flags;
                    loc_fst = first ? z : loc_fst;
                    const float4 probe = lds_probe_at();
                    float vx = probe.x - x, vy = probe.y - y, vz = probe.z - z;
                    const float len2 = fmaf(vx, vx, fmaf(vy, vy, vz * vz));
                    const float rl = __builtin_amdgcn_rsqf(len2);
                    vx *= rl;
                    vy *= rl;
                    vz *= rl;
                    // A photon biased before, or reflected backwards: F = 2, or 1 when its next
                    // step could reach the probe inside the acceptance cone.
                    const bool later = !first & ((flags & (F_FST | F_NBS)) != 0);
                    const float next = S.next_step_num * rcp(muas);
                    const float dist = len2 * rl;  // |x - probe|
                    const int F = first ? 1 : (later ? (((next >= dist) & (-vz >= S.cos_angle)) ? 1 : 2) : 0);
                    const float atmp = (first | later) ? S.bias_a : 0.0f;
                    const uint64_t sS1 = (sS & 0xffffffffull) * (uint64_t)aS + (sS >> 32);
                    const float u_sel = unit_rz24((uint32_t)sS1);
                    sS = (F != 0) ? sS1 : sS;
                    const int sel = (F != 0) ? ((u_sel <= S.p_add) ? 1 : 2) : 0;
                    const bool biased = (F == 1) | (sel == 1);
                    const float u_c = rand_oc(sS, aS);
                    const float cb = bias_cos(atmp, S.inv_1ma, S.inv_sq, S.inv_2a, S.opa2, u_c);
                    const float ch = hg_cos(RG.g, RG.hg_a, RG.hg_b, u_c);
                    const float cost = fminf(fmaxf(biased ? cb : ch, -1.0f), 1.0f);
                    const float fx = biased ? vx : ux, fy = biased ? vy : uy, fz = biased ? vz : uz;
                    rotate(fx, fy, fz, cost, rand_co(sS, aS), 1, ux, uy, uz);
                    const float costg = fmaf(ux0, ux, fmaf(uy0, uy, uz0 * uz));
                    const float g = RG.g;
                    // Equation (8), src/octmps_kernel.cuh:1047-1050.
                    const float e1 = fmaf(-2.0f * atmp, cost, S.opa2);
                    const float e2 = fmaf(-2.0f * g, costg, fmaf(g, g, 1.0f));
                    const float factor8 = RG.lr8 * (e1 * fsqrt(e1)) * rcp(e2 * fsqrt(e2));
                    // Equation (9), src/octmps_kernel.cuh:1071-1078.
                    const float costg1 = fmaf(vx, ux, fmaf(vy, uy, vz * uz));
                    const float k1 = sel == 1 ? cost : costg1;
                    const float k2 = sel == 1 ? costg : cost;
                    const float t1 = fmaf(-2.0f * atmp, k1, S.opa2);
                    const float pdf1 = S.pdfb_k * rcp(t1 * fsqrt(t1));
                    const float t2 = fsqrt(fmaf(-2.0f * g, k2, fmaf(g, g, 1.0f)));
                    const float pdf2 = RG.hg_a * rcp(2.0f * (t2 * t2 * t2));
                    const float factor9 = pdf2 * rcp(fmaf(S.p_add, pdf1, (1.0f - S.p_add) * pdf2));
                    // The likelihood ratio changes only for a biased state (atmp != 0).
                    const bool apply = atmp != 0.0f;
                    lr = apply ? lr * (F == 1 ? factor8 : factor9) : lr;
                    lr_fst = (apply & first) ? lr : lr_fst;
                }
            }
            // Roulette (src/octmps_kernel.cuh:1376-1392) after the parked lanes' events.
            if (parked && w < kWeightF) finish = roulette();
            // ComputeStepSize's draw for the next step of a scattered or restored photon,
            // made here: its next General draw is that one.
            if (parked && !finish && __float_as_uint(tau) == 0u) tau = neg_ln(rand_oc(sG, aG));
        }
        if (finish) {
            flags = 0;
            atomicAdd((unsigned long long *)&A.slot_done[slot], 1ull);
            if (killed) atomicAdd((unsigned long long *)&A.ctrl[CTRL_KILLED], 1ull);
            if (LOG) {
                PhotonLogRec r;
                r.steps = steps;
                r.detections = ndet;
                r.contrib = csum;
                r.killed = killed ? 1u : 0u;
                A.plog[id] = r;
            }
        }
        // COOP: a photon that moved into another tet (crossing a face, or an interface in the
        // parked phase, or a restored continuation photon) and keeps crossing fetches its tet
        // here, ahead of the next crossing phase. Fetching at the end of the crossing phase
        // instead made the parked phase wait for those loads (its reads of the stage, and of
        // spilled registers, wait for all of the wave's vector memory operations).
        if (COOP) {
            coop_fetch((flags & (F_ALIVE | F_PEND | F_IFACE | F_STAGED)) == F_ALIVE);
            __builtin_amdgcn_s_setprio(0);
        }
    }

    // A slice can end with stage loads of the last iteration in flight: they complete into
    // this block's LDS before the wave ends.
    if (COOP) {
        __builtin_amdgcn_s_waitcnt(0x0F70);
        asm volatile("" ::: "memory");
    }
    if (SLICES && A.slice_mode) {
        SavedLane &s = A.save[gl];
        s.flags = flags;
        if (flags & F_ALIVE) {
            s.x = x; s.y = y; s.z = z; s.ux = ux; s.uy = uy; s.uz = uz; s.w = w; s.tau = tau;
            s.op = op; s.maxz = maxz; s.lr = lr; s.lr_fst = lr_fst; s.loc_fst = loc_fst;
            s.tet = tet; s.region = region; s.steps = steps;
            s.aG = aG; s.aR = aR; s.aS = aS; s.sG = sG; s.sR = sR; s.sS = sS;
            s.slot = slot; s.id = id; s.ndet = ndet; s.csum = csum;
        }
    }
    if (lane == 0 && wave_steps) atomicAdd((unsigned long long *)&A.ctrl[CTRL_STEPS], (unsigned long long)wave_steps);
    if (lane == 0 && iters) atomicAdd((unsigned long long *)&A.ctrl[CTRL_ITERS], (unsigned long long)iters);
}

__global__ void fill_u64_kernel(uint64_t *p, uint64_t n, uint64_t v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void clear_saved_kernel(SavedLane *p, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i].flags = 0;
}

__global__ void probe_rng_kernel(const uint32_t *mult, uint32_t K, uint64_t seed, uint64_t salt,
                                 uint64_t photon0, int n_photons, int n_draws, float *out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_photons) return;
    uint64_t pidx = photon0 + (uint64_t)i;
    uint64_t h = photon_hash(seed, salt, pidx);
    uint32_t k = (uint32_t)pidx & (K - 1u);
    for (int s = 0; s < 3; ++s) {
        uint32_t b = mult[s], a = mult[3 + 3 * k + s];
        uint64_t st = derive_state(seeding_state(h, (uint32_t)s, b), b, a);
        for (int d = 0; d < n_draws; ++d) out[((size_t)i * 3 + s) * n_draws + d] = rand_co(st, a);
    }
}

// One detection per lane into scan slot 0's tally (the detector tally on chosen inputs).
__global__ void probe_tally_kernel(KArgs A, DevSet S, int n, const float *op, const float *maxz, const float *c) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) tally_detection(A, S, 0, op[i], maxz[i], c[i]);
}

__global__ void probe_fresnel_kernel(int n, const float *ni, const float *nt, const float *ca1, float *r) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) r[i] = fresnel_f(ni[i], nt[i], ca1[i]);
}

__global__ void probe_sampler_kernel(int kind, float p, int n, const float *u, float *out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (kind == 0) {
        out[i] = hg_cos(p, 1.0f - p * p, p != 0.0f ? 1.0f / (2.0f * p) : 0.0f, u[i]);
    } else {
        float a = p;
        float sq = sqrtf(1.0f + a * a);
        out[i] = bias_cos(a, a != 1.0f ? 1.0f / (1.0f - a) : 0.0f, 1.0f / sq, a != 0.0f ? 1.0f / (2.0f * a) : 0.0f,
                          1.0f + a * a, u[i]);
    }
}

}  // namespace

namespace {
// The single-set variants come with and without the entry skip; the multi-set (serving)
// variant never skips (its register budget is full: with the skip it spills 14 registers).
// Dynamic LDS of the cooperative-fetch variant: a probe per lane, then the region table.
size_t coop_lds_bytes(int block, uint32_t n_regions) {
    return (size_t)block * sizeof(float4) + (size_t)n_regions * sizeof(DevRegion);
}

template <bool LOG>
void launch_single(const KArgs &a, int grid, int block, hipStream_t stream) {
    if (a.coop_fetch)
        hipLaunchKernelGGL((transport_kernel<false, LOG, false, false, true>), dim3(grid), dim3(block),
                           coop_lds_bytes(block, a.n_regions), stream, a);
    else if (a.skip_entry) hipLaunchKernelGGL((transport_kernel<false, LOG, true, false>), dim3(grid), dim3(block), 0, stream, a);
    else hipLaunchKernelGGL((transport_kernel<false, LOG, false, false>), dim3(grid), dim3(block), 0, stream, a);
}

// Multi-set variant with the parameter sets in LDS when the launch allows it, and with the
// cooperative record fetch when the launch asks for it (sets in LDS: their regions fit the
// smaller table beside the stage).
template <bool LOG>
void launch_multi(const KArgs &a, int grid, int block, hipStream_t stream) {
    if (a.sets_in_lds && a.coop_fetch)
        hipLaunchKernelGGL((transport_kernel<true, LOG, false, true, true>), dim3(grid), dim3(block), 0, stream, a);
    else if (a.coop_fetch)
        hipLaunchKernelGGL((transport_kernel<true, LOG, false, false, true>), dim3(grid), dim3(block), 0, stream, a);
    else if (a.sets_in_lds) hipLaunchKernelGGL((transport_kernel<true, LOG, false, true>), dim3(grid), dim3(block), 0, stream, a);
    else hipLaunchKernelGGL((transport_kernel<true, LOG, false, false>), dim3(grid), dim3(block), 0, stream, a);
}

// Blocks per CU of the tightest multi-set variant: a server's slices may switch variants, and
// its saved lanes need the same grid in every slice.
template <bool LOG>
int occupancy_multi(int block) {
    int a = 0, b = 0, c = 0, d = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, transport_kernel<true, LOG, false, true>, block, 0) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, transport_kernel<true, LOG, false, false>, block, 0) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&c, transport_kernel<true, LOG, false, true, true>, block, 0) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&d, transport_kernel<true, LOG, false, false, true>, block, 0) != hipSuccess)
        return 0;
    return std::min(std::min(a, b), std::min(c, d));
}

// Blocks per CU of the tighter of the two single-set variants, or of the cooperative one
// (coop_regions > 0: its LDS holds that many regions).
template <bool LOG>
int occupancy_single(int block, uint32_t coop_regions) {
    int a = 0, b = 0;
    if (coop_regions)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, transport_kernel<false, LOG, false, false, true>, block,
                                                            coop_lds_bytes(block, coop_regions)) == hipSuccess ? a : 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, transport_kernel<false, LOG, true, false>, block, 0) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, transport_kernel<false, LOG, false, false>, block, 0) != hipSuccess)
        return 0;
    return a < b ? a : b;
}
}  // namespace

hipError_t launch_transport(const KArgs &a, int grid, int block, bool multi, bool log, hipStream_t stream) {
    if (multi) {
        if (log) launch_multi<true>(a, grid, block, stream);
        else launch_multi<false>(a, grid, block, stream);
    } else {
        if (log) launch_single<true>(a, grid, block, stream);
        else launch_single<false>(a, grid, block, stream);
    }
    return hipGetLastError();
}

int transport_blocks_per_cu(int block, bool multi, bool log, uint32_t coop_regions) {
    if (multi) return log ? occupancy_multi<true>(block) : occupancy_multi<false>(block);
    return log ? occupancy_single<true>(block, coop_regions) : occupancy_single<false>(block, coop_regions);
}

hipError_t launch_fill_u64(uint64_t *p, uint64_t n, uint64_t v, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    uint64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(fill_u64_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, p, n, v);
    return hipGetLastError();
}

__global__ void fold_tallies_kernel(uint64_t *p, uint64_t n_words, uint64_t words, uint64_t bins,
                                    uint32_t copies, uint64_t stride) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_words;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t q = i % words;  // word within one A-scan's tally
        const uint64_t r = q >= 4 ? (q - 4) % (OCTMPS_TALLY_CLASS_WORDS * bins) : 0;
        const bool is_max = q >= 4 && r >= (OCTMPS_TALLY_CLASS_WORDS - 1) * bins;
        uint64_t acc = p[i];
        for (uint32_t c = 1; c < copies; ++c) {
            const uint64_t v = p[i + c * stride];
            acc = is_max ? (v > acc ? v : acc) : acc + v;  // Max words: non-negative IEEE bits
        }
        p[i] = acc;
    }
}

hipError_t launch_fold_tallies(uint64_t *p, uint64_t n_scans, int bins, uint32_t copies, uint64_t stride,
                               hipStream_t stream) {
    if (copies <= 1 || n_scans == 0) return hipSuccess;
    const uint64_t words = 4 + 2 * OCTMPS_TALLY_CLASS_WORDS * (uint64_t)bins;  // octmps_tally_words
    const uint64_t n = n_scans * words;
    uint64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(fold_tallies_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, p, n, words,
                       (uint64_t)bins, copies, stride);
    return hipGetLastError();
}

hipError_t launch_clear_saved(SavedLane *p, uint64_t n, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    uint64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(clear_saved_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, p, n);
    return hipGetLastError();
}

hipError_t launch_probe_rng(const uint32_t *mult, uint32_t K, uint64_t seed, uint64_t salt, uint64_t photon0,
                            int n_photons, int n_draws, float *out, hipStream_t stream) {
    hipLaunchKernelGGL(probe_rng_kernel, dim3((n_photons + 63) / 64), dim3(64), 0, stream, mult, K, seed, salt,
                       photon0, n_photons, n_draws, out);
    return hipGetLastError();
}

hipError_t launch_probe_fresnel(int n, const float *ni, const float *nt, const float *ca1, float *r,
                                hipStream_t stream) {
    hipLaunchKernelGGL(probe_fresnel_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, n, ni, nt, ca1, r);
    return hipGetLastError();
}

hipError_t launch_probe_tally(const KArgs &a, const DevSet &S, int n, const float *op, const float *maxz,
                              const float *c, hipStream_t stream) {
    hipLaunchKernelGGL(probe_tally_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, a, S, n, op, maxz, c);
    return hipGetLastError();
}

hipError_t launch_probe_sampler(int kind, float p, int n, const float *u, float *out, hipStream_t stream) {
    hipLaunchKernelGGL(probe_sampler_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, kind, p, n, u, out);
    return hipGetLastError();
}

}  // namespace octmps

namespace octmps {
namespace {
__global__ void copy_kernel(const uint4 *__restrict__ src, uint4 *__restrict__ dst, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}
}  // namespace

namespace {
// Divergent 16-byte gathers from 128-byte records that stay in cache: each lane visits
// pseudo-random records of a small set and reads 5 x 16 B of each, as a photon-step reads
// its tet record. Every offset is (index mod n) * 128 + 16 * j with j < 5, inside the
// n * 128 bytes of the buffer.
__global__ void __launch_bounds__(256) gather_probe_kernel(const TetRec *recs, uint32_t n, int iters,
                                                           uint32_t *out) {
    const __amdgpu_buffer_rsrc_t r =
        __builtin_amdgcn_make_buffer_rsrc((void *)recs, (short)0, (int)(n * 128u), kBufferWord3);
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t idx = g * 2654435761u, acc = 0;
    for (int i = 0; i < iters; ++i) {
        idx = idx * 1664525u + 1013904223u;
        const uint32_t rec = (idx % n) * 128u;
        const u4v a = __builtin_amdgcn_raw_buffer_load_b128(r, rec, 0, 0);
        const u4v b = __builtin_amdgcn_raw_buffer_load_b128(r, rec, 16, 0);
        const u4v c = __builtin_amdgcn_raw_buffer_load_b128(r, rec, 32, 0);
        const u4v d = __builtin_amdgcn_raw_buffer_load_b128(r, rec, 48, 0);
        const u4v e = __builtin_amdgcn_raw_buffer_load_b128(r, rec, 64, 0);
        acc += a.x ^ b.y ^ c.z ^ d.w ^ e.x;
    }
    out[g] = acc;
}

// L1 cost of one 16-byte vector load per access pattern: blocks x 256 lanes issue iters x 8
// loads each from a 16 KiB buffer (128 lines of 128 B, L1-resident). Patterns: 0 divergent
// (the 64 lanes of a load hit 64 lines), 1 quad-coalesced (lanes 4q..4q+3 read the four
// 16-byte quarters of one 64-byte piece, 16 lines per load), 2 divergent with every other
// lane masked off, 3 divergent with every other lane out of range, 4 quad-coalesced DMA into
// LDS (buffer_load ... lds, the cooperative record fetch's form). Every offset is below
// 16 KiB except pattern 3's out-of-range ones (past the buffer resource's size).
template <int P>
__global__ void __launch_bounds__(256) l1_probe_kernel(const uint8_t *buf, int iters, float *out) {
    __shared__ float4 dma_dst[4 * 128];
    const uint32_t l = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc((void *)buf, (short)0, 16384, kBufferWord3);
    float acc = 0.0f;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (uint32_t u = 0; u < 8; ++u) {
            uint32_t off;
            if (P == 1 || P == 4) off = (((l >> 2) * 9u + u * 5u + (uint32_t)it) & 127u) * 128u + ((u >> 2) & 1u) * 64u + (l & 3u) * 16u;
            else off = (((l * 5u + u * 17u + (uint32_t)it * 3u) & 127u) * 128u) + (u & 3u) * 16u;
            if (P == 3 && (l & 1u)) off = kOutOfRange;
            if (P == 4) {
                typedef __attribute__((address_space(3))) void *lds_ptr;
                __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (lds_ptr)(dma_dst + wave * 128u + (u & 1u) * 64u), 16, off, 0, 0, 0);
            } else if (P != 2 || (l & 1u) == 0) {
                const u4v v = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
                acc += __uint_as_float(v.x) + __uint_as_float(v.y) + __uint_as_float(v.z) + __uint_as_float(v.w);
            }
        }
    }
    if (P == 4) {
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
        asm volatile("" ::: "memory");
        acc += dma_dst[wave * 128u + l].x;
    }
    out[blockIdx.x * 256u + threadIdx.x] = acc;
}

// VALU issue rate by instruction class: every lane runs iters x 8 independent operations on
// eight accumulators (no dependency between consecutive instructions of a wave), at 6 waves
// per SIMD. Class 0: v_fma_f32 (the common FP32 operation); 1: v_rcp_f32 (the transcendental
// class: reciprocal, square root, log, sin and cos); 2: one MWC step (v_mad_u64_u32 and the
// move of the previous high word into its addend, the two instructions the transport kernel
// spends per draw).
template <int K>
__global__ void __launch_bounds__(256) valu_probe_kernel(int iters, float *out) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    float a[8];
    uint64_t b[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        a[k] = 1.0f + 1e-3f * (float)((g + 7u * k) & 255u);
        b[k] = ((uint64_t)(g + k) << 32) | (0x9E3779B9u ^ (uint32_t)k);
    }
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if (K == 0) a[k] = fmaf(a[k], 0.999f, 1e-3f);
            if (K == 1) a[k] = rcp(a[k]);
            if (K == 2) b[k] = (b[k] & 0xffffffffull) * 4294967118ull + (b[k] >> 32);
        }
    }
    float acc = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; ++k) acc += a[k] + (float)(uint32_t)b[k];
    out[g] = acc;
}
}  // namespace

hipError_t launch_valu_probe(int kind, int iters, float *out, int blocks, hipStream_t stream) {
    switch (kind) {
    case 0: hipLaunchKernelGGL(valu_probe_kernel<0>, dim3(blocks), dim3(256), 0, stream, iters, out); break;
    case 1: hipLaunchKernelGGL(valu_probe_kernel<1>, dim3(blocks), dim3(256), 0, stream, iters, out); break;
    case 2: hipLaunchKernelGGL(valu_probe_kernel<2>, dim3(blocks), dim3(256), 0, stream, iters, out); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_l1_probe(int pattern, const uint8_t *buf, int iters, float *out, int blocks, hipStream_t stream) {
    switch (pattern) {
    case 0: hipLaunchKernelGGL(l1_probe_kernel<0>, dim3(blocks), dim3(256), 0, stream, buf, iters, out); break;
    case 1: hipLaunchKernelGGL(l1_probe_kernel<1>, dim3(blocks), dim3(256), 0, stream, buf, iters, out); break;
    case 2: hipLaunchKernelGGL(l1_probe_kernel<2>, dim3(blocks), dim3(256), 0, stream, buf, iters, out); break;
    case 3: hipLaunchKernelGGL(l1_probe_kernel<3>, dim3(blocks), dim3(256), 0, stream, buf, iters, out); break;
    case 4: hipLaunchKernelGGL(l1_probe_kernel<4>, dim3(blocks), dim3(256), 0, stream, buf, iters, out); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_gather_probe(const TetRec *recs, uint32_t n, int iters, uint32_t *out, int blocks,
                               hipStream_t stream) {
    hipLaunchKernelGGL(gather_probe_kernel, dim3(blocks), dim3(256), 0, stream, recs, n, iters, out);
    return hipGetLastError();
}

hipError_t launch_copy(const void *src, void *dst, uint64_t bytes, hipStream_t stream) {
    uint64_t n = bytes / sizeof(uint4);
    hipLaunchKernelGGL(copy_kernel, dim3(8192), dim3(256), 0, stream, (const uint4 *)src, (uint4 *)dst, n);
    return hipGetLastError();
}
}  // namespace octmps
