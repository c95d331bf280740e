// FP64 CPU oracle for the OCT-MPS photon transport path.
//
// TEST INFRASTRUCTURE ONLY. This file restates, in plain C++ and IEEE-754 double, the
// reference's per-photon Monte Carlo loop (src/octmps_kernel.cuh:10-29, 448-1086,
// 1329-1402) so that tests, smoke() and the CPU-baseline leg of bench.py can check the HIP
// kernels. Nothing in the product library links or calls it.
//
// The reference's own kernel (src/octmps_kernel.cuh) cannot be built here: it needs a CUDA
// compiler and runtime (device qualifiers, atomics, __ddiv_rd, __uint2float_rz,
// cudaMemcpyToSymbol), and stand-ins for them are not allowed. Its arithmetic is
// restated from the IEEE-754 definitions instead:
//   __ddiv_rd(a, b)      binary64 quotient rounded toward -inf (div_rd below);
//   __uint2float_rz(u)   u truncated to 24 significant bits (rz24 below);
//   rsqrt(x)             1/sqrt(x) (CUDA's rsqrt may differ by ulps);
//   log, sqrt, sincos, acos, fabs, fmin, fmax, copysign   host libm.
//
// Deliberate, documented differences from the reference (the same ones as the HIP path):
//   * photon streams are per photon (seeded from (seed, salt, photon index)) instead of per
//     thread with a racing photon counter (src/octmps_kernel.cuh:1388);
//   * every photon launches at the probe (the reference's first wave uses an unset probe,
//     src/main.cu:534);
//   * Max is an exact maximum (the reference's compare-and-store races,
//     src/octmps_kernel.cuh:710-713);
//   * a photon exceeding max_steps transport steps is stopped and counted (no guard exists
//     in the reference);
//   * a negative bin position saturates to 0 like CUDA's double-to-unsigned conversion.
//
// Pins (tests/test_cpu_rng.py, tests/test_cpu_oracle_pins.py, tests/test_cpu_formats.py):
// MWC draws and stream seeding against big-integer arithmetic; Fresnel against the textbook
// formula; the HG and biased samplers as inverse CDFs; equations (8) and (9) through their
// sphere integrals; Spin's mean direction; HitBoundary against a barycentric brute force;
// the detector tally's window, counters and class rule on enumerated edge cases; SpinBias's
// branch choice (src/octmps_kernel.cuh:918-968), the detector filter rule (:667-681) and the
// roulette with the continuation-photon restore (:1376-1391) on enumerated states, one per
// side of each condition (tests/test_cpu_oracle_branches.py); the whole loop against a
// glass-slab known answer. The reference ships no golden outputs and its kernel cannot run
// here, so none of these is compared with reference output. ComputeStepSize and Drop are
// pinned through event traces (tests/test_cpu_oracle_transport_pins.py): the optical depth of
// every complete free path follows Exp(1) truncated at the ray's optical distance to the mesh
// boundary, across faces between regions of different mu_t, and the weight ratio between
// consecutive scatterings is mus/(mua+mus). The ORACLE_PLANT_* macros build deliberately
// wrong variants that those tests must reject; liboracle.so is built without them.
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../include/octmps.h"

namespace {

// ---------------------------------------------------------------- IEEE helpers
// Quotient rounded toward -inf: the round-to-nearest quotient is stepped down one ulp when
// the exact residual a - q*b (an fma, exact for normal operands) shows it is too large.
inline double div_rd(double a, double b) {
    double q = a / b;
    if (!std::isfinite(q) || b == 0.0) return q;
    double r = std::fma(-q, b, a);
    if (r != 0.0 && ((r < 0.0) != (b < 0.0))) q = std::nextafter(q, -INFINITY);
    return q;
}

inline uint32_t rz24(uint32_t u) {
    if (u < (1u << 24)) return u;
    int bits = 32 - __builtin_clz(u);
    return u & (0xffffffffu << (bits - 24));
}

const double kCosZero = (1.0F - 1.0E-14);      // src/octmps.h:29
const double kCos90 = 1.0E-6;                  // src/octmps.h:28
const double kPi = 3.14159265359;              // src/octmps.h:20
const double kWeight = 1E-4;                   // src/octmps.h:22
const double kChance = 0.1;                    // src/octmps.h:23
const int kSub = 6;                            // src/octmps.h:34

struct Rng {
    uint64_t x;
    uint32_t a;
};

// rand_MWC_co / rand_MWC_oc, src/octmps_kernel.cuh:10-29.
inline double rand_co(Rng &r) {
    r.x = (r.x & 0xffffffffull) * r.a + (r.x >> 32);
    return div_rd((double)(float)rz24((uint32_t)r.x), (double)0x100000000ull);
}
inline double rand_oc(Rng &r) { return 1.0 - rand_co(r); }

// Rspecular, src/main.cu:29-37.
inline double rspecular(double ni, double nt) {
    double t = (ni - nt) / (ni + nt);
    return t * t;
}

// ---------------------------------------------------------------- streams
inline uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
inline uint64_t mwc(uint64_t x, uint32_t a) { return (x & 0xffffffffull) * a + (x >> 32); }
inline bool bad_state(uint64_t x, uint32_t a) {
    return (x == 0) | ((uint32_t)(x >> 32) >= (a - 1u)) | ((uint32_t)x >= 0xffffffffu);
}

void streams_for(uint64_t seed, uint64_t salt, uint64_t photon, const uint32_t *mult, uint32_t K,
                 Rng out[3]) {
    uint64_t h = splitmix64(splitmix64(splitmix64(seed) ^ salt) ^ photon);
    uint32_t k = (uint32_t)(photon % K);
    for (uint32_t s = 0; s < 3; ++s) {
        uint32_t begin = mult[s], a = mult[3 + 3 * k + s];
        uint64_t z = splitmix64(h + 0x9E3779B97F4A7C15ull * (uint64_t)(s + 1));
        uint32_t hi = (uint32_t)(((uint64_t)(uint32_t)(z >> 32) * (uint64_t)(begin - 1u)) >> 32);
        uint32_t lo = (uint32_t)z;
        if (lo == 0xffffffffu) lo = 0;
        uint64_t xinit = ((uint64_t)hi << 32) | lo;
        if (xinit == 0) xinit = 1;
        uint64_t x = 0;
        int pass = 0;
        // init_RNG's loop, src/octmps_kernel.cuh:109-127.
        while (bad_state(x, a) && pass < 64) {
            xinit = mwc(xinit, begin);
            x = (uint32_t)floor((((double)((uint32_t)xinit)) / (double)0x100000000ull) * a);
            x = x << 32;
            xinit = mwc(xinit, begin);
            x += (uint32_t)xinit;
            ++pass;
        }
        if (bad_state(x, a)) x = (1ull << 32) | 1ull;
        out[s].x = x;
        out[s].a = a;
    }
}

// ---------------------------------------------------------------- model
struct Region {  // RegionStructGPU, src/octmps_kernel.h:132-140
    double n, muas, rmuas, mua_muas, g;
};

struct Sim {  // SimParamGPU, src/octmps_kernel.h:112-130
    double Rspecular, TargetDepthMin, TargetDepthMax, BackwardBiasCoefficient;
    double rndStepSizeInTissue, MaxCollectingAngleDeg, MaxCollectingRadius;
    double ProbabilityAdditionalBias, OpticalDepthShift, CoherenceLengthSource;
    uint64_t NumOpticalDepthLengthSteps;
    int TypeBias;
    uint32_t rootIdx;
    Region reg[OCTMPS_MAX_REGIONS];
    const double *planes;
    const int32_t *links;
    const int32_t *regions;
};

// InitDeviceConstantMemory, src/octmps_kernel.cuh:177-220, with the glass rule of
// read_simulation_data (src/octmps_io.cpp:624-626).
void make_sim(const octmps_params *p, Sim &s) {
    std::memset(&s, 0, sizeof(s));
    s.TargetDepthMin = p->target_depth_min;
    s.TargetDepthMax = p->target_depth_max;
    s.BackwardBiasCoefficient = p->backward_bias_coefficient;
    s.rndStepSizeInTissue = p->rnd_step_size_in_tissue;
    s.MaxCollectingAngleDeg = p->max_collecting_angle_deg;
    s.MaxCollectingRadius = p->max_collecting_radius;
    s.ProbabilityAdditionalBias = p->probability_additional_bias;
    s.OpticalDepthShift = p->optical_depth_shift;
    s.CoherenceLengthSource = p->coherence_length_source;
    s.NumOpticalDepthLengthSteps = (uint64_t)p->num_depth_bins;
    s.TypeBias = p->type_bias;
    s.reg[0].n = p->regions[0].n;
    for (uint32_t i = 1; i < p->n_regions && i < OCTMPS_MAX_REGIONS; ++i) {
        const octmps_region &r = p->regions[i];
        double mutr = (r.mus == 0.0) ? DBL_MAX : 1.0 / (r.mua + r.mus);
        s.reg[i].n = r.n;
        s.reg[i].muas = 1.0 / mutr;
        s.reg[i].rmuas = mutr;
        s.reg[i].mua_muas = r.mua * mutr;
        s.reg[i].g = r.g;
    }
}

struct Photon {  // PhotonStructGPU, src/octmps_kernel.h:219-255
    int NextTetrahedron;
    uint32_t hit, dead, FstBackReflectionFlag, rootIdx, NumBackwardsSpecularReflections;
    double x, y, z, ux, uy, uz, w, s, sleft, MinCos, LocationFstBias, OpticalPath, MaxDepth;
    double LikelihoodRatio, LikelihoodRatioAfterFstBias;
#ifdef ORACLE_PLANT_LEFTOVER_OLD_REGION
    double plant_mut;
#endif
};

struct Tally {
    int B;
    std::vector<double> sum, sumsq, max;  // [2][B]
    std::vector<uint64_t> count;          // [2][B]
    uint64_t nfilt = 0, nc1 = 0, nc2 = 0;
    explicit Tally(int b) : B(b), sum(2 * b), sumsq(2 * b), max(2 * b), count(2 * b) {}
};

struct PhotonLog {  // per-photon summary used to pair GPU and CPU photons
    uint64_t steps = 0, detections = 0;
    double contrib = 0;
};

inline int region_of(const Sim &s, uint32_t t) { return s.regions[t]; }

// CopyPhotonStruct, src/octmps_kernel.cuh:448-473 (hit and LikelihoodRatioAfterFstBias
// are not copied).
void copy_photon(const Photon &o, Photon &d) {
    d.x = o.x; d.y = o.y; d.z = o.z;
    d.ux = o.ux; d.uy = o.uy; d.uz = o.uz;
    d.w = o.w; d.dead = o.dead; d.MinCos = o.MinCos;
    d.rootIdx = o.rootIdx; d.NextTetrahedron = o.NextTetrahedron;
    d.s = o.s; d.sleft = o.sleft;
    d.OpticalPath = o.OpticalPath; d.MaxDepth = o.MaxDepth; d.LikelihoodRatio = o.LikelihoodRatio;
    d.FstBackReflectionFlag = o.FstBackReflectionFlag;
    d.LocationFstBias = o.LocationFstBias;
    d.NumBackwardsSpecularReflections = o.NumBackwardsSpecularReflections;
}

// LaunchPhoton, src/octmps_kernel.cuh:475-503.
void launch(const Sim &s, Photon &p, double px, double py, double pz) {
    p.x = px; p.y = py; p.z = pz;
    p.ux = p.uy = 0.0;
    p.uz = 1.0;
    p.w = 1.0 - s.Rspecular;
    p.dead = 0; p.hit = 0;
    p.MinCos = 0.0; p.s = 0.0; p.sleft = 0.0;
    p.OpticalPath = 0.0; p.MaxDepth = 0.0;
    p.LikelihoodRatio = 1.0; p.LikelihoodRatioAfterFstBias = 1.0;
    p.FstBackReflectionFlag = 0; p.LocationFstBias = -1.0;
    p.NumBackwardsSpecularReflections = 0;
    p.NextTetrahedron = -1;
    p.rootIdx = s.rootIdx;
}

// ComputeStepSize, src/octmps_kernel.cuh:505-515.
void compute_step(const Sim &s, Photon &p, Rng &g) {
    if (p.sleft == 0.0) {
        double r = rand_oc(g);
        p.s = -log(r) * s.reg[region_of(s, p.rootIdx)].rmuas;
    } else {
#ifdef ORACLE_PLANT_LEFTOVER_OLD_REGION
        // Planted mistake (pin tests only): the leftover converted with the old region's mu_t.
        p.s = p.sleft / p.plant_mut;
#else
        p.s = p.sleft * s.reg[region_of(s, p.rootIdx)].rmuas;
#endif
        p.sleft = 0.0;
    }
}

// HitBoundary, src/octmps_kernel.cuh:517-572 (planes are N*sign per tet face).
uint32_t hit_boundary(const Sim &s, Photon &p) {
    uint32_t t = p.rootIdx;
    double min_distance = 1E10;
    int imin = -1;
    double cosn[4];
    const double *pl = s.planes + 16 * (int64_t)t;
    for (int i = 0; i < 4; ++i)
        cosn[i] = pl[4 * i] * p.ux + pl[4 * i + 1] * p.uy + pl[4 * i + 2] * p.uz;
    for (int i = 0; i < 4; ++i) {
        if (cosn[i] < 0.0) {
            double perp = pl[4 * i] * p.x + pl[4 * i + 1] * p.y + pl[4 * i + 2] * p.z + pl[4 * i + 3];
            double dist = div_rd(-perp, cosn[i]);
            if (dist < min_distance) {
                min_distance = dist;
                imin = i;
                p.MinCos = cosn[i];
            }
        }
    }
    if (p.s > min_distance) {
        double mut = s.reg[region_of(s, t)].muas;
        p.sleft = (p.s - min_distance) * mut;
#ifdef ORACLE_PLANT_LEFTOVER_OLD_REGION
        p.plant_mut = mut;
#endif
        p.s = min_distance;
        p.NextTetrahedron = imin;
        return 1;
    }
    p.NextTetrahedron = -1;
    return 0;
}

// Hop, src/octmps_kernel.cuh:574-582.
void hop(Photon &p) {
    p.x += p.s * p.ux;
    p.y += p.s * p.uy;
    p.z += p.s * p.uz;
    p.OpticalPath += p.s;
    if (p.MaxDepth < p.z) p.MaxDepth = p.z;
}

// Drop, src/octmps_kernel.cuh:584-588.
void drop(const Sim &s, Photon &p) {
#ifdef ORACLE_PLANT_DROP_MUA_OVER_MUS
    // Planted mistake (pin tests only): absorbed fraction mua/mus instead of mua/(mua+mus).
    const Region &r = s.reg[region_of(s, p.rootIdx)];
    const double mua = r.mua_muas * r.muas, mus = r.muas - mua;
    double dwa = p.w * (mua / mus);
#else
    double dwa = p.w * s.reg[region_of(s, p.rootIdx)].mua_muas;
#endif
    p.w -= dwa;
}

// Fresnel reflectance of FastReflectTransmit, src/octmps_kernel.cuh:608-645.
double fresnel(double ni, double nt, double ca1) {
    double ni_nt = div_rd(ni, nt);
    double sa1 = sqrt(1.0 - ca1 * ca1);
    if (ca1 > kCosZero) sa1 = 0.0;
    double sa2 = fmin(ni_nt * sa1, 1.0);
    double ca2 = sqrt(1.0 - sa2 * sa2);
    double ca1ca2 = ca1 * ca2, sa1sa2 = sa1 * sa2, sa1ca2 = sa1 * ca2, ca1sa2 = ca1 * sa2;
    if (ca1 > kCosZero) {
        sa1ca2 = 1.0;
        ca1sa2 = ni_nt;
    }
    double cam = ca1ca2 + sa1sa2, sap = sa1ca2 + ca1sa2, sam = sa1ca2 - ca1sa2;
    double r = div_rd(sam, sap * cam);
    r *= r;
    r *= (ca1ca2 * ca1ca2 + sa1sa2 * sa1sa2);
    if (ca1 < kCos90 || sa2 == 1.0) r = 1.0;
    return r;
}

void tally_detection(const Sim &s, const Photon &p, double rF, Tally &T, PhotonLog *log) {
    // src/octmps_kernel.cuh:684-741.
    T.nfilt++;
    double op = p.OpticalPath;
    double depth = div_rd(op, 2.0);
    double q = div_rd(div_rd(op, 2.0) - s.OpticalDepthShift, div_rd(s.CoherenceLengthSource, (double)kSub));
    unsigned int vp = q <= 0.0 ? 0u : (q >= 4294967295.0 ? 4294967295u : (unsigned int)q);
    double c = p.w * (1.0 - rF) * p.LikelihoodRatio;
    if (log) {
        log->detections++;
        log->contrib += c;
    }
    int lo = (int)((double)vp - div_rd(kSub, 2.0));
    for (int ii = lo; (double)ii < (double)vp + div_rd(kSub, 2.0); ++ii) {
        if (ii >= 0 && (uint64_t)ii < s.NumOpticalDepthLengthSteps) {
            bool c1 = p.MaxDepth > (depth - div_rd(s.CoherenceLengthSource, 2.0));
            int cls = c1 ? 0 : 1;
            if ((unsigned int)ii == vp) {
                if (c1) T.nc1++;
                else T.nc2++;
            }
            int k = cls * T.B + ii;
            T.sum[k] += c;
            if (T.max[k] < c) T.max[k] = c;
            T.sumsq[k] += c * c;
            T.count[k]++;
        }
    }
}

// FastReflectTransmit, src/octmps_kernel.cuh:590-805 (PARTIAL_REFLECTION = 1).
// The spatial filter of a photon leaving through the top surface (uz < 0),
// src/octmps_kernel.cuh:667-681: inside the collecting radius (strict) and cone (strict),
// and then TypeBias 0 or 3, or the TypeBias-37 photon was biased backwards, or (TypeBias
// other than 37) its first bias happened at its maximum depth, or it met a backward
// specular reflection.
bool detect_filter(const Sim &s, const Photon &p, double px, double py) {
    if (!(sqrt((p.x - px) * (p.x - px) + (p.y - py) * (p.y - py)) < s.MaxCollectingRadius &&
          acos(-p.uz) < s.MaxCollectingAngleDeg * kPi / 180))
        return false;
    if (s.TypeBias == 0 || s.TypeBias == 3) return true;
    return (p.FstBackReflectionFlag && s.TypeBias == 37) || (p.LocationFstBias == p.MaxDepth && s.TypeBias != 37) ||
           p.NumBackwardsSpecularReflections > 0;
}

void reflect_transmit(const Sim &s, Photon &p, Rng &rr, double px, double py, double pz, Tally &T,
                      PhotonLog *log) {
    uint32_t t = p.rootIdx;
    double ca1 = -(p.MinCos);
    double ni = s.reg[region_of(s, t)].n, nt;
    int next = p.NextTetrahedron;
    int nidx = s.links[4 * (int64_t)t + next];
    if (nidx == -1) nt = s.reg[0].n;
    else nt = s.reg[region_of(s, (uint32_t)nidx)].n;
    double rF = fresnel(ni, nt, ca1);
    const double *pl = s.planes + 16 * (int64_t)t + 4 * next;
    double uxr = p.ux, uyr = p.uy, uzr = p.uz;
    double uxf = 2 * ca1 * pl[0] + p.ux;
    double uyf = 2 * ca1 * pl[1] + p.uy;
    double uzf = 2 * ca1 * pl[2] + p.uz;
    double rnd = rand_co(rr);
    if (nidx == -1 && rF < 1.0) {
        p.ux = uxr; p.uy = uyr; p.uz = uzr;
        if (p.uz < 0) {
            if (detect_filter(s, p, px, py)) tally_detection(s, p, rF, T, log);
        } else {
            p.NumBackwardsSpecularReflections++;
        }
        p.w *= rF;
        p.ux = uxf; p.uy = uyf; p.uz = uzf;
    } else if (rnd > rF) {
        p.rootIdx = (uint32_t)nidx;
        p.ux = uxr; p.uy = uyr; p.uz = uzr;
        p.NextTetrahedron = -1;
    } else {
        p.NextTetrahedron = -1;
        if (p.uz > 0.0) p.NumBackwardsSpecularReflections++;
        p.ux = uxf; p.uy = uyf; p.uz = uzf;
    }
}

// SpinTheta, src/octmps_kernel.cuh:807-820.
double spin_theta(double g, Rng &r) {
    double cost;
    if (g == 0.0) {
        cost = 2.0 * rand_oc(r) - 1.0;
    } else {
        double temp = div_rd((1.0 - g * g), (1.0 - g + 2.0 * g * rand_oc(r)));
        cost = div_rd((1.0 + g * g - temp * temp), (2.0 * g));
        if (cost < -1.0) cost = -1.0;
        else if (cost > 1.0) cost = 1.0;
    }
    return cost;
}

// Spin, src/octmps_kernel.cuh:822-879 (SINGLE_PRECISION off: no renormalisation).
void spin(double g, Photon &p, Rng &r) {
    double cost = spin_theta(g, r);
    double sint = sqrt(1.0 - cost * cost);
    double psi = 2.0 * kPi * rand_co(r);
    double sinp = sin(psi), cosp = cos(psi);
    double stcp = sint * cosp, stsp = sint * sinp;
    double ux = p.ux, uy = p.uy, uz = p.uz;
    if (fabs(uz) > kCosZero) {
        p.ux = stcp;
        p.uy = stsp;
        p.uz = cost * ((uz >= 0.0) ? 1.0 : -1.0);
    } else {
        double temp = 1.0 / sqrt(1.0 - uz * uz);
        p.ux = (stcp * ux * uz - stsp * uy) * temp + ux * cost;
        p.uy = (stcp * uy * uz + stsp * ux) * temp + uy * cost;
        p.uz = div_rd(-stcp, temp) + uz * cost;
    }
}

// SpinThetaForwardFstBias, src/octmps_kernel.cuh:881-896.
double spin_theta_bias(double g, Rng &r) {
    double cost;
    if (g == 0.0) {
        cost = rand_oc(r);
    } else {
        double u = rand_oc(r);
        double temp = div_rd(u, (1.0 - g)) + div_rd((1.0 - u), sqrt(g * g + 1.0));
        cost = div_rd(g * g + 1.0 - div_rd(1.0, temp * temp), (2.0 * g));
        if (cost < -1.0) cost = -1.0;
        else if (cost > 1.0) cost = 1.0;
    }
    return cost;
}

struct BiasInfo {  // diagnostics of one SpinBias call
    int flag;      // ReachedTargetOpticalDepthAndGoingBackwarFlag
    int selected;  // BiasFunctionRandomlySelected
    double factor; // LikelihoodRatioIncreaseFactor (1 when no update)
};

// SpinBias, src/octmps_kernel.cuh:898-1086.
void spin_bias(const Sim &s, double g, Photon &p, Photon &cont, Rng &r, double px, double py,
               double pz, BiasInfo *info) {
    double g2 = g * g;
    double cost, sint, cosp, sinp, costg, costg1, costg2, psi, temp;
    double ux = p.ux, uy = p.uy, uz = p.uz;
    double uxo = ux, uyo = uy, uzo = uz;
    double a = s.BackwardBiasCoefficient, atmp = 0.0;
    int F = 0, first = 0;
    if (p.z > s.TargetDepthMin && p.z < s.TargetDepthMax && p.uz > 0.0 && !p.FstBackReflectionFlag) {
        copy_photon(p, cont);
        F = 1;
        p.FstBackReflectionFlag = 1;
        p.LocationFstBias = p.z;
        atmp = a;
        first = 1;
    }
    double vx = px - p.x, vy = py - p.y, vz = pz - p.z;
    double len = sqrt(vx * vx + vy * vy + vz * vz);
    vx = div_rd(vx, len);
    vy = div_rd(vy, len);
    vz = div_rd(vz, len);
    if ((p.FstBackReflectionFlag || p.NumBackwardsSpecularReflections > 0) && !F) {
        F = 2;
        double mut = s.reg[region_of(s, p.rootIdx)].muas;
        double next = div_rd(-log(s.rndStepSizeInTissue), mut);
        double dist = sqrt((p.x - px) * (p.x - px) + (p.y - py) * (p.y - py) + (p.z - pz) * (p.z - pz));
        if (next >= dist && acos(-vz) <= (s.MaxCollectingAngleDeg * kPi / 180)) F = 1;
        atmp = a;
    }
    int sel = 0;
    if (F) {
        sel = (rand_co(r) <= s.ProbabilityAdditionalBias) ? 1 : 2;
        if (F == 1 || sel == 1) {
            cost = spin_theta_bias(atmp, r);
            ux = vx;
            uy = vy;
            uz = vz;
        } else {
            cost = spin_theta(g, r);
        }
    } else {
        cost = spin_theta(g, r);
    }
    cost = fmax(cost, -1.0);
    cost = fmin(cost, 1.0);
    sint = sqrt(1.0 - cost * cost);
    psi = 2.0 * kPi * rand_co(r);
    sinp = sin(psi);
    cosp = cos(psi);
    double stcp = sint * cosp, stsp = sint * sinp;
    if (fabs(uz) > kCosZero) {
        p.ux = stcp;
        p.uy = stsp;
        p.uz = copysign(cost, uz * cost);
    } else {
        temp = 1.0 / sqrt(1.0 - uz * uz);
        p.ux = (stcp * ux * uz - stsp * uy) * temp + ux * cost;
        p.uy = (stcp * uy * uz + stsp * ux) * temp + uy * cost;
        p.uz = div_rd(-stcp, temp) + uz * cost;
    }
    costg = uxo * p.ux + uyo * p.uy + uzo * p.uz;
    costg2 = costg;
    costg1 = vx * p.ux + vy * p.uy + vz * p.uz;
    double factor = 1.0;
    if (atmp != 0.0) {
        double opa2 = 1 + atmp * atmp;
        double sq = sqrt(opa2);
        if (F == 1) {
            // Equation (8) of JBO 19(4) 046001.
            double t1 = opa2 - 2.0 * atmp * cost;
            double t2 = 1.0 + g2 - 2.0 * g * costg;
            factor = div_rd((1 - g2) * (sq - 1.0 + atmp) * sqrt(t1 * t1 * t1),
                            2.0 * atmp * (1 - atmp) * sq * sqrt(t2 * t2 * t2));
        } else {
            double c1, c2;
            if (sel == 1) {
                c1 = cost;
                c2 = costg2;
            } else {
                c1 = costg1;
                c2 = cost;
            }
            // Equation (9) of JBO 19(4) 046001.
            double t1 = opa2 - 2.0 * atmp * c1;
            double pdf1 = div_rd(sq * atmp * (1.0 - atmp), (sq - 1.0 + atmp) * sqrt(t1 * t1 * t1));
            double t2 = sqrt(1.0 + g2 - 2.0 * g * c2);
            double pdf2 = div_rd(1.0 - g2, (2.0 * (t2 * t2 * t2)));
            factor = div_rd(pdf2, (s.ProbabilityAdditionalBias * pdf1 + (1.0 - s.ProbabilityAdditionalBias) * pdf2));
        }
        p.LikelihoodRatio *= factor;
        if (first == 1) p.LikelihoodRatioAfterFstBias = p.LikelihoodRatio;
    }
    if (info) {
        info->flag = F;
        info->selected = sel;
        info->factor = factor;
    }
}

// Events recorded by trace_photon.
enum { EV_LAUNCH = 0, EV_SCATTER = 1, EV_TRANSMIT = 2, EV_REFLECT = 3, EV_BOUNDARY = 4, EV_SURVIVE = 5,
       EV_RESTORE = 6, EV_DEATH = 7, EV_KILLED = 8 };

// Roulette of a photon below the weight threshold, src/octmps_kernel.cuh:1376-1391: it
// survives with probability CHANCE (weight / CHANCE; a zero weight never survives);
// otherwise, unless TypeBias is 3, a photon that was biased backwards is replaced by its continuation
// photon (the copy taken before its first bias; LikelihoodRatioAfterFstBias is not part of
// the copy), spun with HG and given LR = 1 - LR_afterFstBias (1 when that is >= 1), and
// any other photon dies.
enum { kUntouched = 0, kSurvived = 1, kRestored = 2, kDied = 3 };
int roulette(const Sim &s, Photon &p, const Photon &cont, Rng rng[3]) {
    if (!(p.w < kWeight)) return kUntouched;
    double rnd = rand_co(rng[0]);
    if (p.w != 0.0 && rnd < kChance) {
        p.w *= div_rd(1.0, kChance);
        return kSurvived;
    }
    if (p.FstBackReflectionFlag && s.TypeBias != 3) {
        double lrt = p.LikelihoodRatioAfterFstBias;
        copy_photon(cont, p);
        spin(s.reg[region_of(s, p.rootIdx)].g, p, rng[2]);
        if (lrt < 1) p.LikelihoodRatio = 1 - lrt;
        else p.LikelihoodRatio = 1;
        return kRestored;
    }
    return kDied;
}

struct TraceSink {
    double *out;
    int cap, n;
    void add(int ev, const Photon &p, int F, double fac) {
        if (!out || n >= cap) {
            ++n;
            return;
        }
        double *r = out + 16 * (int64_t)n;
        r[0] = ev; r[1] = p.rootIdx; r[2] = p.x; r[3] = p.y; r[4] = p.z;
        r[5] = p.ux; r[6] = p.uy; r[7] = p.uz; r[8] = p.w; r[9] = p.LikelihoodRatio;
        r[10] = p.OpticalPath; r[11] = p.MaxDepth; r[12] = p.FstBackReflectionFlag;
        r[13] = p.NumBackwardsSpecularReflections; r[14] = F; r[15] = fac;
        ++n;
    }
};

// One photon from launch to death (OCTMPSKernel's loop body, src/octmps_kernel.cuh:1348-1393).
void run_photon(const Sim &s, Rng rng[3], double px, double py, double pz, uint64_t max_steps, Tally &T,
                PhotonLog &log, bool &killed, TraceSink *tr) {
    Photon p, cont;
    std::memset(&p, 0, sizeof p);
    launch(s, p, px, py, pz);
    copy_photon(p, cont);
    killed = false;
    if (tr) tr->add(EV_LAUNCH, p, 0, 1.0);
    for (;;) {
        if (log.steps >= max_steps) {
            killed = true;
            if (tr) tr->add(EV_KILLED, p, 0, 1.0);
            return;
        }
        log.steps++;
        compute_step(s, p, rng[0]);
        p.hit = hit_boundary(s, p);
        hop(p);
        if (p.hit) {
            uint32_t before = p.rootIdx;
            int nb = s.links[4 * (int64_t)p.rootIdx + p.NextTetrahedron];
            reflect_transmit(s, p, rng[1], px, py, pz, T, &log);
            if (tr) tr->add(nb < 0 ? EV_BOUNDARY : (p.rootIdx != before ? EV_TRANSMIT : EV_REFLECT), p, 0, 1.0);
        } else {
            drop(s, p);
            BiasInfo bi = {0, 0, 1.0};
            int r = region_of(s, p.rootIdx);
            switch (s.TypeBias) {
                case 0: spin(s.reg[r].g, p, rng[2]); break;
                case 37: spin_bias(s, s.reg[r].g, p, cont, rng[2], px, py, pz, &bi); break;
            }
            if (tr) tr->add(EV_SCATTER, p, bi.flag, bi.factor);
        }
        const int fate = roulette(s, p, cont, rng);
        if (fate == kSurvived && tr) tr->add(EV_SURVIVE, p, 0, 1.0);
        if (fate == kRestored && tr) tr->add(EV_RESTORE, p, 0, 1.0);
        if (fate == kDied) {
            if (tr) tr->add(EV_DEATH, p, 0, 1.0);
            return;
        }
    }
}

struct RunArgs {
    Sim sim;
    double px, py, pz;
    uint64_t seed, salt, off, n, max_steps;
    const uint32_t *mult;
    uint32_t K;
};

}  // namespace

extern "C" {

// Runs photons [scan->photon_offset, +scan->n_photons) of one A-scan on n_threads threads
// (static contiguous split, per-thread tallies merged in thread order).
// tally_out[2 classes][4][B]: Sum, SumSq, Max, Count (count as double).
// counters[6]: NumFiltered, ClassI, ClassII, photons, photon-steps, killed.
// photon_log (optional) [n_photons][3]: steps, detections, summed contribution.
int oracle_run_ascan(const double *planes, const int32_t *links, const int32_t *regions, int64_t n_tets,
                     const octmps_params *p, const octmps_ascan *scan, uint64_t seed, const uint32_t *mult,
                     uint32_t n_triplets, uint64_t max_steps, int n_threads, double *tally_out,
                     uint64_t *counters, double *photon_log) {
    if (!planes || !links || !regions || !p || !scan || !mult || n_triplets == 0 || !tally_out || !counters)
        return 1;
    if (scan->root_tet < 0 || scan->root_tet >= n_tets) return 2;
    RunArgs a;
    make_sim(p, a.sim);
    a.sim.planes = planes;
    a.sim.links = links;
    a.sim.regions = regions;
    a.sim.rootIdx = (uint32_t)scan->root_tet;
    a.sim.Rspecular = rspecular(p->regions[0].n, p->regions[regions[scan->root_tet]].n);
    a.px = scan->x; a.py = scan->y; a.pz = scan->z;
    a.seed = seed; a.salt = scan->stream_salt; a.off = scan->photon_offset; a.n = scan->n_photons;
    a.max_steps = max_steps ? max_steps : (1ull << 22);
    a.mult = mult; a.K = n_triplets;
    const int B = p->num_depth_bins;
    if (n_threads < 1) n_threads = 1;
    std::vector<Tally> tallies(n_threads, Tally(B));
    std::vector<uint64_t> steps(n_threads, 0), killed(n_threads, 0), done(n_threads, 0);
    auto work = [&](int tid) {
        uint64_t lo = a.n * tid / n_threads, hi = a.n * (tid + 1) / n_threads;
        for (uint64_t i = lo; i < hi; ++i) {
            Rng rng[3];
            streams_for(a.seed, a.salt, a.off + i, a.mult, a.K, rng);
            PhotonLog log;
            bool k = false;
            run_photon(a.sim, rng, a.px, a.py, a.pz, a.max_steps, tallies[tid], log, k, nullptr);
            steps[tid] += log.steps;
            killed[tid] += k;
            done[tid]++;
            if (photon_log) {
                photon_log[3 * i] = (double)log.steps;
                photon_log[3 * i + 1] = (double)log.detections;
                photon_log[3 * i + 2] = log.contrib;
            }
        }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < n_threads; ++t) th.emplace_back(work, t);
    work(0);
    for (auto &x : th) x.join();
    Tally &T = tallies[0];
    for (int t = 1; t < n_threads; ++t) {
        for (int k = 0; k < 2 * B; ++k) {
            T.sum[k] += tallies[t].sum[k];
            T.sumsq[k] += tallies[t].sumsq[k];
            T.max[k] = fmax(T.max[k], tallies[t].max[k]);
            T.count[k] += tallies[t].count[k];
        }
        T.nfilt += tallies[t].nfilt;
        T.nc1 += tallies[t].nc1;
        T.nc2 += tallies[t].nc2;
    }
    for (int c = 0; c < 2; ++c)
        for (int b = 0; b < B; ++b) {
            int k = c * B + b;
            tally_out[(c * 4 + 0) * B + b] = T.sum[k];
            tally_out[(c * 4 + 1) * B + b] = T.sumsq[k];
            tally_out[(c * 4 + 2) * B + b] = T.max[k];
            tally_out[(c * 4 + 3) * B + b] = (double)T.count[k];
        }
    uint64_t st = 0, ki = 0, dn = 0;
    for (int t = 0; t < n_threads; ++t) {
        st += steps[t];
        ki += killed[t];
        dn += done[t];
    }
    counters[0] = T.nfilt;
    counters[1] = T.nc1;
    counters[2] = T.nc2;
    counters[3] = dn;
    counters[4] = st;
    counters[5] = ki;
    return 0;
}

// Event trace of one photon: out[max_events][16] (see TraceSink::add); returns the number
// of events (may exceed max_events, in which case the tail is dropped).
int oracle_trace_photon(const double *planes, const int32_t *links, const int32_t *regions, int64_t n_tets,
                        const octmps_params *p, const octmps_ascan *scan, uint64_t seed, const uint32_t *mult,
                        uint32_t n_triplets, uint64_t photon, uint64_t max_steps, int max_events, double *out) {
    if (!planes || !links || !regions || !p || !scan || !mult) return -1;
    if (scan->root_tet < 0 || scan->root_tet >= n_tets) return -2;
    Sim s;
    make_sim(p, s);
    s.planes = planes;
    s.links = links;
    s.regions = regions;
    s.rootIdx = (uint32_t)scan->root_tet;
    s.Rspecular = rspecular(p->regions[0].n, p->regions[regions[scan->root_tet]].n);
    Rng rng[3];
    streams_for(seed, scan->stream_salt, photon, mult, n_triplets, rng);
    Tally T(p->num_depth_bins);
    PhotonLog log;
    bool killed;
    TraceSink tr{out, max_events, 0};
    run_photon(s, rng, scan->x, scan->y, scan->z, max_steps ? max_steps : (1ull << 22), T, log, killed, &tr);
    return tr.n;
}

// ---- unit entry points -------------------------------------------------------------
double oracle_div_rd(double a, double b) { return div_rd(a, b); }

// n co-draws from (x, a); returns the final state in *x.
void oracle_rand_co(uint64_t *x, uint32_t a, int n, double *out) {
    Rng r{*x, a};
    for (int i = 0; i < n; ++i) out[i] = rand_co(r);
    *x = r.x;
}

void oracle_photon_streams(uint64_t seed, uint64_t salt, uint64_t photon, const uint32_t *mult, uint32_t K,
                           uint64_t x[3], uint32_t a[3]) {
    Rng r[3];
    streams_for(seed, salt, photon, mult, K, r);
    for (int i = 0; i < 3; ++i) {
        x[i] = r[i].x;
        a[i] = r[i].a;
    }
}

double oracle_fresnel(double ni, double nt, double ca1) { return fresnel(ni, nt, ca1); }

// SpinTheta / SpinThetaForwardFstBias with the uniform draw forced to u_oc = rand_oc.
// The draw is fed through a one-step MWC state that yields exactly u_oc when u_oc is a
// representable draw; other values are applied directly.
double oracle_spin_theta(double g, double u_oc) {
    if (g == 0.0) return 2.0 * u_oc - 1.0;
    double temp = div_rd((1.0 - g * g), (1.0 - g + 2.0 * g * u_oc));
    double cost = div_rd((1.0 + g * g - temp * temp), (2.0 * g));
    if (cost < -1.0) cost = -1.0;
    else if (cost > 1.0) cost = 1.0;
    return cost;
}

double oracle_spin_theta_bias(double g, double u_oc) {
    if (g == 0.0) return u_oc;
    double temp = div_rd(u_oc, (1.0 - g)) + div_rd((1.0 - u_oc), sqrt(g * g + 1.0));
    double cost = div_rd(g * g + 1.0 - div_rd(1.0, temp * temp), (2.0 * g));
    if (cost < -1.0) cost = -1.0;
    else if (cost > 1.0) cost = 1.0;
    return cost;
}

// HitBoundary on one tet: planes16 (pre-signed), position, direction, step s.
// out[0] hit, [1] face, [2] step after, [3] sleft (with mut), [4] MinCos.
void oracle_hit_boundary(const double *planes16, const double *pos, const double *dir, double s, double mut,
                         double *out) {
    Sim sim;
    std::memset(&sim, 0, sizeof sim);
    int32_t link[4] = {-1, -1, -1, -1}, reg[1] = {1};
    sim.planes = planes16;
    sim.links = link;
    sim.regions = reg;
    sim.reg[1].muas = mut;
    Photon p;
    std::memset(&p, 0, sizeof p);
    p.x = pos[0]; p.y = pos[1]; p.z = pos[2];
    p.ux = dir[0]; p.uy = dir[1]; p.uz = dir[2];
    p.s = s;
    uint32_t h = hit_boundary(sim, p);
    out[0] = h;
    out[1] = p.NextTetrahedron;
    out[2] = p.s;
    out[3] = p.sleft;
    out[4] = p.MinCos;
}

// One detection through the detector tally (src/octmps_kernel.cuh:684-741) of an empty
// tally: sums[2][B] (Class I, II), counts[2][B], counters[3] = filtered, Class I, Class II.
void oracle_tally_detection(const octmps_params *prm, double op, double maxz, double w, double rF, double lr,
                            double *sums, uint64_t *counts, uint64_t *counters) {
    Sim s;
    make_sim(prm, s);
    Photon p;
    std::memset(&p, 0, sizeof p);
    p.OpticalPath = op;
    p.MaxDepth = maxz;
    p.w = w;
    p.LikelihoodRatio = lr;
    Tally T(prm->num_depth_bins);
    tally_detection(s, p, rF, T, nullptr);
    for (int k = 0; k < 2 * T.B; ++k) {
        sums[k] = T.sum[k];
        counts[k] = T.count[k];
    }
    counters[0] = T.nfilt;
    counters[1] = T.nc1;
    counters[2] = T.nc2;
}

// n independent SpinBias calls on copies of one photon state; Spin stream per call from
// (seed, call index). state[12] = x,y,z,ux,uy,uz,w,LR,FstBack,NumBackSpec,region g,
// region muas. out[n][8] = ux',uy',uz', factor, F, selected, LR', FstBack'.
void oracle_spin_bias_batch(const octmps_params *prm, const double *state, double px, double py, double pz,
                            uint64_t seed, const uint32_t *mult, uint32_t K, int n, double *out) {
    Sim s;
    make_sim(prm, s);
    int32_t link[4] = {-1, -1, -1, -1}, reg[1] = {1};
    double planes[16] = {0};
    s.planes = planes;
    s.links = link;
    s.regions = reg;
    s.reg[1].g = state[10];
    s.reg[1].muas = state[11];
    for (int i = 0; i < n; ++i) {
        Photon p, cont;
        std::memset(&p, 0, sizeof p);
        p.x = state[0]; p.y = state[1]; p.z = state[2];
        p.ux = state[3]; p.uy = state[4]; p.uz = state[5];
        p.w = state[6]; p.LikelihoodRatio = state[7];
        p.FstBackReflectionFlag = (uint32_t)state[8];
        p.NumBackwardsSpecularReflections = (uint32_t)state[9];
        p.LocationFstBias = -1.0;
        p.LikelihoodRatioAfterFstBias = 1.0;
        p.rootIdx = 0;
        Rng r[3];
        streams_for(seed, 0, (uint64_t)i, mult, K, r);
        BiasInfo bi;
        spin_bias(s, state[10], p, cont, r[2], px, py, pz, &bi);
        double *o = out + 8 * (int64_t)i;
        o[0] = p.ux; o[1] = p.uy; o[2] = p.uz; o[3] = bi.factor; o[4] = bi.flag; o[5] = bi.selected;
        o[6] = p.LikelihoodRatio; o[7] = p.FstBackReflectionFlag;
    }
}

// Detector filter on one exit state: st[7] = x, y, uz, FstBack, LocationFstBias, MaxDepth,
// NumBackSpec. Returns 1 when the photon would be tallied.
int oracle_detect_filter(const octmps_params *prm, const double *st, double px, double py) {
    Sim s;
    make_sim(prm, s);
    Photon p;
    std::memset(&p, 0, sizeof p);
    p.x = st[0]; p.y = st[1]; p.uz = st[2];
    p.FstBackReflectionFlag = (uint32_t)st[3];
    p.LocationFstBias = st[4];
    p.MaxDepth = st[5];
    p.NumBackwardsSpecularReflections = (uint32_t)st[6];
    return detect_filter(s, p, px, py) ? 1 : 0;
}

// n independent roulettes of copies of one photon and its continuation photon, in a
// one-tet mesh of region 1 with HG parameter g; streams per call from (seed, call index).
// ph[10] and ct[10] = x,y,z,ux,uy,uz,w,LR,LR_afterFstBias,FstBack.
// out[n][10] = fate (0 untouched, 1 survived, 2 restored, 3 died), x,y,z,ux,uy,uz,w,LR,FstBack.
void oracle_roulette_batch(const octmps_params *prm, const double *ph, const double *ct, double g, uint64_t seed,
                           const uint32_t *mult, uint32_t K, int n, double *out) {
    Sim s;
    make_sim(prm, s);
    int32_t link[4] = {-1, -1, -1, -1}, reg[1] = {1};
    double planes[16] = {0};
    s.planes = planes;
    s.links = link;
    s.regions = reg;
    s.reg[1].g = g;
    auto load = [](const double *v, Photon &q) {
        std::memset(&q, 0, sizeof q);
        q.x = v[0]; q.y = v[1]; q.z = v[2];
        q.ux = v[3]; q.uy = v[4]; q.uz = v[5];
        q.w = v[6]; q.LikelihoodRatio = v[7]; q.LikelihoodRatioAfterFstBias = v[8];
        q.FstBackReflectionFlag = (uint32_t)v[9];
        q.LocationFstBias = -1.0;
        q.rootIdx = 0;
    };
    for (int i = 0; i < n; ++i) {
        Photon p, cont;
        load(ph, p);
        load(ct, cont);
        Rng r[3];
        streams_for(seed, 0, (uint64_t)i, mult, K, r);
        const int fate = roulette(s, p, cont, r);
        double *o = out + 10 * (int64_t)i;
        o[0] = fate; o[1] = p.x; o[2] = p.y; o[3] = p.z; o[4] = p.ux; o[5] = p.uy; o[6] = p.uz;
        o[7] = p.w; o[8] = p.LikelihoodRatio; o[9] = p.FstBackReflectionFlag;
    }
}

}  // extern "C"
