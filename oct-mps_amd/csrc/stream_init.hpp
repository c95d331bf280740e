// Photon stream seeding shared by the host library and the HIP kernels.
//
// The reference draws from three multiply-with-carry (MWC) streams per GPU thread -
// General (step size, roulette), ReflectT (Fresnel draw) and Spin (all scattering draws) -
// seeded once per thread by init_RNG (src/octmps_kernel.cuh:52-172), and hands photons to
// threads through a racing atomic counter (src/octmps_kernel.cuh:1388). Here every photon
// owns its three streams: a photon's state is a pure function of (seed, salt, photon index),
// so a result does not depend on which lane, slice or launch ran it, and the GPU and the CPU
// oracle consume identical numbers per photon.
//
// Derivation per photon p and stream s (0 General, 1 ReflectT, 2 Spin):
//   h      = splitmix64(splitmix64(splitmix64(seed) ^ salt) ^ p)
//   xinit  = a valid seeding state for multiplier table[s] made from splitmix64(h + G*(s+1))
//   a      = table[3 + 3*(p mod K) + s]
//   x      = init_RNG's rejection loop (src/octmps_kernel.cuh:109-127) driven by xinit.
#pragma once

#include <stdint.h>
#include <math.h>

#if defined(__HIP__)
#define OCTMPS_HD __host__ __device__ inline
#else
#define OCTMPS_HD inline
#endif

namespace octmps {

OCTMPS_HD uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// One MWC step, base 2^32 (rand_MWC_co, src/octmps_kernel.cuh:15).
OCTMPS_HD uint64_t mwc_next(uint64_t x, uint32_t a) {
    return (x & 0xffffffffull) * (uint64_t)a + (x >> 32);
}

OCTMPS_HD uint64_t photon_hash(uint64_t seed, uint64_t salt, uint64_t photon) {
    return splitmix64(splitmix64(splitmix64(seed) ^ salt) ^ photon);
}

// A seeding state accepted by init_RNG's check (src/octmps_kernel.cuh:86-92) for
// multiplier `begin`: carry < begin-1, low word < 2^32-1, state != 0.
OCTMPS_HD uint64_t seeding_state(uint64_t h, uint32_t s, uint32_t begin) {
    uint64_t z = splitmix64(h + 0x9E3779B97F4A7C15ull * (uint64_t)(s + 1));
    uint32_t hi = (uint32_t)(((uint64_t)(uint32_t)(z >> 32) * (uint64_t)(begin - 1u)) >> 32);
    uint32_t lo = (uint32_t)z;
    if (lo == 0xffffffffu) lo = 0u;
    uint64_t x = ((uint64_t)hi << 32) | lo;
    return x == 0 ? 1ull : x;
}

// init_RNG's per-stream loop (src/octmps_kernel.cuh:109-127): draw (carry, x) from the
// seeding MWC until the state is valid for multiplier `a`. The loop ends after the first
// pass except with probability ~1/a; a fixed valid state caps it at 64 passes.
OCTMPS_HD uint64_t derive_state(uint64_t xinit, uint32_t begin, uint32_t a) {
    uint64_t x = 0;
    for (int pass = 0; pass < 64; ++pass) {
        if (!((x == 0) | ((uint32_t)(x >> 32) >= (a - 1u)) | ((uint32_t)x >= 0xffffffffu)))
            return x;
        xinit = mwc_next(xinit, begin);
        x = (uint64_t)(uint32_t)floor((((double)((uint32_t)xinit)) / 4294967296.0) * (double)a);
        x = x << 32;
        xinit = mwc_next(xinit, begin);
        x += (uint32_t)xinit;
    }
    if (!((x == 0) | ((uint32_t)(x >> 32) >= (a - 1u)) | ((uint32_t)x >= 0xffffffffu)))
        return x;
    return (1ull << 32) | 1ull;
}

}  // namespace octmps
