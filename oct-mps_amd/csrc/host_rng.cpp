// Safe-prime multiplier table and photon stream seeding (host side).
//
// init_RNG (src/octmps_kernel.cuh:52-172) reads MWC multipliers from
// safeprimes_base32.txt, which the reference does not ship (.MISSING_LARGE_BLOBS:1). The
// table is regenerated here from its defining property: multipliers a < 2^32 for which
// a*2^32-1 and (a*2^32-2)/2 = a*2^31-1 are both prime (the maximal-period condition of
// the MWC generator, base b = 2^32), listed in descending order from 2^32-1.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "host_common.hpp"
#include "stream_init.hpp"

namespace octmps {
namespace {

typedef unsigned __int128 u128;

uint64_t mulmod(uint64_t a, uint64_t b, uint64_t m) { return (uint64_t)((u128)a * b % m); }

uint64_t powmod(uint64_t b, uint64_t e, uint64_t m) {
    uint64_t r = 1 % m;
    b %= m;
    while (e) {
        if (e & 1) r = mulmod(r, b, m);
        b = mulmod(b, b, m);
        e >>= 1;
    }
    return r;
}

// Deterministic Miller-Rabin for n < 3.3e24 (bases 2..37).
bool is_prime(uint64_t n) {
    static const uint64_t bases[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    if (n < 2) return false;
    for (uint64_t p : bases) {
        if (n % p == 0) return n == p;
    }
    uint64_t d = n - 1;
    int s = 0;
    while ((d & 1) == 0) {
        d >>= 1;
        ++s;
    }
    for (uint64_t a : bases) {
        uint64_t x = powmod(a, d, n);
        if (x == 1 || x == n - 1) continue;
        bool comp = true;
        for (int r = 1; r < s; ++r) {
            x = mulmod(x, x, n);
            if (x == n - 1) {
                comp = false;
                break;
            }
        }
        if (comp) return false;
    }
    return true;
}

// Generates the first n multipliers below 2^32 (descending). A segmented sieve over small
// primes q removes a with a*2^32 = 1 or a*2^31 = 1 (mod q) before Miller-Rabin.
void generate(uint32_t n, std::vector<uint32_t> &out) {
    out.clear();
    out.reserve(n);
    std::vector<uint32_t> small;
    for (uint32_t q = 3; q < 4096; q += 2) {
        bool pr = true;
        for (uint32_t d = 3; d * d <= q; d += 2)
            if (q % d == 0) {
                pr = false;
                break;
            }
        if (pr) small.push_back(q);
    }
    // residues r1 = inv(2^32) mod q, r2 = inv(2^31) mod q: a == r is excluded.
    std::vector<uint32_t> r1(small.size()), r2(small.size());
    for (size_t i = 0; i < small.size(); ++i) {
        uint64_t q = small[i];
        uint64_t t32 = powmod(2, 32, q), t31 = powmod(2, 31, q);
        r1[i] = (uint32_t)powmod(t32, q - 2, q);
        r2[i] = (uint32_t)powmod(t31, q - 2, q);
    }
    const uint32_t SEG = 1u << 16;
    std::vector<uint8_t> bad(SEG);
    uint64_t top = 0xffffffffull;  // a = top, top-1, ...
    while (out.size() < n && top > SEG) {
        uint64_t lo = top - SEG + 1;  // segment [lo, top]
        std::fill(bad.begin(), bad.end(), 0);
        for (size_t i = 0; i < small.size(); ++i) {
            uint64_t q = small[i];
            for (uint64_t r : {(uint64_t)r1[i], (uint64_t)r2[i]}) {
                uint64_t start = lo + ((r + q - lo % q) % q);
                for (uint64_t a = start; a <= top; a += q) bad[a - lo] = 1;
            }
        }
        for (uint64_t a = top; a >= lo && out.size() < n; --a) {
            if (bad[a - lo]) continue;
            uint64_t p1 = (a << 32) - 1, p2 = (a << 31) - 1;
            if (is_prime(p2) && is_prime(p1)) out.push_back((uint32_t)a);
        }
        top = lo - 1;
    }
}

std::mutex g_mu;
StreamTable g_table;
bool g_ready = false;
std::string g_file;
uint32_t g_triplets = 4096;

octmps_status load_file(const std::string &path, uint32_t need, std::vector<uint32_t> &out) {
    FILE *fp = fopen(path.c_str(), "r");
    if (!fp) return fail(OCTMPS_E_IO, "cannot open safeprimes file " + path);
    out.clear();
    unsigned a, t1, t2;
    while (out.size() < need && fscanf(fp, "%u %u %u", &a, &t1, &t2) == 3) out.push_back(a);
    fclose(fp);
    if (out.size() < 3 + 3 * 1) return fail(OCTMPS_E_PARSE, "safeprimes file too short: " + path);
    return OCTMPS_OK;
}

}  // namespace

octmps_status stream_table(const StreamTable **out) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ready) {
        uint32_t need = 3 + 3 * g_triplets;
        StreamTable t;
        if (!g_file.empty()) {
            octmps_status st = load_file(g_file, need, t.mult);
            if (st != OCTMPS_OK) return st;
            // Triplet counts are powers of two (the device indexes them with a mask).
            uint32_t avail = (uint32_t)((t.mult.size() - 3) / 3), k = 1;
            while (k * 2 <= avail && k * 2 <= g_triplets) k *= 2;
            t.n_triplets = k;
            t.mult.resize(3 + 3 * t.n_triplets);
        } else {
            generate(need, t.mult);
            if (t.mult.size() < need) return fail(OCTMPS_E_RANGE, "safe-prime search exhausted");
            t.n_triplets = g_triplets;
        }
        g_table = t;
        g_ready = true;
    }
    *out = &g_table;
    return OCTMPS_OK;
}

bool seed_valid(uint64_t seed, uint32_t begin) {
    return !((seed == 0ull) | (((uint32_t)(seed >> 32)) >= (begin - 1)) |
             (((uint32_t)seed) >= 0xfffffffful));
}

void photon_streams(const StreamTable &t, uint64_t seed, uint64_t salt, uint64_t photon,
                    uint64_t x[3], uint32_t a[3]) {
    uint64_t h = photon_hash(seed, salt, photon);
    uint32_t k = (uint32_t)(photon % t.n_triplets);
    for (uint32_t s = 0; s < 3; ++s) {
        uint32_t begin = t.mult[s];
        uint32_t am = t.mult[3 + 3 * k + s];
        uint64_t xi = seeding_state(h, s, begin);
        x[s] = derive_state(xi, begin, am);
        a[s] = am;
    }
}

}  // namespace octmps

using namespace octmps;

extern "C" octmps_status octmps_safeprimes(uint32_t n, uint32_t *out) {
    if (!out && n) return fail(OCTMPS_E_INVALID, "null output");
    std::vector<uint32_t> v;
    generate(n, v);
    if (v.size() < n) return fail(OCTMPS_E_RANGE, "safe-prime search exhausted");
    std::memcpy(out, v.data(), n * sizeof(uint32_t));
    return OCTMPS_OK;
}

extern "C" octmps_status octmps_set_safeprimes_file(const char *path) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_file = path ? path : "";
    g_ready = false;
    return OCTMPS_OK;
}

extern "C" octmps_status octmps_set_stream_triplets(uint32_t n) {
    if (n == 0 || n > (1u << 20) || (n & (n - 1)) != 0)
        return fail(OCTMPS_E_RANGE, "triplet count must be a power of two <= 2^20");
    std::lock_guard<std::mutex> lk(g_mu);
    g_triplets = n;
    g_ready = false;
    return OCTMPS_OK;
}

extern "C" octmps_status octmps_photon_streams(uint64_t seed, uint64_t salt, uint64_t photon,
                                               uint64_t x_out[3], uint32_t a_out[3]) {
    const StreamTable *t;
    octmps_status st = stream_table(&t);
    if (st != OCTMPS_OK) return st;
    photon_streams(*t, seed, salt, photon, x_out, a_out);
    return OCTMPS_OK;
}
