// Internal declarations shared by the host-side translation units of liboctmps.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "octmps.h"

namespace octmps {

// Thread-local error detail returned by octmps_last_error().
void set_error(const std::string &msg);
octmps_status fail(octmps_status st, const std::string &msg);

// Serialized mesh kept on the host. Planes are FP64, pre-signed so that n.x + d >= 0 holds
// inside the tet; links hold the neighbour across face j (opposite vertex 3-j) or -1.
struct HostMesh {
    std::vector<double> xyz;      // vertices, 3 per vertex [cm]
    std::vector<int32_t> tv;      // 4 vertex ids per tet (input order)
    std::vector<int32_t> region;  // region id per tet
    std::vector<double> planes;   // 16 per tet: 4 x (nx, ny, nz, d)
    std::vector<int32_t> links;   // 4 per tet
    // Point-location grid over tets that touch the boundary (built on first use).
    bool grid_built = false;
    double gx0 = 0, gy0 = 0, gdx = 1, gdy = 1;
    int gnx = 0, gny = 0;
    std::vector<int64_t> gstart;  // CSR offsets per grid cell
    std::vector<int32_t> gtets;   // tet ids, ascending within a cell
    int32_t max_region = 0;
    int64_t ntets() const { return (int64_t)region.size(); }
};

}  // namespace octmps

struct octmps_mesh {
    octmps::HostMesh m;
};

namespace octmps {

// Builds planes/links from xyz, tv, region (reference plane arithmetic).
octmps_status build_topology(HostMesh &m);

// Multiplier table (generated or loaded) used by the photon streams.
struct StreamTable {
    std::vector<uint32_t> mult;  // [0..2] seeding multipliers, then triplets
    uint32_t n_triplets = 0;
};
// Returns the current table (generated on first use); thread-safe.
octmps_status stream_table(const StreamTable **out);

// Reference seed check of init_RNG (src/octmps_kernel.cuh:86-92) against `begin`.
bool seed_valid(uint64_t seed, uint32_t begin);
// Photon stream initial states (host restatement of the device code).
void photon_streams(const StreamTable &t, uint64_t seed, uint64_t salt, uint64_t photon,
                    uint64_t x[3], uint32_t a[3]);

// Tally layout helpers (see octmps.h).
inline uint64_t tally_words(int32_t b) { return 4ull + 2ull * OCTMPS_TALLY_CLASS_WORDS * (uint64_t)b; }

}  // namespace octmps
