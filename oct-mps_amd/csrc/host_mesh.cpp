// Mesh ingestion: .mesh parsing, face matching, oriented planes, neighbour links, root
// finding, and structured 6-tet meshes for synthetic configurations.
//
// Plane arithmetic follows MeshGraph/SerializeGraph (src/mesh.cpp:86-359) so that the
// FP64 planes equal the reference's d_faces[faces[i]] * signs[i] bit for bit:
//  * a face's normal comes from the first tet (lowest index) holding it, with that tet's
//    vertex order: N = (V0-V1) x (V0-V2) in long double, normalised by a double norm,
//    d = (-N).V0 in double (src/mesh.cpp:178-196);
//  * the face is turned inward for that tet using the iterated-halving point
//    m = (((v0+v1)/2+v2)/2+v3)/2 (src/mesh.cpp:216-245);
//  * each tet's sign is -1 where N.m + d < 0 for its own m (src/mesh.cpp:325-349).
// Face matching is a sort-based join on vertex-id triples: O(T log T) instead of the
// reference's O(F^2) coordinate scan (src/mesh.cpp:51-80), with 64-bit counts instead of
// the reference's short num_tetrahedrons (src/octmps.h:95).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>

#include "host_common.hpp"

namespace octmps {
namespace {

const int kFaceVerts[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};

struct FaceKey {
    int32_t v[3];   // sorted vertex ids
    int32_t tet;
    int32_t face;
    bool operator<(const FaceKey &o) const {
        if (v[0] != o.v[0]) return v[0] < o.v[0];
        if (v[1] != o.v[1]) return v[1] < o.v[1];
        if (v[2] != o.v[2]) return v[2] < o.v[2];
        return tet < o.tet;
    }
    bool same(const FaceKey &o) const { return v[0] == o.v[0] && v[1] == o.v[1] && v[2] == o.v[2]; }
};

void midpoint(const HostMesh &m, int64_t t, double &mx, double &my, double &mz) {
    const int32_t *tv = &m.tv[t * 4];
    mx = m.xyz[3 * tv[0]];
    my = m.xyz[3 * tv[0] + 1];
    mz = m.xyz[3 * tv[0] + 2];
    for (int k = 1; k < 4; ++k) {
        mx += m.xyz[3 * tv[k]];
        my += m.xyz[3 * tv[k] + 1];
        mz += m.xyz[3 * tv[k] + 2];
        mx /= 2;
        my /= 2;
        mz /= 2;
    }
}

}  // namespace

octmps_status build_topology(HostMesh &m) {
    const int64_t T = m.ntets();
    const int64_t V = (int64_t)m.xyz.size() / 3;
    if (T <= 0) return fail(OCTMPS_E_PARSE, "mesh has no tetrahedra");
    if (T > (int64_t)INT32_MAX / 4) return fail(OCTMPS_E_RANGE, "too many tetrahedra");
    for (int64_t i = 0; i < 4 * T; ++i)
        if (m.tv[i] < 0 || m.tv[i] >= V) return fail(OCTMPS_E_PARSE, "tet references a missing vertex");
    m.max_region = 0;
    for (int64_t t = 0; t < T; ++t) {
        if (m.region[t] < 1 || m.region[t] >= OCTMPS_MAX_REGIONS)
            return fail(OCTMPS_E_RANGE, "tet region must be in 1..99");
        m.max_region = std::max(m.max_region, m.region[t]);
    }
    // Faces are matched by vertex coordinates, like CheckFaceExists: vertices with equal
    // coordinates (distinct ids do occur in sample meshes) count as one.
    std::vector<int32_t> canon((size_t)V);
    {
        std::vector<int32_t> ord((size_t)V);
        std::iota(ord.begin(), ord.end(), 0);
        auto key = [&](int32_t v, int q) { return m.xyz[3 * (int64_t)v + q] + 0.0; };  // -0 == +0
        std::sort(ord.begin(), ord.end(), [&](int32_t a, int32_t b) {
            for (int q = 0; q < 3; ++q)
                if (key(a, q) != key(b, q)) return key(a, q) < key(b, q);
            return a < b;
        });
        for (size_t i = 0; i < ord.size(); ++i) {
            bool same = i > 0;
            for (int q = 0; q < 3 && same; ++q) same = key(ord[i], q) == key(ord[i - 1], q);
            canon[ord[i]] = same ? canon[ord[i - 1]] : ord[i];
        }
    }
    std::vector<FaceKey> keys((size_t)(4 * T));
    for (int64_t t = 0; t < T; ++t) {
        for (int j = 0; j < 4; ++j) {
            FaceKey &k = keys[(size_t)(4 * t + j)];
            for (int q = 0; q < 3; ++q) k.v[q] = canon[m.tv[4 * t + kFaceVerts[j][q]]];
            std::sort(k.v, k.v + 3);
            k.tet = (int32_t)t;
            k.face = j;
        }
    }
    std::sort(keys.begin(), keys.end(), [](const FaceKey &a, const FaceKey &b) {
        if (a < b) return true;
        if (b < a) return false;
        return a.face < b.face;
    });
    m.links.assign((size_t)(4 * T), -1);
    m.planes.assign((size_t)(16 * T), 0.0);
    std::vector<double> mids((size_t)(3 * T));
    for (int64_t t = 0; t < T; ++t) midpoint(m, t, mids[3 * t], mids[3 * t + 1], mids[3 * t + 2]);

    // Plane of face (tet, j) from that tet's vertex order, turned inward for that tet. A
    // zero-area face yields NaN like the reference's division by a zero norm.
    auto face_plane = [&](int32_t tet, int j, double out[4]) {
        const int32_t *tv = &m.tv[4 * (int64_t)tet];
        const double *V0 = &m.xyz[3 * tv[kFaceVerts[j][0]]];
        const double *V1 = &m.xyz[3 * tv[kFaceVerts[j][1]]];
        const double *V2 = &m.xyz[3 * tv[kFaceVerts[j][2]]];
        long double x1 = V0[0] - V1[0], y1 = V0[1] - V1[1], z1 = V0[2] - V1[2];
        long double x2 = V0[0] - V2[0], y2 = V0[1] - V2[1], z2 = V0[2] - V2[2];
        long double Nx = y1 * z2 - z1 * y2;
        long double Ny = z1 * x2 - x1 * z2;
        long double Nz = x1 * y2 - y1 * x2;
        double norm = (double)sqrtl(Nx * Nx + Ny * Ny + Nz * Nz);
        Nx /= norm;
        Ny /= norm;
        Nz /= norm;
        double nx = (double)Nx, ny = (double)Ny, nz = (double)Nz;
        double mnx = (double)(-Nx), mny = (double)(-Ny), mnz = (double)(-Nz);
        double d = mnx * V0[0] + mny * V0[1] + mnz * V0[2];
        const double *mc = &mids[3 * (int64_t)tet];
        long double h = nx * mc[0] + ny * mc[1] + nz * mc[2] + d;
        if (h < 0) {
            nx *= -1;
            ny *= -1;
            nz *= -1;
            d *= -1;
        }
        out[0] = nx;
        out[1] = ny;
        out[2] = nz;
        out[3] = d;
    };

    size_t i = 0;
    while (i < keys.size()) {
        size_t j = i + 1;
        while (j < keys.size() && keys[j].same(keys[i])) ++j;
        // Entries i..j-1 share one face, ordered by (tet, face). The first tet creates it
        // (each of its own entries creates a face; later tets reuse the first one). Later
        // tets link to the first tet; the first tet links to the last one
        // (src/mesh.cpp:198-208).
        const int32_t first = keys[i].tet, last = keys[j - 1].tet;
        double base[4];
        face_plane(first, keys[i].face, base);
        for (size_t k = i; k < j; ++k) {
            const FaceKey &e = keys[k];
            double *pl = &m.planes[16 * (int64_t)e.tet + 4 * e.face];
            double f[4];
            if (e.tet == first) face_plane(first, e.face, f);
            else std::copy(base, base + 4, f);
            const double *mt = &mids[3 * (int64_t)e.tet];
            long double h = f[0] * mt[0] + f[1] * mt[1] + f[2] * mt[2] + f[3];
            double s = (h < 0) ? -1.0 : 1.0;
            pl[0] = f[0] * s;
            pl[1] = f[1] * s;
            pl[2] = f[2] * s;
            pl[3] = f[3] * s;
            int32_t link = (e.tet == first) ? (last != first ? last : -1) : first;
            m.links[4 * (int64_t)e.tet + e.face] = link;
        }
        i = j;
    }
    m.grid_built = false;
    return OCTMPS_OK;
}

namespace {

// Point-location grid over tets with at least one vertex on a boundary face (every tet
// that contains a point of the tissue surface has one).
void build_grid(HostMesh &m) {
    const int64_t T = m.ntets();
    const int64_t V = (int64_t)m.xyz.size() / 3;
    std::vector<uint8_t> bvert((size_t)V, 0);
    for (int64_t t = 0; t < T; ++t)
        for (int j = 0; j < 4; ++j)
            if (m.links[4 * t + j] < 0)
                for (int q = 0; q < 3; ++q) bvert[m.tv[4 * t + kFaceVerts[j][q]]] = 1;
    std::vector<int32_t> cand;
    double x0 = 1e300, x1 = -1e300, y0 = 1e300, y1 = -1e300;
    for (int64_t t = 0; t < T; ++t) {
        bool b = false;
        for (int k = 0; k < 4; ++k) b |= bvert[m.tv[4 * t + k]] != 0;
        if (!b) continue;
        cand.push_back((int32_t)t);
        for (int k = 0; k < 4; ++k) {
            const double *p = &m.xyz[3 * m.tv[4 * t + k]];
            x0 = std::min(x0, p[0]);
            x1 = std::max(x1, p[0]);
            y0 = std::min(y0, p[1]);
            y1 = std::max(y1, p[1]);
        }
    }
    int n = std::max(1, (int)std::sqrt((double)cand.size() / 4.0));
    n = std::min(n, 4096);
    m.gnx = m.gny = n;
    m.gx0 = x0;
    m.gy0 = y0;
    m.gdx = std::max((x1 - x0) / n, 1e-12);
    m.gdy = std::max((y1 - y0) / n, 1e-12);
    std::vector<std::vector<int32_t>> cells((size_t)n * n);
    for (int32_t t : cand) {
        double bx0 = 1e300, bx1 = -1e300, by0 = 1e300, by1 = -1e300;
        for (int k = 0; k < 4; ++k) {
            const double *p = &m.xyz[3 * m.tv[4 * (int64_t)t + k]];
            bx0 = std::min(bx0, p[0]);
            bx1 = std::max(bx1, p[0]);
            by0 = std::min(by0, p[1]);
            by1 = std::max(by1, p[1]);
        }
        const double eps = 1e-9;
        int i0 = std::max(0, (int)std::floor((bx0 - eps - x0) / m.gdx));
        int i1 = std::min(n - 1, (int)std::floor((bx1 + eps - x0) / m.gdx));
        int j0 = std::max(0, (int)std::floor((by0 - eps - y0) / m.gdy));
        int j1 = std::min(n - 1, (int)std::floor((by1 + eps - y0) / m.gdy));
        for (int jj = j0; jj <= j1; ++jj)
            for (int ii = i0; ii <= i1; ++ii) cells[(size_t)jj * n + ii].push_back(t);
    }
    m.gstart.assign((size_t)n * n + 1, 0);
    m.gtets.clear();
    for (size_t c = 0; c < cells.size(); ++c) {
        m.gstart[c] = (int64_t)m.gtets.size();
        m.gtets.insert(m.gtets.end(), cells[c].begin(), cells[c].end());
    }
    m.gstart[cells.size()] = (int64_t)m.gtets.size();
    m.grid_built = true;
}

void grid_candidates(const HostMesh &m, double x, double y, std::vector<int32_t> &out) {
    out.clear();
    int i = (int)std::floor((x - m.gx0) / m.gdx), j = (int)std::floor((y - m.gy0) / m.gdy);
    i = std::min(std::max(i, 0), m.gnx - 1);
    j = std::min(std::max(j, 0), m.gny - 1);
    size_t c = (size_t)j * m.gnx + i;
    out.assign(m.gtets.begin() + m.gstart[c], m.gtets.begin() + m.gstart[c + 1]);
    std::sort(out.begin(), out.end());
}

double plane_dist(const HostMesh &m, int64_t t, int j, double x, double y, double z) {
    const double *p = &m.planes[16 * t + 4 * j];
    // Dotproduct(probe, N) + d, src/mesh.cpp:268-272.
    return x * p[0] + y * p[1] + z * p[2] + p[3];
}

}  // namespace
}  // namespace octmps

using namespace octmps;

extern "C" octmps_status octmps_mesh_build(int64_t nv, const double *xyz, int64_t nt,
                                           const int32_t *tets, octmps_mesh **out) {
    if (!out || nv <= 0 || nt <= 0 || !xyz || !tets) return fail(OCTMPS_E_INVALID, "bad mesh arrays");
    octmps_mesh *h = new (std::nothrow) octmps_mesh();
    if (!h) return fail(OCTMPS_E_NOMEM, "host allocation failed");
    try {
        h->m.xyz.assign(xyz, xyz + 3 * nv);
        h->m.tv.resize((size_t)(4 * nt));
        h->m.region.resize((size_t)nt);
        for (int64_t t = 0; t < nt; ++t) {
            for (int k = 0; k < 4; ++k) h->m.tv[4 * t + k] = tets[5 * t + k];
            h->m.region[t] = tets[5 * t + 4];
        }
    } catch (...) {
        delete h;
        return fail(OCTMPS_E_NOMEM, "host allocation failed");
    }
    octmps_status st = build_topology(h->m);
    if (st != OCTMPS_OK) {
        delete h;
        return st;
    }
    *out = h;
    return OCTMPS_OK;
}

extern "C" octmps_status octmps_mesh_read(const char *path, octmps_mesh **out) {
    // .mesh layout read by MeshGraph (src/mesh.cpp:86-160): vertex count, tet count, one
    // "x y z" line per vertex (parsed as float, src/mesh.cpp:136), then one
    // "v1 v2 v3 v4 region" line per tet with 1-based vertex ids. '#' comments and blank
    // lines are skipped like readints/readfloats do.
    if (!path || !out) return fail(OCTMPS_E_INVALID, "null argument");
    FILE *fp = fopen(path, "r");
    if (!fp) return fail(OCTMPS_E_IO, std::string("cannot open ") + path);
    auto next_line = [&](char *buf, int n, float *f, int *iv, bool ints) -> int {
        for (;;) {
            if (!fgets(buf, n, fp)) return -1;
            int k = ints ? sscanf(buf, "%d %d %d %d %d", &iv[0], &iv[1], &iv[2], &iv[3], &iv[4])
                         : sscanf(buf, "%f %f %f %f %f", &f[0], &f[1], &f[2], &f[3], &f[4]);
            if (k > 0) return k;
        }
    };
    char buf[512];
    float f[5];
    int iv[5];
    int64_t nv, nt;
    if (next_line(buf, sizeof buf, f, iv, true) != 1) {
        fclose(fp);
        return fail(OCTMPS_E_PARSE, "error reading number of vertices");
    }
    nv = iv[0];
    if (next_line(buf, sizeof buf, f, iv, true) != 1) {
        fclose(fp);
        return fail(OCTMPS_E_PARSE, "error reading number of tetrahedrons");
    }
    nt = iv[0];
    if (nv <= 0 || nt <= 0) {
        fclose(fp);
        return fail(OCTMPS_E_PARSE, "non-positive vertex or tet count");
    }
    std::vector<double> xyz((size_t)(3 * nv));
    for (int64_t i = 0; i < nv; ++i) {
        if (next_line(buf, sizeof buf, f, iv, false) != 3) {
            fclose(fp);
            return fail(OCTMPS_E_PARSE, "error reading vertices coordinates");
        }
        xyz[3 * i] = f[0];
        xyz[3 * i + 1] = f[1];
        xyz[3 * i + 2] = f[2];
    }
    std::vector<int32_t> tets((size_t)(5 * nt));
    for (int64_t i = 0; i < nt; ++i) {
        if (next_line(buf, sizeof buf, f, iv, true) != 5) {
            fclose(fp);
            return fail(OCTMPS_E_PARSE, "error reading tetrahedron points and regions");
        }
        for (int k = 0; k < 4; ++k) tets[5 * i + k] = iv[k] - 1;
        tets[5 * i + 4] = iv[4];
    }
    fclose(fp);
    return octmps_mesh_build(nv, xyz.data(), nt, tets.data(), out);
}

extern "C" void octmps_mesh_free(octmps_mesh *m) { delete m; }

extern "C" int64_t octmps_mesh_num_tets(const octmps_mesh *m) { return m ? m->m.ntets() : 0; }

extern "C" int64_t octmps_mesh_num_vertices(const octmps_mesh *m) {
    return m ? (int64_t)m->m.xyz.size() / 3 : 0;
}

extern "C" octmps_status octmps_mesh_get_soa(const octmps_mesh *m, double *planes, int32_t *links,
                                             int32_t *regions) {
    if (!m) return fail(OCTMPS_E_INVALID, "null mesh");
    if (planes) std::memcpy(planes, m->m.planes.data(), m->m.planes.size() * sizeof(double));
    if (links) std::memcpy(links, m->m.links.data(), m->m.links.size() * sizeof(int32_t));
    if (regions) std::memcpy(regions, m->m.region.data(), m->m.region.size() * sizeof(int32_t));
    return OCTMPS_OK;
}

extern "C" octmps_status octmps_mesh_write(const octmps_mesh *mh, const char *path) {
    if (!mh || !path) return fail(OCTMPS_E_INVALID, "null argument");
    const HostMesh &m = mh->m;
    FILE *fp = fopen(path, "w");
    if (!fp) return fail(OCTMPS_E_IO, std::string("cannot open ") + path);
    int64_t nv = (int64_t)m.xyz.size() / 3, nt = m.ntets();
    fprintf(fp, "%lld\n%lld\n", (long long)nv, (long long)nt);
    for (int64_t i = 0; i < nv; ++i)
        fprintf(fp, "%.9g %.9g %.9g\n", m.xyz[3 * i], m.xyz[3 * i + 1], m.xyz[3 * i + 2]);
    for (int64_t t = 0; t < nt; ++t)
        fprintf(fp, "%d %d %d %d %d\n", m.tv[4 * t] + 1, m.tv[4 * t + 1] + 1, m.tv[4 * t + 2] + 1,
                m.tv[4 * t + 3] + 1, m.region[t]);
    if (fclose(fp) != 0) return fail(OCTMPS_E_IO, "write failed");
    return OCTMPS_OK;
}

extern "C" octmps_status octmps_mesh_find_root(const octmps_mesh *mh, double x, double y, double z,
                                               int32_t project, int32_t *root, double *z_out) {
    if (!mh || !root) return fail(OCTMPS_E_INVALID, "null argument");
    HostMesh &m = const_cast<octmps_mesh *>(mh)->m;
    if (!m.grid_built) build_grid(m);
    std::vector<int32_t> cand;
    grid_candidates(m, x, y, cand);
    if (project) {
        // Top-most boundary face whose xy projection holds (x, y): the tissue surface.
        double best = 1e300;
        bool found = false;
        for (int32_t t : cand) {
            for (int j = 0; j < 4; ++j) {
                if (m.links[4 * (int64_t)t + j] >= 0) continue;
                const double *p = &m.planes[16 * (int64_t)t + 4 * j];
                if (!(p[2] > 1e-12)) continue;  // inward normal must point into +z (a top face)
                const int32_t *tv = &m.tv[4 * (int64_t)t];
                const double *a = &m.xyz[3 * tv[kFaceVerts[j][0]]];
                const double *b = &m.xyz[3 * tv[kFaceVerts[j][1]]];
                const double *c = &m.xyz[3 * tv[kFaceVerts[j][2]]];
                double det = (b[0] - a[0]) * (c[1] - a[1]) - (c[0] - a[0]) * (b[1] - a[1]);
                if (std::fabs(det) < 1e-300) continue;
                double l1 = ((x - a[0]) * (c[1] - a[1]) - (c[0] - a[0]) * (y - a[1])) / det;
                double l2 = ((b[0] - a[0]) * (y - a[1]) - (x - a[0]) * (b[1] - a[1])) / det;
                const double e = 1e-12;
                if (l1 < -e || l2 < -e || l1 + l2 > 1 + e) continue;
                double zz = -(p[0] * x + p[1] * y + p[3]) / p[2];
                if (zz < best) {
                    best = zz;
                    found = true;
                }
            }
        }
        if (!found) return fail(OCTMPS_E_PROBE, "no tissue surface above/below the probe");
        z = best;
    }
    if (z_out) *z_out = z;
    // 1) The reference rule (FindRootofGraph, src/mesh.cpp:265-283): first tet (index order)
    //    with all plane distances >= 0 and a zero distance; its first zero face must be a
    //    boundary face.
    bool ref_failed = false;
    for (int32_t t : cand) {
        double d[4];
        bool inside = true;
        for (int j = 0; j < 4 && inside; ++j) {
            d[j] = plane_dist(m, t, j, x, y, z);
            if (d[j] < 0) inside = false;
        }
        if (!inside) continue;
        int zf = -1;
        for (int j = 0; j < 4; ++j)
            if (d[j] == 0) {
                zf = j;
                break;
            }
        if (zf < 0) continue;
        if (m.links[4 * (int64_t)t + zf] < 0) {
            *root = t;
            return OCTMPS_OK;
        }
        ref_failed = true;  // the reference would stop with "probe is not placed on the surface"
        break;
    }
    // 2) Tolerant rule: tets containing the probe within eps that own a boundary face through
    //    it, preferring one the launch direction (0,0,1) does not leave at once; lowest index.
    const double eps = 1e-9;
    int32_t fallback = -1;
    for (int32_t t : cand) {
        double d[4];
        bool inside = true;
        for (int j = 0; j < 4 && inside; ++j) {
            d[j] = plane_dist(m, t, j, x, y, z);
            if (d[j] < -eps) inside = false;
        }
        if (!inside) continue;
        bool on_boundary = false, enters = true;
        for (int j = 0; j < 4; ++j) {
            if (std::fabs(d[j]) > eps) continue;
            if (m.links[4 * (int64_t)t + j] < 0) on_boundary = true;
            if (m.planes[16 * (int64_t)t + 4 * j + 2] < 0) enters = false;
        }
        if (!on_boundary) continue;
        if (enters) {
            *root = t;
            return OCTMPS_OK;
        }
        if (fallback < 0) fallback = t;
    }
    if (fallback >= 0) {
        *root = fallback;
        return OCTMPS_OK;
    }
    (void)ref_failed;
    return fail(OCTMPS_E_PROBE, "the probe is not on the tissue surface");
}

// ---------------------------------------------------------------------------------------
// Structured meshes.
// ---------------------------------------------------------------------------------------
namespace {

int region_of(const octmps_grid_spec &s, double cx, double cy, double cz) {
    switch (s.kind) {
        case OCTMPS_GEOM_HOMOGENEOUS: return 1;
        case OCTMPS_GEOM_SPHERE: {
            double dx = cx - s.p[0], dy = cy - s.p[1], dz = cz - s.p[2];
            return (dx * dx + dy * dy + dz * dz <= s.p[3] * s.p[3]) ? 1 : 2;
        }
        case OCTMPS_GEOM_LAYERS: {
            for (int i = 0; i < s.n_layers - 1; ++i)
                if (cz < s.p[i]) return i + 1;
            return s.n_layers;
        }
        case OCTMPS_GEOM_ELLIPSOID2S: {
            double th = s.p[6] * M_PI / 180.0;
            double px = cx - s.p[0], py = cy - s.p[1], pz = cz - s.p[2];
            // ellipsoid frame: rotate by -theta about y
            double qx = px * cos(th) - pz * sin(th);
            double qy = py;
            double qz = px * sin(th) + pz * cos(th);
            double ax = cos(th), az = -sin(th);  // major axis in world frame
            for (int sg = -1; sg <= 1; sg += 2) {
                double sx = s.p[0] + sg * s.p[8] * ax, sz = s.p[2] + sg * s.p[8] * az;
                double dx = cx - sx, dy = cy - s.p[1], dz = cz - sz;
                if (dx * dx + dy * dy + dz * dz <= s.p[7] * s.p[7]) return 3;
            }
            double e = (qx / s.p[3]) * (qx / s.p[3]) + (qy / s.p[4]) * (qy / s.p[4]) +
                       (qz / s.p[5]) * (qz / s.p[5]);
            return e <= 1.0 ? 2 : 1;
        }
        case OCTMPS_GEOM_BUMP: {
            double zs = -s.p[0] * exp(-(cx * cx + cy * cy) / (s.p[1] * s.p[1]));
            double depth = cz - zs;
            if (depth < s.p[2]) return 1;
            if (depth < s.p[3]) return 2;
            return 3;
        }
    }
    return 1;
}

}  // namespace

extern "C" octmps_status octmps_mesh_generate(const octmps_grid_spec *s, octmps_mesh **out) {
    if (!s || !out) return fail(OCTMPS_E_INVALID, "null argument");
    if (s->nx <= 0 || s->ny <= 0 || s->nz <= 0 || !(s->size_x > 0) || !(s->size_y > 0) ||
        !(s->size_z > 0))
        return fail(OCTMPS_E_INVALID, "bad grid spec");
    if (s->kind == OCTMPS_GEOM_LAYERS && (s->n_layers < 1 || s->n_layers > 16))
        return fail(OCTMPS_E_INVALID, "bad layer count");
    const int64_t nx = s->nx, ny = s->ny, nz = s->nz;
    const int64_t NV = (nx + 1) * (ny + 1) * (nz + 1), NT = 6 * nx * ny * nz;
    if (NT > (int64_t)INT32_MAX / 4) return fail(OCTMPS_E_RANGE, "grid too large");
    std::vector<double> xyz((size_t)(3 * NV));
    const double hx = s->size_x / nx, hy = s->size_y / ny;
    auto vid = [&](int64_t i, int64_t j, int64_t k) { return (k * (ny + 1) + j) * (nx + 1) + i; };
    for (int64_t k = 0; k <= nz; ++k)
        for (int64_t j = 0; j <= ny; ++j)
            for (int64_t i = 0; i <= nx; ++i) {
                double x = (double)(2 * i - nx) * 0.5 * hx;
                double y = (double)(2 * j - ny) * 0.5 * hy;
                double z = (double)k * s->size_z / nz;
                if (s->kind == OCTMPS_GEOM_BUMP) {
                    double zs = -s->p[0] * exp(-(x * x + y * y) / (s->p[1] * s->p[1]));
                    z = zs + (double)k * (s->size_z - zs) / nz;
                }
                // Round through float like a .mesh written and read back (src/mesh.cpp:136).
                int64_t v = vid(i, j, k);
                xyz[3 * v] = (float)x;
                xyz[3 * v + 1] = (float)y;
                xyz[3 * v + 2] = (float)z;
            }
    // Six tets per cell along the diagonal (0,0,0)-(1,1,1), one per axis permutation.
    static const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    std::vector<int32_t> tets((size_t)(5 * NT));
    int64_t t = 0;
    for (int64_t k = 0; k < nz; ++k)
        for (int64_t j = 0; j < ny; ++j)
            for (int64_t i = 0; i < nx; ++i)
                for (int p = 0; p < 6; ++p) {
                    int64_t c[3] = {0, 0, 0};
                    int64_t ids[4];
                    ids[0] = vid(i, j, k);
                    for (int q = 0; q < 2; ++q) {
                        c[perm[p][q]] = 1;
                        ids[q + 1] = vid(i + c[0], j + c[1], k + c[2]);
                    }
                    ids[3] = vid(i + 1, j + 1, k + 1);
                    double cx = 0, cy = 0, cz = 0;
                    for (int q = 0; q < 4; ++q) {
                        tets[5 * t + q] = (int32_t)ids[q];
                        cx += xyz[3 * ids[q]];
                        cy += xyz[3 * ids[q] + 1];
                        cz += xyz[3 * ids[q] + 2];
                    }
                    tets[5 * t + 4] = region_of(*s, cx / 4, cy / 4, cz / 4);
                    ++t;
                }
    return octmps_mesh_build(NV, xyz.data(), NT, tets.data(), out);
}
