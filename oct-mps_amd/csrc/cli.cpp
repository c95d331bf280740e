// OCT-MPS drop-in command line: OCT-MPS [-S<seed>] [-G<n>] [--batched] opt mesh mci
//
// Mirrors main() of the reference (src/main.cu:850-1056): read the .opt/.mci/.mesh inputs,
// step the probe along x over run 0's probe line from start while x < end
// (src/main.cu:948-951, probe y = z = 0), and at every position run each run of the .opt in
// turn (src/main.cu:1032-1036), appending its rows to the Class I/II CSVs and rewriting
// BiasParam.mco in the working directory. Each run has its own regions and photon count; the
// .mci's bias parameters apply to every run (the reference's Bias_Read fills run 0 only and
// leaves the others unset). -G is accepted and ignored (one GPU). Every A-scan of every run
// uses the same photon streams, as the reference re-seeds with the same seed per position
// and never copies stream state back (src/main.cu:1009); --batched runs each run's A-scans
// in one launch with per-A-scan streams instead.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "octmps.h"

static void usage(const char *prog) {
    printf("\nUsage: %s [-S<seed>] [-G<num GPUs>] [--batched] <input opt file> <input mesh file> <input bias file>\n\n",
           prog);
    printf("  -S: seed for random number generation\n");
    printf("  -G: number of GPUs (accepted for compatibility; one GPU is used)\n");
    printf("  --batched: run the whole B-scan in one persistent launch\n\n");
}

static int check(octmps_status s, const char *what) {
    if (s == OCTMPS_OK) return 0;
    fprintf(stderr, "%s failed: %s (%s)\n", what, octmps_status_string(s), octmps_last_error());
    return 1;
}

int main(int argc, char **argv) {
    unsigned long long seed = OCTMPS_SEED0;
    unsigned int ngpu = 1;
    bool batched = false;
    const char *fopt = nullptr, *fmesh = nullptr, *fbias = nullptr;
    for (int i = 1; i < argc; ++i) {  // interpret_arg, src/octmps_io.cpp:105-145
        char *arg = argv[i];
        if (strcmp(arg, "--batched") == 0) {
            batched = true;
            continue;
        }
        if (arg[0] != '-') {
            if (i + 2 >= argc) break;
            fopt = arg;
            fmesh = argv[++i];
            fbias = argv[++i];
            break;
        }
        ++arg;
        if (sscanf(arg, "S%llu", &seed) == 1) continue;
        if (sscanf(arg, "G%u", &ngpu) == 1) continue;
    }
    if (!fopt || !fmesh || !fbias) {
        usage(argv[0]);
        return 1;
    }
    octmps_params p;
    octmps_params_default(&p);
    int32_t nruns = 0;
    if (check(octmps_read_opt(fopt, 0, &p, &nruns), "read_simulation_data")) return 1;
    if (check(octmps_read_mci(fbias, &p), "Bias_Read")) return 1;
    // Run k: run 0's parameters with run k's regions and photon count.
    std::vector<octmps_params> runs((size_t)nruns, p);
    for (int32_t k = 1; k < nruns; ++k) {
        if (check(octmps_read_opt(fopt, k, &runs[k], nullptr), "read_simulation_data")) return 1;
        runs[k].probe_start_x = p.probe_start_x;
        runs[k].probe_step_x = p.probe_step_x;
        runs[k].probe_end_x = p.probe_end_x;
    }
    octmps_mesh *mesh = nullptr;
    if (check(octmps_mesh_read(fmesh, &mesh), "MeshGraph")) return 1;
    printf("Mesh: %lld vertices, %lld tetrahedra; %d run(s); %d depth bins\n", (long long)octmps_mesh_num_vertices(mesh),
           (long long)octmps_mesh_num_tets(mesh), nruns, p.num_depth_bins);
    std::vector<double> xs;
    std::vector<int32_t> roots;
    for (double x = p.probe_start_x; x < p.probe_end_x; x += p.probe_step_x) {
        int32_t root = -1;
        if (check(octmps_mesh_find_root(mesh, x, 0, 0, 0, &root, nullptr), "FindRootofGraph")) return 1;
        xs.push_back(x);
        roots.push_back(root);
        if (p.probe_step_x <= 0) break;
    }
    octmps_ctx *ctx = nullptr;
    if (check(octmps_ctx_create(0, &ctx), "GPU setup")) return 1;
    int32_t mid = -1;
    if (check(octmps_ctx_upload_mesh(ctx, mesh, &mid), "mesh upload")) return 1;
    const uint64_t words = octmps_tally_words(p.num_depth_bins);
    // tallies[(k * positions + i) * words]: run k at position i
    std::vector<uint64_t> tallies(words * xs.size() * runs.size());
    octmps_run_stats st;
    double total_ms = 0;
    for (size_t k = 0; k < runs.size(); ++k) {
        std::vector<octmps_ascan> scans(xs.size());
        for (size_t i = 0; i < xs.size(); ++i) {
            octmps_ascan &s = scans[i];
            memset(&s, 0, sizeof s);
            s.x = xs[i];
            s.root_tet = roots[i];
            s.n_photons = runs[k].number_of_photons;
            s.stream_salt = batched ? (uint64_t)i : 0;
        }
        uint64_t *out = &tallies[k * xs.size() * words];
        if (batched) {
            octmps_batch_opts o;
            memset(&o, 0, sizeof o);
            o.use_graph = 1;
            if (check(octmps_run_bscan(ctx, mid, &runs[k], (int32_t)scans.size(), scans.data(), seed, &o, out, &st),
                      "B-scan"))
                return 1;
            total_ms += st.wall_ms;
        } else {
            for (size_t i = 0; i < scans.size(); ++i) {
                if (check(octmps_run_ascan(ctx, mid, &runs[k], &scans[i], seed, nullptr, out + i * words, &st),
                          "A-scan"))
                    return 1;
                total_ms += st.wall_ms;
                printf("Run %zu A-scan x=%G: %llu photons, %.3f ms\n", k, xs[i], (unsigned long long)st.photons,
                       st.wall_ms);
            }
        }
    }
    // The reference's output order: position by position, every run at each.
    for (size_t i = 0; i < xs.size(); ++i)
        for (size_t k = 0; k < runs.size(); ++k)
            if (check(octmps_write_ascan_outputs("", &runs[k], xs[i], &tallies[(k * xs.size() + i) * words]),
                      "WriteResult"))
                return 1;
    printf("\n>>>>>>Simulation time: %f (ms)\n", total_ms);
    octmps_ctx_destroy(ctx);
    octmps_mesh_free(mesh);
    return 0;
}
