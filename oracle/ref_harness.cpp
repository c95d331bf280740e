// C wrapper over the reference's own host routines (src/mesh.cpp, src/bias.cpp,
// src/octmps_io.cpp), compiled from the reference sources into oracle/_ref/libref_host.so.
// TEST INFRASTRUCTURE ONLY: used by CPU tests to check the product's .opt/.mci parsers,
// mesh builder and CSV writer against the reference's.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <unistd.h>

#include "octmps.h"  // the reference's header (src/octmps.h), via -I

extern "C" {

// read_simulation_data + Bias_Read: out[0..] = n_regions, then per region (n, mua, mus, g,
// mutr), then photons, probe start/step/end, TypeSimulation, TypeBias, a, TDmin, TDmax, Lc,
// RecordPhotonsFlag, radius, angle, p, MRC, MFOD, bins, shift. Returns the number of values.
int ref_read_inputs(const char *opt, const char *mci, double *out, int cap) {
    SimulationStruct *sim = nullptr;
    int n = read_simulation_data((char *)opt, &sim);
    if (n <= 0) return -1;
    if (!Bias_Read((char *)mci, &sim)) return -2;
    int k = 0;
    auto put = [&](double v) {
        if (k < cap) out[k] = v;
        ++k;
    };
    put(sim->n_regions);
    for (unsigned i = 0; i < sim->n_regions; ++i) {
        put(sim->regions[i].n);
        put(i ? sim->regions[i].mua : 0);
        put(i ? sim->regions[i].mus : 0);
        put(i ? sim->regions[i].g : 0);
        put(i ? sim->regions[i].mutr : 0);
    }
    put(sim->number_of_photons);
    put(sim->probe.start_x);
    put(sim->probe.distance_Ascans);
    put(sim->probe.end_x);
    put(sim->TypeSimulation);
    put(sim->TypeBias);
    put(sim->BackwardBiasCoefficient);
    put(sim->TargetDepthMin);
    put(sim->TargetDepthMax);
    put(sim->CoherenceLengthSource);
    put(sim->RecordPhotonsFlag);
    put(sim->MaxCollectingRadius);
    put(sim->MaxCollectingAngleDeg);
    put(sim->ProbabilityAdditionalBias);
    put(sim->MaxRelativeContributionToBinPerPostProcessedSample);
    put(sim->MaxFilteredOpticalDepth);
    put(sim->NumOpticalDepthLengthSteps);
    put(sim->OpticalDepthShift);
    FreeSimulationStruct(sim, 1);
    return k;
}

// read_simulation_data for run `run` of a multi-run .opt: out[0..] = n_regions, then per
// region (n, mua, mus, g, mutr), photons, probe start/step/end. Returns the number of values.
// read_simulation_data stores run i's probe line through simulations[i] - the i-th pointer
// after the one it fills with the run array (src/octmps_io.cpp:636-640) - so it is given an
// array of pointers whose entries 1.. point to structs that receive those probe lines.
int ref_read_run(const char *opt, int run, double *out, int cap) {
    constexpr int kMaxRuns = 64;
    static SimulationStruct probe_sinks[kMaxRuns];
    SimulationStruct *ptrs[kMaxRuns];
    ptrs[0] = nullptr;
    for (int i = 1; i < kMaxRuns; ++i) ptrs[i] = &probe_sinks[i];
    // The run count is the second data line; refuse files with more runs than sinks.
    int n = read_simulation_data((char *)opt, ptrs);
    if (n <= 0 || n > kMaxRuns) return -1;
    if (run < 0 || run >= n) return -3;
    const SimulationStruct &s = ptrs[0][run];
    const SimulationStruct &pr = run == 0 ? ptrs[0][0] : probe_sinks[run];
    int k = 0;
    auto put = [&](double v) {
        if (k < cap) out[k] = v;
        ++k;
    };
    put(s.n_regions);
    for (unsigned i = 0; i < s.n_regions; ++i) {
        put(s.regions[i].n);
        put(i ? s.regions[i].mua : 0);
        put(i ? s.regions[i].mus : 0);
        put(i ? s.regions[i].g : 0);
        put(i ? s.regions[i].mutr : 0);
    }
    put(s.number_of_photons);
    put(pr.probe.start_x);
    put(pr.probe.distance_Ascans);
    put(pr.probe.end_x);
    FreeSimulationStruct(ptrs[0], n);
    return k;
}

// MeshGraph + FindRootofGraph + SerializeGraph for a probe at (x, 0, 0).
// planes[T*16] = d_faces[faces[f]].N*signs[f], d*signs[f]; links[T*4]; regions[T].
// Returns the root index, or -1 (no root) / -2 (bad mesh).
int ref_serialize_mesh(const char *path, double probe_x, double *planes, int *links, int *regions,
                       int *n_tets_out) {
    int nv = 0, nt = 0;
    if (!read_mesh_param((char *)path, &nv, &nt)) return -2;
    SimulationStruct *sim = (SimulationStruct *)calloc(1, sizeof(SimulationStruct));
    TetrahedronStruct *g = MeshGraph((char *)path, &sim);
    if (!g) return -2;
    sim->probe_x = probe_x;
    sim->probe_y = 0;
    sim->probe_z = 0;
    TetrahedronStruct *root = FindRootofGraph(g, &sim);
    int nf = 4 * nt;
    Vertex *dv = new Vertex[nv];
    TriangleFaces *df = new TriangleFaces[nf];
    TetrahedronStructGPU *dr = new TetrahedronStructGPU[nt];
    int rootIdx = -1;
    if (root) {
        rootIdx = root[0].index;
        SerializeGraph(root, nv, nf, nt, dv, df, dr);
        for (int i = 0; i < nt; ++i) {
            for (int f = 0; f < 4; ++f) {
                const TriangleFaces &F = df[dr[i].faces[f]];
                double s = dr[i].signs[f];
                planes[16 * i + 4 * f + 0] = F.Nx * s;
                planes[16 * i + 4 * f + 1] = F.Ny * s;
                planes[16 * i + 4 * f + 2] = F.Nz * s;
                planes[16 * i + 4 * f + 3] = F.d * s;
                links[4 * i + f] = dr[i].adjTetras[f];
            }
            regions[i] = dr[i].region;
        }
    }
    *n_tets_out = nt;
    delete[] dv;
    delete[] df;
    delete[] dr;
    free(sim);
    return rootIdx;
}

// WriteResult for one A-scan from FP64 tallies (Sum, SumSq, Max per class and bin, counts).
void ref_write_result(const char *dir_cstr, const char *opt, const char *mci, double probe_x,
                      const double *sum1, const double *sq1, const double *max1, const unsigned long long *cnt1,
                      const double *sum2, const double *sq2, const double *max2, const unsigned long long *cnt2,
                      unsigned long long nf, unsigned long long nf1, unsigned long long nf2) {
    SimulationStruct *sim = nullptr;
    read_simulation_data((char *)opt, &sim);
    Bias_Read((char *)mci, &sim);
    sim->probe.current_scanning_x = probe_x;
    int B = sim->NumOpticalDepthLengthSteps;
    SimState s;
    memset(&s, 0, sizeof s);
    s.NumFilteredPhotons = nf;
    s.NumFilteredPhotonsClassI = nf1;
    s.NumFilteredPhotonsClassII = nf2;
    double *a[12];
    for (int i = 0; i < 12; ++i) a[i] = (double *)calloc(B + 2, sizeof(double));
    s.ReflectanceClassI_Sum = a[0];
    s.ReflectanceClassI_SumSq = a[1];
    s.ReflectanceClassI_Max = a[2];
    s.ReflectanceClassII_Sum = a[3];
    s.ReflectanceClassII_SumSq = a[4];
    s.ReflectanceClassII_Max = a[5];
    s.ReflectanceClassI = a[6];
    s.ReflectanceClassII = a[7];
    s.MeanReflectanceClassI_Sum = a[8];
    s.MeanReflectanceClassI_SumSq = a[9];
    s.MeanReflectanceClassII_Sum = a[10];
    s.MeanReflectanceClassII_SumSq = a[11];
    s.NumClassI_PhotonsFilteredInRange = (UINT64 *)calloc(B + 1, 8);
    s.NumClassII_PhotonsFilteredInRange = (UINT64 *)calloc(B + 1, 8);
    for (int b = 0; b < B; ++b) {
        a[0][b] = sum1[b];
        a[1][b] = sq1[b];
        a[2][b] = max1[b];
        a[3][b] = sum2[b];
        a[4][b] = sq2[b];
        a[5][b] = max2[b];
        s.NumClassI_PhotonsFilteredInRange[b] = cnt1[b];
        s.NumClassII_PhotonsFilteredInRange[b] = cnt2[b];
    }
    char cwd[4096];
    if (!getcwd(cwd, sizeof cwd)) cwd[0] = 0;
    if (dir_cstr && dir_cstr[0] && chdir(dir_cstr) != 0) return;
    Write_Bias_Parameters(&s, sim, 0.0f);
    if (sim->output_file_ptr) fclose(sim->output_file_ptr);
    WriteResult(sim, &s, nullptr);
    if (cwd[0] && chdir(cwd) != 0) return;
    for (int i = 0; i < 12; ++i) free(a[i]);
    free(s.NumClassI_PhotonsFilteredInRange);
    free(s.NumClassII_PhotonsFilteredInRange);
    FreeSimulationStruct(sim, 1);
}

}  // extern "C"
