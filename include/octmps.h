/*
 * octmps.h - C ABI of the MI355X (gfx950) OCT-MPS photon transport library.
 *
 * The reference (SSSherif/OCT-MPS) has no plugin or operator API: its seam is the
 * host<->device contract inside RunGPUi (src/main.cu:43-660), called once per A-scan
 * from DoOneSimulation (src/main.cu:665-848) inside the B-scan loop of main
 * (src/main.cu:948-1051). Every entry point below names the reference routine it
 * replaces. Plain pointers and sizes only; every call returns an octmps_status and
 * never exits the process (the reference aborts through checkCudaErrors/exit(1)).
 *
 * Units follow the reference: lengths in cm, coefficients in 1/cm.
 */
#ifndef OCTMPS_H_
#define OCTMPS_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Constants kept from the reference. */
#define OCTMPS_MAX_REGIONS 100              /* src/octmps_kernel.h:110 (incl. ambient) */
#define OCTMPS_NUM_SUBSTEPS_RESOLUTION 6    /* src/octmps.h:34 (bins per coherence length) */
#define OCTMPS_WEIGHT 1e-4                  /* src/octmps.h:22 roulette threshold */
#define OCTMPS_CHANCE 0.1                   /* src/octmps.h:23 roulette survival */
#define OCTMPS_SEED0 985456376ull           /* src/octmps_kernel.h:7 default -S seed */

typedef enum {
    OCTMPS_OK = 0,
    OCTMPS_E_INVALID = 1,   /* bad argument or inconsistent sizes */
    OCTMPS_E_IO = 2,        /* file could not be opened/written */
    OCTMPS_E_PARSE = 3,     /* malformed .opt/.mci/.mesh/safeprimes content */
    OCTMPS_E_PROBE = 4,     /* probe not on the tissue surface (src/mesh.cpp:281-288) */
    OCTMPS_E_RANGE = 5,     /* regions/bins/tets beyond supported limits */
    OCTMPS_E_SEED = 6,      /* invalid MWC seed (src/octmps_kernel.cuh:86-92) */
    OCTMPS_E_NOMEM = 7,     /* host/device allocation failed or HBM budget exceeded */
    OCTMPS_E_DEVICE = 8,    /* HIP runtime error or no GPU */
    OCTMPS_E_BUSY = 9,      /* server queue full */
    OCTMPS_E_CANCELLED = 10,/* request cancelled */
    OCTMPS_E_NOT_FOUND = 11,/* unknown mesh or request id */
    OCTMPS_E_TIMEOUT = 12,  /* wait timed out */
    OCTMPS_E_FAILED = 13    /* request failed (device fault injected or real) */
} octmps_status;

const char *octmps_status_string(int status);
/* Detail of the last error raised on the calling thread ("" if none). */
const char *octmps_last_error(void);
/* Library version string. */
const char *octmps_version(void);

/* ------------------------------------------------------------------------- */
/* Simulation parameters: SimulationStruct (src/octmps.h:77-128) minus the     */
/* mesh, plus the values Bias_Read derives (src/bias.cpp:92-104).              */
/* ------------------------------------------------------------------------- */
typedef struct {
    double n;    /* refractive index */
    double mua;  /* absorption [1/cm] */
    double mus;  /* scattering [1/cm]; 0 marks a glass region (src/octmps_io.cpp:625) */
    double g;    /* anisotropy */
} octmps_region;

typedef struct {
    uint32_t n_regions;                        /* including ambient region 0 */
    octmps_region regions[OCTMPS_MAX_REGIONS]; /* [0].n = ambient index; others used by tets */
    int32_t type_simulation;                   /* .mci line 1 */
    int32_t type_bias;                         /* .mci line 2: 0 = HG only, 37 = biased */
    double backward_bias_coefficient;          /* a */
    double target_depth_min;                   /* [cm] */
    double target_depth_max;                   /* [cm] */
    double coherence_length_source;            /* Lc [cm] */
    int32_t record_photons_flag;
    double max_collecting_radius;              /* [cm] */
    double max_collecting_angle_deg;           /* [deg] */
    double probability_additional_bias;        /* p */
    double max_relative_contribution;          /* outlier filter of the *_Filt outputs */
    double rnd_step_size_in_tissue;            /* never set by the reference; 0 here */
    double max_filtered_optical_depth;         /* derived, src/bias.cpp:92-96 */
    int32_t num_depth_bins;                    /* NumOpticalDepthLengthSteps */
    double optical_depth_shift;                /* fmod(TargetDepthMin, Lc) */
    uint64_t number_of_photons;                /* per A-scan (.opt) */
    double probe_start_x;                      /* B-scan line (.opt ProbePositionData) */
    double probe_step_x;
    double probe_end_x;
} octmps_params;

/* Parse the .opt file (read_simulation_data, src/octmps_io.cpp:504-645). Fills the
 * regions, photon count and probe line of run `run_index`; *n_runs gets the run count.
 * Other fields are left untouched. */
octmps_status octmps_read_opt(const char *path, int32_t run_index, octmps_params *p,
                              int32_t *n_runs);
/* Parse the .mci bias file (Bias_Read, src/bias.cpp:9-110), keeping its float parse and
 * short truncation of the bin count, and derive bins and shift. */
octmps_status octmps_read_mci(const char *path, octmps_params *p);
/* Fill defaults: ambient n=1, TypeBias 37, a=0.925, p=0.5, radius 10e-4 cm, 5 deg,
 * MaxRelContrib 0.01, rndStepSizeInTissue 0. */
void octmps_params_default(octmps_params *p);
/* Specular reflectance at launch (Rspecular, src/main.cu:29-37). */
double octmps_rspecular(double n_ambient, double n_tissue);

/* ------------------------------------------------------------------------- */
/* Host mesh (MeshGraph / FindRootofGraph / SerializeGraph, src/mesh.cpp).    */
/* A tet's face j is the face opposite vertex 3-j (three_combinations order,  */
/* src/mesh.cpp:164-167); planes are pre-signed so that n.x + d >= 0 inside.  */
/* ------------------------------------------------------------------------- */
typedef struct octmps_mesh octmps_mesh;

/* Read a Netgen-style .mesh file (MeshGraph, src/mesh.cpp:86-250): vertices parsed as
 * float like the reference, 1-based vertex ids, region per tet. Face matching is a hash
 * join (the reference's CheckFaceExists is O(F^2)). */
octmps_status octmps_mesh_read(const char *path, octmps_mesh **out);
/* Build from arrays: xyz[n_vertices*3] (cm), tets[n_tets*5] = 4 zero-based vertex ids +
 * region id. */
octmps_status octmps_mesh_build(int64_t n_vertices, const double *xyz, int64_t n_tets,
                                const int32_t *tets, octmps_mesh **out);

/* Structured tetrahedral meshes: nx*ny*nz cells, each split into 6 tets sharing the cell
 * diagonal; x,y centred on 0, tissue surface at z=0 (z grows into tissue). */
typedef enum {
    OCTMPS_GEOM_HOMOGENEOUS = 0, /* region 1 everywhere */
    OCTMPS_GEOM_SPHERE = 1,      /* p[0..2] centre, p[3] radius: region 1 inside, 2 outside */
    OCTMPS_GEOM_LAYERS = 2,      /* p[0..n-1] layer bottoms (cm), n = n_layers; regions 1.. */
    OCTMPS_GEOM_ELLIPSOID2S = 3, /* p[0..2] centre, p[3..5] semi-axes, p[6] tilt about y (deg),
                                    p[7] sphere radius, p[8] sphere offset along major axis:
                                    background 1, ellipsoid 2, spheres 3 */
    OCTMPS_GEOM_BUMP = 4         /* p[0] bump height, p[1] bump radius, p[2..3] layer depths
                                    below the surface: surface z = -h exp(-r^2/rb^2); layers 1..3 */
} octmps_geom_kind;

typedef struct {
    int32_t kind;      /* octmps_geom_kind */
    int32_t nx, ny, nz;
    double size_x, size_y, size_z; /* cm */
    int32_t n_layers;
    double p[16];
} octmps_grid_spec;

octmps_status octmps_mesh_generate(const octmps_grid_spec *spec, octmps_mesh **out);
void octmps_mesh_free(octmps_mesh *m);
int64_t octmps_mesh_num_tets(const octmps_mesh *m);
int64_t octmps_mesh_num_vertices(const octmps_mesh *m);
/* Copy out the serialized mesh. planes[T*16] = 4 x (nx,ny,nz,d) FP64 pre-signed inward
 * planes (reference d_faces[faces[i]] times signs[i], src/mesh.cpp:335-349); links[T*4]
 * neighbour tet or -1 (adjTetras, src/mesh.cpp:351-357); regions[T]. NULL skips. */
octmps_status octmps_mesh_get_soa(const octmps_mesh *m, double *planes, int32_t *links,
                                  int32_t *regions);
/* Root tet for a probe on the tissue surface (FindRootofGraph, src/mesh.cpp:257-289, with
 * a tolerance and a deterministic tie-break). With project_to_surface != 0 the probe z is
 * moved to the surface above/below (x,y) first; *z_out gets the z used. */
octmps_status octmps_mesh_find_root(const octmps_mesh *m, double x, double y, double z,
                                    int32_t project_to_surface, int32_t *root_tet,
                                    double *z_out);
/* Write the mesh in the reference .mesh text format (1-based vertex ids). */
octmps_status octmps_mesh_write(const octmps_mesh *m, const char *path);

/* ------------------------------------------------------------------------- */
/* RNG: multiply-with-carry, base 2^32 (rand_MWC_co, src/octmps_kernel.cuh:10-29) */
/* ------------------------------------------------------------------------- */
/* First n multipliers a (descending, a < 2^32) with a*2^32-1 and a*2^31-1 prime: the
 * table init_RNG reads from safeprimes_base32.txt (src/octmps_kernel.cuh:52-172), which
 * the reference does not ship. */
octmps_status octmps_safeprimes(uint32_t n, uint32_t *out);
/* Use a safeprimes_base32.txt (reference format, first column) instead of the generated
 * table for contexts created afterwards. NULL restores the generated table. */
octmps_status octmps_set_safeprimes_file(const char *path);
/* Number of multiplier triplets used for photon streams (default 4096). */
octmps_status octmps_set_stream_triplets(uint32_t n_triplets);
/* Initial (x, a) of the three streams (General, ReflectT, Spin) of one photon: hashed
 * seed for the photon, then init_RNG's derivation loop (src/octmps_kernel.cuh:109-167). */
octmps_status octmps_photon_streams(uint64_t seed, uint64_t salt, uint64_t photon,
                                    uint64_t x_out[3], uint32_t a_out[3]);

/* ------------------------------------------------------------------------- */
/* Tallies: SimState (src/octmps.h:133-176) as exact integers.                */
/* One A-scan tally is octmps_tally_words(B) = 4 + 2*29*B uint64 words:       */
/*   [0] NumFilteredPhotons [1] ...ClassI [2] ...ClassII                      */
/*   [3] contributions this tally could not hold (not finite or negative; 0)  */
/*   then for class c in {0: Class I, 1: Class II}, base = 4 + c*29*B:         */
/*   base + k*B,      k < 9:  Sum digit k   Sum   = sum_k S_k * 2^(32k - 160)  */
/*   base + (9+k)*B,  k < 18: SumSq digit k SumSq = sum_k Q_k * 2^(32k - 320)  */
/*   base + 27*B count (NumClass*_PhotonsFilteredInRange)                      */
/*   base + 28*B max   (IEEE-754 binary64 bits of Max)                          */
/* A contribution c = m * 2^(e-150) (FP32 fields) adds m << (e+10) to Sum and   */
/* m*m << (2e+20) to SumSq, cut into 32-bit pieces, each added to its digit     */
/* word: every positive finite FP32 c and its square are held exactly, with no  */
/* carry between words, for up to 2^32 contributions per bin.                   */
/* ------------------------------------------------------------------------- */
#define OCTMPS_TALLY_SUM_DIGITS 9
#define OCTMPS_TALLY_SQ_DIGITS 18
#define OCTMPS_TALLY_CLASS_WORDS 29   /* per bin and class: 9 + 18 digits, count, max */
#define OCTMPS_TALLY_SUM_EXP (-160)   /* Sum digit 0 counts units of 2^-160 */
#define OCTMPS_TALLY_SQ_EXP (-320)    /* SumSq digit 0 counts units of 2^-320 */
uint64_t octmps_tally_words(int32_t num_bins);

/* Mean and standard error per bin, raw and filtered
 * (WriteReflectanceClassI_and_ClassII[_Filtered], src/octmps_io.cpp:302-420).
 * out[((f*2 + c)*2 + k)*B + b]: f=0 raw / 1 filtered, c = class, k=0 mean / 1 SE. */
octmps_status octmps_tally_finalize(const uint64_t *tally, int32_t num_bins,
                                    uint64_t n_photons, double max_rel_contrib,
                                    double *out);
/* Append one A-scan to ClassI_Scatt.out, ClassII_Scatt.out, ClassI_ScattFilt.out,
 * ClassII_ScattFilt.out in `dir` and rewrite BiasParam.mco (WriteResult and
 * Write_Bias_Parameters, src/octmps_io.cpp:237-437). */
octmps_status octmps_write_ascan_outputs(const char *dir, const octmps_params *p,
                                         double probe_x, const uint64_t *tally);

/* ------------------------------------------------------------------------- */
/* Device context: one GPU, meshes resident in HBM.                           */
/* ------------------------------------------------------------------------- */
typedef struct octmps_ctx octmps_ctx;

octmps_status octmps_ctx_create(int32_t device, octmps_ctx **out);
void octmps_ctx_destroy(octmps_ctx *ctx);
/* Upload (replaces RunGPUi's per-A-scan cudaMalloc+H2D of the mesh, src/main.cu:100-465):
 * the mesh stays resident until released. All resident meshes share one pool of at most
 * 2^25 - 1 tets (neighbour ids are packed with the neighbour's region), and regions must
 * be < 100 (OCTMPS_E_RANGE otherwise). */
octmps_status octmps_ctx_upload_mesh(octmps_ctx *ctx, const octmps_mesh *m, int32_t *mesh_id);
/* Register a mesh without uploading it (no counterpart in the reference): its 128-byte tet
 * records are kept in host memory and made resident on demand - by a run on it, or by a
 * request server when it admits a request on it - and a registered mesh that nothing uses
 * may be evicted from the device again when a server needs the room (it is then uploaded
 * again when next used). Results do not depend on where or how often it was uploaded. */
octmps_status octmps_ctx_register_mesh(octmps_ctx *ctx, const octmps_mesh *m, int32_t *mesh_id);
/* Release a mesh (src/main.cu:638's cudaFree): OCTMPS_E_BUSY while a submitted request or a
 * run still uses it. */
octmps_status octmps_ctx_release_mesh(octmps_ctx *ctx, int32_t mesh_id);
/* Whether a mesh's records are on the device now, and their size in bytes. */
octmps_status octmps_ctx_mesh_info(octmps_ctx *ctx, int32_t mesh_id, int32_t *resident,
                                   uint64_t *device_bytes);
/* Device bytes in use (meshes + work buffers) and HBM total. */
octmps_status octmps_ctx_memory(octmps_ctx *ctx, uint64_t *used, uint64_t *total);

typedef struct {
    double x, y, z;          /* probe [cm] */
    int32_t root_tet;        /* tet holding the probe (octmps_mesh_find_root) */
    int32_t reserved;
    uint64_t n_photons;      /* photon budget of this A-scan */
    uint64_t stream_salt;    /* photon streams derive from (seed, salt, photon index) */
    uint64_t photon_offset;  /* index of the first photon (run a sub-range of a scan) */
} octmps_ascan;

typedef struct {
    uint64_t photons;         /* photons completed */
    uint64_t photon_steps;    /* transport-loop iterations of live photons */
    uint64_t killed;          /* photons stopped by the per-photon step guard */
    double kernel_ms;         /* HIP-event time of the transport kernels */
    double wall_ms;           /* host wall time of the call */
    uint32_t launches;
    uint32_t block_size;
    uint32_t grid_size;
    uint32_t refill_quantum;
    uint64_t wave_iterations; /* loop iterations summed over waves; lane utilisation =
                                 photon_steps / (64 * wave_iterations) */
    uint64_t bad_contributions; /* detector contributions the tally could not hold exactly
                                   (expected 0; see the tally layout above) */
    uint32_t record_fetch;    /* record path the launch ran: 1 per-lane gathers, 2 cooperative */
    uint32_t entry_skip;      /* 1 when the per-lane gathers skipped the entry plane */
} octmps_run_stats;

typedef struct {
    int32_t block_size;      /* 0 = autotuned/default */
    int32_t refill_quantum;  /* dead lanes per wave before a refill; 0 = default */
    int32_t blocks_per_cu;   /* 0 = occupancy maximum */
    int32_t use_graph;       /* capture the batched launch in a hipGraph */
    int32_t autotune;        /* pick block size / quantum on a calibration slice */
    int32_t scatter_quantum; /* lanes of a wave parked for scattering before it runs; 0 = default */
    int32_t tally_copies;    /* tally privatisation depth: private copies of every A-scan's
                                tally (workgroup w adds into copy w mod k), folded into one
                                after the launch; 1..16, 0 = autotuned/default (1) */
    int32_t entry_skip;      /* per-lane record fetch only: a photon that crossed into a tet
                                through face k does not fetch plane k: 0 = auto (on when the
                                mesh's 128-byte tet records fit the 256 MB Infinity Cache),
                                1 = on, 2 = off. Results are the same bits either way. */
    uint64_t max_steps_per_photon; /* 0 = default guard (1<<22) */
    int32_t record_fetch;    /* how a wave reads the tet records of its crossing photons:
                                1 = each lane gathers its own record (five 16-byte loads);
                                2 = cooperatively: four lanes read the four face planes of one
                                record together into LDS (one L1 cycle instead of four), and a
                                record stays staged while its photon stays in the tet;
                                0 = auto (2 when its launch holds as many waves per CU as 1,
                                which needs 256-lane blocks). Results are the same bits. */
    int32_t reserved0;       /* 0 */
} octmps_batch_opts;

/* Plain path: one A-scan per launch sequence, synchronous, tallies copied back
 * (RunGPUi's launch loop + CopyDeviceToHostMem, src/main.cu:534-636). tally_out has
 * octmps_tally_words(params->num_depth_bins) words. */
octmps_status octmps_run_ascan(octmps_ctx *ctx, int32_t mesh_id, const octmps_params *p,
                               const octmps_ascan *scan, uint64_t seed,
                               const octmps_batch_opts *opts, uint64_t *tally_out,
                               octmps_run_stats *stats);
/* Batched path: all A-scans of a B-scan in one persistent launch (hipGraph-captured when
 * opts->use_graph); lanes refill across A-scans. Replaces the host B-scan loop
 * (src/main.cu:948-1051). tallies_out has n_scans * octmps_tally_words(B) words. */
octmps_status octmps_run_bscan(octmps_ctx *ctx, int32_t mesh_id, const octmps_params *p,
                               int32_t n_scans, const octmps_ascan *scans, uint64_t seed,
                               const octmps_batch_opts *opts, uint64_t *tallies_out,
                               octmps_run_stats *stats);
/* octmps_run_bscan that allocates no device memory: every device buffer of the run comes
 * from one file-scope device array of 64 MiB per GPU (held until the process ends), so it
 * calls no hipMalloc/hipFree. No counterpart in the reference: it is for callers that cannot
 * allocate device memory (tight serving loops, harnesses that forbid allocation). Same
 * arguments and the same tallies, bit for bit, as octmps_run_bscan; always launched eagerly
 * (opts->use_graph is ignored). A run needing more than 64 MiB of scratch (per-lane state
 * ~19 MB plus n_scans * tally_copies * octmps_tally_words(B) * 8 bytes, e.g. 128 A-scans at
 * 512 bins with one tally copy: ~50 MB) returns OCTMPS_E_RANGE. Calls on one GPU share the
 * array: calls that overlap on one GPU (two contexts on the same device) race. */
octmps_status octmps_run_bscan_noalloc(octmps_ctx *ctx, int32_t mesh_id, const octmps_params *p,
                                       int32_t n_scans, const octmps_ascan *scans, uint64_t seed,
                                       const octmps_batch_opts *opts, uint64_t *tallies_out,
                                       octmps_run_stats *stats);
/* octmps_run_bscan that also writes one 16-byte record per photon, in launch order
 * (photon ids of scan 0 first): {u32 steps, u32 detections, f32 summed contribution,
 * u32 killed}. Diagnostic path used to pair GPU photons with the CPU oracle's; no
 * counterpart in the reference. log_records must be >= the total photon count. */
octmps_status octmps_run_bscan_logged(octmps_ctx *ctx, int32_t mesh_id, const octmps_params *p,
                                      int32_t n_scans, const octmps_ascan *scans, uint64_t seed,
                                      const octmps_batch_opts *opts, uint64_t *tallies_out,
                                      octmps_run_stats *stats, void *photon_log,
                                      uint64_t log_records);
/* Autotune occupancy (blocks per CU), scatter quantum, block size, refill quantum, tally
 * privatisation depth and the record path on a calibration slice of a B-scan (coordinate
 * search from the context defaults, 21 launches: a candidate must be more than 1 % faster to
 * win, and the winner must lead the starting point again on a second measurement of both);
 * writes the winner into *opts (blocks_per_cu, scatter_quantum, block_size, refill_quantum,
 * tally_copies, entry_skip, record_fetch), makes block size, quanta and tally copies the
 * context's default (occupancy and the record path stay per launch, since their best values
 * depend on the mesh: meshes larger than the caches run faster with fewer photons in
 * flight), and returns its
 * photons/s in *rate. The slice is the first calib_photons / n_scans photons of every A-scan,
 * raised to at least 8 photons per lane of a full-occupancy grid (about 3.1e6 on 256 CUs)
 * where the B-scan holds that many: a smaller slice is mostly the launch's tail. */
octmps_status octmps_autotune(octmps_ctx *ctx, int32_t mesh_id, const octmps_params *p,
                              int32_t n_scans, const octmps_ascan *scans, uint64_t seed,
                              uint64_t calib_photons, octmps_batch_opts *opts, double *rate);

/* ------------------------------------------------------------------------- */
/* Serving: continuous batching of A-/B-scan requests.                        */
/* ------------------------------------------------------------------------- */
typedef struct octmps_server octmps_server;

typedef struct {
    int32_t max_inflight_scans;  /* scan slots on the device (0 = 32768) */
    int32_t block_size;          /* 0 = default */
    int32_t refill_quantum;      /* 0 = default */
    int32_t fail_request_index;  /* test hook: submit index to fail (-1 = none) */
    uint64_t hbm_budget_bytes;   /* admission budget for the requests' HBM footprint: each
                                    admitted request's tallies, scan slots and parameter set,
                                    plus the device records of every mesh the server keeps
                                    resident for them (0 = 3/4 of the device's memory). The
                                    server's own per-lane state (~176 B per lane) is extra. */
    uint64_t slice_photons;      /* photon ids handed out per slice (0 = auto) */
    int32_t policy;              /* order of photon hand-out among requests in flight:
                                    0 = least remaining photons first, aged by waiting time
                                    (halved after 10 s, a third after 20 s, ...);
                                    1 = equal shares per A-scan. Results are the same bits. */
    int32_t record_fetch;        /* record path of the slices: 0 = auto (cooperative, with the
                                    parameter sets in flight in its LDS table when they fit
                                    it: 16 sets of up to 8 regions or 8 of up to 16, read
                                    from memory when they fit no LDS table; per-lane gathers
                                    with the sets in LDS when they fit only the per-lane
                                    table of 16 sets of 16 regions), 1 = per-lane gathers
                                    always, 2 = the same as 0.
                                    Results are the same bits. */
} octmps_serve_opts;

typedef struct {
    int32_t mesh_id;
    int32_t n_scans;
    const octmps_params *params;   /* copied at submit */
    const octmps_ascan *scans;     /* copied at submit */
    uint64_t seed;
    uint64_t *tallies_out;         /* n_scans * octmps_tally_words(B); written by the server's
                                      completion thread before the request is reported done,
                                      so it must stay valid until then (or until close) */
} octmps_request;

typedef struct {
    int32_t status;          /* octmps_status of the request */
    int32_t done;
    double t_submit_ms, t_admit_ms, t_done_ms; /* server clock, ms */
    uint64_t photons;
} octmps_request_info;

octmps_status octmps_serve_open(octmps_ctx *ctx, const octmps_serve_opts *opts,
                                octmps_server **out);
octmps_status octmps_serve_submit(octmps_server *s, const octmps_request *req,
                                  uint64_t *req_id);
octmps_status octmps_serve_poll(octmps_server *s, uint64_t req_id, octmps_request_info *info);
octmps_status octmps_serve_wait(octmps_server *s, uint64_t req_id, double timeout_s,
                                octmps_request_info *info);
octmps_status octmps_serve_cancel(octmps_server *s, uint64_t req_id);
/* Forget a finished request (its id becomes unknown and its record is freed), so a
 * long-running server does not keep every request it has served. OCTMPS_E_BUSY while the
 * request is still running. No counterpart in the reference. */
octmps_status octmps_serve_release(octmps_server *s, uint64_t req_id);
/* Mesh uploads and evictions the server made since open, and the footprint bytes admitted
 * now against its budget. */
octmps_status octmps_serve_residency(octmps_server *s, uint64_t *uploads, uint64_t *evictions,
                                     uint64_t *used_bytes, uint64_t *budget_bytes);
/* Aggregate counters since open: photons, photon-steps, busy kernel ms, slices. */
octmps_status octmps_serve_stats(octmps_server *s, uint64_t *photons, uint64_t *steps,
                                 double *kernel_ms, uint64_t *slices);
octmps_status octmps_serve_close(octmps_server *s);

/* ------------------------------------------------------------------------- */
/* Device unit probes (parity of single device functions against the oracle). */
/* ------------------------------------------------------------------------- */
/* n_draws co-draws of each of the 3 streams of photons [photon0, photon0+n_photons):
 * out[(p*3 + s)*n_draws + k]. Device rand_MWC_co (src/octmps_kernel.cuh:10-22) on the
 * per-photon streams. */
octmps_status octmps_probe_rng(octmps_ctx *ctx, uint64_t seed, uint64_t salt, uint64_t photon0,
                               int32_t n_photons, int32_t n_draws, float *out);
/* Fresnel reflectance of the device code for n cases (ni, nt, ca1 = cos incidence): the
 * MCML formula of FastReflectTransmit (src/octmps_kernel.cuh:601-646). */
octmps_status octmps_probe_fresnel(octmps_ctx *ctx, int32_t n, const float *ni, const float *nt,
                                   const float *ca1, float *r_out);
/* cos(theta) from the HG sampler (kind 0, parameter g; SpinTheta, src/octmps_kernel.cuh:807-820)
 * or the biased forward sampler (kind 1, coefficient a; SpinThetaForwardFstBias, :881-896)
 * for uniform draws u in (0,1]. */
octmps_status octmps_probe_sampler(octmps_ctx *ctx, int32_t kind, float param, int32_t n,
                                   const float *u, float *cost_out);

/* The device's detector tally (FastReflectTransmit's tally, src/octmps_kernel.cuh:684-741)
 * run on n chosen detections (optical path op[i], max depth maxz[i], contribution
 * contrib[i]) into one zeroed A-scan tally of p's bins: tally_out has
 * octmps_tally_words(p->num_depth_bins) words; *bad_out (may be NULL) gets the number of
 * contributions the tally could not hold. */
octmps_status octmps_probe_tally(octmps_ctx *ctx, const octmps_params *p, int32_t n, const float *op,
                                 const float *maxz, const float *contrib, uint64_t *tally_out,
                                 uint64_t *bad_out);

/* Device-to-device copy of `bytes` through a plain streaming kernel (no counterpart in the
 * reference): the calibration load for hardware byte counters. Returns the kernel time. */
octmps_status octmps_probe_copy(octmps_ctx *ctx, uint64_t bytes, int32_t repeats, double *ms);

/* Divergent gathers of 5 x 16 B from pseudo-random 128-byte records of an n_records set
 * that stays in cache, 6 waves per SIMD (no counterpart in the reference): the ceiling of
 * the vector-memory path for the access pattern of a photon-step's tet visit. Returns 16-byte
 * lane-loads per second. n_records in 1..2^24. */
octmps_status octmps_probe_gather(octmps_ctx *ctx, int32_t n_records, int32_t iters, double *lane_loads_per_s);

/* L1 cost of a 16-byte vector load by access pattern (no counterpart in the reference): every
 * lane of a 6-waves-per-SIMD grid issues iters x 8 loads from a 16 KiB (L1-resident) buffer.
 * pattern 0: divergent (the 64 lanes of a load hit 64 cache lines); 1: quad-coalesced (each
 * four lanes read the four quarters of one 64-byte piece); 2: divergent, every other lane
 * masked off; 3: divergent, every other lane out of range of the buffer resource; 4: pattern
 * 1 as DMA loads into LDS (the cooperative record fetch). Returns wave-level load
 * instructions per second over the GPU and CU cycles per load at the nominal clock. */
octmps_status octmps_probe_l1(octmps_ctx *ctx, int32_t pattern, int32_t iters, double *loads_per_s,
                              double *cu_cycles_per_load);

/* VALU issue rate by instruction class (no counterpart in the reference): every lane of a
 * 6-waves-per-SIMD grid runs iters x 8 independent operations of one class. kind 0: FP32
 * fused multiply-add; 1: FP32 reciprocal (the transcendental class); 2: an MWC step (64-bit
 * integer multiply-add and a move: two instructions, counted as one). Returns wave-level instructions per second over the GPU and
 * SIMD cycles per instruction at the nominal clock: the denominators of the transport
 * kernel's VALU issue fraction. */
octmps_status octmps_probe_valu(octmps_ctx *ctx, int32_t kind, int32_t iters, double *insts_per_s,
                                double *simd_cycles_per_inst);

#ifdef __cplusplus
}
#endif
#endif /* OCTMPS_H_ */
