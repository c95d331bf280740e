"""ctypes binding of liboctmps (include/octmps.h) for tests, benchmarks and scripting.

The product is the C ABI in ``lib/liboctmps.so``; this module only declares its structs and
functions and adds numpy views of the integer tallies. Importing it raises if the library
has not been built: there is no Python fallback for any computation.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
LIB_PATH = os.path.join(PKG, "lib", "liboctmps.so")
CLI_PATH = os.path.join(PKG, "bin", "OCT-MPS")

MAX_REGIONS = 100
SUBSTEPS = 6
SEED0 = 985456376

OK = 0
STATUS = {
    0: "OK", 1: "E_INVALID", 2: "E_IO", 3: "E_PARSE", 4: "E_PROBE", 5: "E_RANGE", 6: "E_SEED",
    7: "E_NOMEM", 8: "E_DEVICE", 9: "E_BUSY", 10: "E_CANCELLED", 11: "E_NOT_FOUND",
    12: "E_TIMEOUT", 13: "E_FAILED",
}


class OctmpsError(RuntimeError):
    def __init__(self, status, where, detail=""):
        self.status = status
        super().__init__(f"{where}: {STATUS.get(status, status)} {detail}".strip())


class Region(C.Structure):
    _fields_ = [("n", C.c_double), ("mua", C.c_double), ("mus", C.c_double), ("g", C.c_double)]


class Params(C.Structure):
    _fields_ = [
        ("n_regions", C.c_uint32),
        ("regions", Region * MAX_REGIONS),
        ("type_simulation", C.c_int32),
        ("type_bias", C.c_int32),
        ("backward_bias_coefficient", C.c_double),
        ("target_depth_min", C.c_double),
        ("target_depth_max", C.c_double),
        ("coherence_length_source", C.c_double),
        ("record_photons_flag", C.c_int32),
        ("max_collecting_radius", C.c_double),
        ("max_collecting_angle_deg", C.c_double),
        ("probability_additional_bias", C.c_double),
        ("max_relative_contribution", C.c_double),
        ("rnd_step_size_in_tissue", C.c_double),
        ("max_filtered_optical_depth", C.c_double),
        ("num_depth_bins", C.c_int32),
        ("optical_depth_shift", C.c_double),
        ("number_of_photons", C.c_uint64),
        ("probe_start_x", C.c_double),
        ("probe_step_x", C.c_double),
        ("probe_end_x", C.c_double),
    ]


class GridSpec(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("size_x", C.c_double), ("size_y", C.c_double), ("size_z", C.c_double),
        ("n_layers", C.c_int32), ("p", C.c_double * 16),
    ]


GEOM_HOMOGENEOUS, GEOM_SPHERE, GEOM_LAYERS, GEOM_ELLIPSOID2S, GEOM_BUMP = range(5)


class AScan(C.Structure):
    _fields_ = [
        ("x", C.c_double), ("y", C.c_double), ("z", C.c_double),
        ("root_tet", C.c_int32), ("reserved", C.c_int32),
        ("n_photons", C.c_uint64), ("stream_salt", C.c_uint64), ("photon_offset", C.c_uint64),
    ]


class RunStats(C.Structure):
    _fields_ = [
        ("photons", C.c_uint64), ("photon_steps", C.c_uint64), ("killed", C.c_uint64),
        ("kernel_ms", C.c_double), ("wall_ms", C.c_double),
        ("launches", C.c_uint32), ("block_size", C.c_uint32), ("grid_size", C.c_uint32),
        ("refill_quantum", C.c_uint32), ("wave_iterations", C.c_uint64), ("bad_contributions", C.c_uint64),
        ("record_fetch", C.c_uint32), ("entry_skip", C.c_uint32),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class BatchOpts(C.Structure):
    _fields_ = [
        ("block_size", C.c_int32), ("refill_quantum", C.c_int32), ("blocks_per_cu", C.c_int32),
        ("use_graph", C.c_int32), ("autotune", C.c_int32), ("scatter_quantum", C.c_int32),
        ("tally_copies", C.c_int32), ("entry_skip", C.c_int32), ("max_steps_per_photon", C.c_uint64),
        ("record_fetch", C.c_int32), ("reserved0", C.c_int32),
    ]


class ServeOpts(C.Structure):
    _fields_ = [
        ("max_inflight_scans", C.c_int32), ("block_size", C.c_int32), ("refill_quantum", C.c_int32),
        ("fail_request_index", C.c_int32), ("hbm_budget_bytes", C.c_uint64), ("slice_photons", C.c_uint64),
        ("policy", C.c_int32), ("record_fetch", C.c_int32),
    ]


class Request(C.Structure):
    _fields_ = [
        ("mesh_id", C.c_int32), ("n_scans", C.c_int32), ("params", C.POINTER(Params)),
        ("scans", C.POINTER(AScan)), ("seed", C.c_uint64), ("tallies_out", C.POINTER(C.c_uint64)),
    ]


class RequestInfo(C.Structure):
    _fields_ = [
        ("status", C.c_int32), ("done", C.c_int32), ("t_submit_ms", C.c_double),
        ("t_admit_ms", C.c_double), ("t_done_ms", C.c_double), ("photons", C.c_uint64),
    ]


class PhotonLogRec(C.Structure):
    _fields_ = [("steps", C.c_uint32), ("detections", C.c_uint32), ("contrib", C.c_float), ("killed", C.c_uint32)]


# (name, restype, argtypes) of every exported function; include/octmps.h declares them.
_P = C.c_void_p
_u64p = C.POINTER(C.c_uint64)
SIGNATURES = [
    ("octmps_status_string", C.c_char_p, [C.c_int]),
    ("octmps_last_error", C.c_char_p, []),
    ("octmps_version", C.c_char_p, []),
    ("octmps_read_opt", C.c_int, [C.c_char_p, C.c_int32, C.POINTER(Params), C.POINTER(C.c_int32)]),
    ("octmps_read_mci", C.c_int, [C.c_char_p, C.POINTER(Params)]),
    ("octmps_params_default", None, [C.POINTER(Params)]),
    ("octmps_rspecular", C.c_double, [C.c_double, C.c_double]),
    ("octmps_mesh_read", C.c_int, [C.c_char_p, C.POINTER(_P)]),
    ("octmps_mesh_build", C.c_int, [C.c_int64, C.POINTER(C.c_double), C.c_int64, C.POINTER(C.c_int32), C.POINTER(_P)]),
    ("octmps_mesh_generate", C.c_int, [C.POINTER(GridSpec), C.POINTER(_P)]),
    ("octmps_mesh_free", None, [_P]),
    ("octmps_mesh_num_tets", C.c_int64, [_P]),
    ("octmps_mesh_num_vertices", C.c_int64, [_P]),
    ("octmps_mesh_get_soa", C.c_int, [_P, C.POINTER(C.c_double), C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    ("octmps_mesh_find_root", C.c_int, [_P, C.c_double, C.c_double, C.c_double, C.c_int32, C.POINTER(C.c_int32),
                                        C.POINTER(C.c_double)]),
    ("octmps_mesh_write", C.c_int, [_P, C.c_char_p]),
    ("octmps_safeprimes", C.c_int, [C.c_uint32, C.POINTER(C.c_uint32)]),
    ("octmps_set_safeprimes_file", C.c_int, [C.c_char_p]),
    ("octmps_set_stream_triplets", C.c_int, [C.c_uint32]),
    ("octmps_photon_streams", C.c_int, [C.c_uint64, C.c_uint64, C.c_uint64, _u64p, C.POINTER(C.c_uint32)]),
    ("octmps_tally_words", C.c_uint64, [C.c_int32]),
    ("octmps_tally_finalize", C.c_int, [_u64p, C.c_int32, C.c_uint64, C.c_double, C.POINTER(C.c_double)]),
    ("octmps_write_ascan_outputs", C.c_int, [C.c_char_p, C.POINTER(Params), C.c_double, _u64p]),
    ("octmps_ctx_create", C.c_int, [C.c_int32, C.POINTER(_P)]),
    ("octmps_ctx_destroy", None, [_P]),
    ("octmps_ctx_upload_mesh", C.c_int, [_P, _P, C.POINTER(C.c_int32)]),
    ("octmps_ctx_register_mesh", C.c_int, [_P, _P, C.POINTER(C.c_int32)]),
    ("octmps_ctx_release_mesh", C.c_int, [_P, C.c_int32]),
    ("octmps_ctx_mesh_info", C.c_int, [_P, C.c_int32, C.POINTER(C.c_int32), _u64p]),
    ("octmps_ctx_memory", C.c_int, [_P, _u64p, _u64p]),
    ("octmps_run_ascan", C.c_int, [_P, C.c_int32, C.POINTER(Params), C.POINTER(AScan), C.c_uint64,
                                   C.POINTER(BatchOpts), _u64p, C.POINTER(RunStats)]),
    ("octmps_run_bscan", C.c_int, [_P, C.c_int32, C.POINTER(Params), C.c_int32, C.POINTER(AScan), C.c_uint64,
                                   C.POINTER(BatchOpts), _u64p, C.POINTER(RunStats)]),
    ("octmps_run_bscan_noalloc", C.c_int, [_P, C.c_int32, C.POINTER(Params), C.c_int32, C.POINTER(AScan), C.c_uint64,
                                           C.POINTER(BatchOpts), _u64p, C.POINTER(RunStats)]),
    ("octmps_run_bscan_logged", C.c_int, [_P, C.c_int32, C.POINTER(Params), C.c_int32, C.POINTER(AScan), C.c_uint64,
                                          C.POINTER(BatchOpts), _u64p, C.POINTER(RunStats), _P, C.c_uint64]),
    ("octmps_autotune", C.c_int, [_P, C.c_int32, C.POINTER(Params), C.c_int32, C.POINTER(AScan), C.c_uint64,
                                  C.c_uint64, C.POINTER(BatchOpts), C.POINTER(C.c_double)]),
    ("octmps_serve_open", C.c_int, [_P, C.POINTER(ServeOpts), C.POINTER(_P)]),
    ("octmps_serve_submit", C.c_int, [_P, C.POINTER(Request), _u64p]),
    ("octmps_serve_poll", C.c_int, [_P, C.c_uint64, C.POINTER(RequestInfo)]),
    ("octmps_serve_wait", C.c_int, [_P, C.c_uint64, C.c_double, C.POINTER(RequestInfo)]),
    ("octmps_serve_cancel", C.c_int, [_P, C.c_uint64]),
    ("octmps_serve_release", C.c_int, [_P, C.c_uint64]),
    ("octmps_serve_residency", C.c_int, [_P, _u64p, _u64p, _u64p, _u64p]),
    ("octmps_serve_stats", C.c_int, [_P, _u64p, _u64p, C.POINTER(C.c_double), _u64p]),
    ("octmps_serve_close", C.c_int, [_P]),
    ("octmps_probe_rng", C.c_int, [_P, C.c_uint64, C.c_uint64, C.c_uint64, C.c_int32, C.c_int32,
                                   C.POINTER(C.c_float)]),
    ("octmps_probe_fresnel", C.c_int, [_P, C.c_int32, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                       C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    ("octmps_probe_sampler", C.c_int, [_P, C.c_int32, C.c_float, C.c_int32, C.POINTER(C.c_float),
                                       C.POINTER(C.c_float)]),
    ("octmps_probe_tally", C.c_int, [_P, C.POINTER(Params), C.c_int32, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                     C.POINTER(C.c_float), _u64p, _u64p]),
    ("octmps_probe_copy", C.c_int, [_P, C.c_uint64, C.c_int32, C.POINTER(C.c_double)]),
    ("octmps_probe_gather", C.c_int, [_P, C.c_int32, C.c_int32, C.POINTER(C.c_double)]),
    ("octmps_probe_l1", C.c_int, [_P, C.c_int32, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    ("octmps_probe_valu", C.c_int, [_P, C.c_int32, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
]


def load(path=LIB_PATH):
    if not os.path.exists(path):
        raise ImportError(f"liboctmps is not built ({path}); run __graft_entry__.build()")
    lib = C.CDLL(path)
    for name, res, args in SIGNATURES:
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = load()


def check(status, where):
    if status != OK:
        raise OctmpsError(status, where, lib.octmps_last_error().decode(errors="replace"))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _up(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint64))


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


# ------------------------------------------------------------------ parameters
def default_params():
    p = Params()
    lib.octmps_params_default(C.byref(p))
    return p


def make_params(regions, n_ambient=1.0, num_bins=256, depth=0.1, **kw):
    """Parameters with explicit bins: Lc = 6 * depth / num_bins so that the reference's
    binning (width Lc/6, src/octmps_io.cpp:333-335) spans [0, depth)."""
    p = default_params()
    p.n_regions = len(regions) + 1
    p.regions[0].n = n_ambient
    for i, (n, mua, mus, g) in enumerate(regions, start=1):
        p.regions[i].n, p.regions[i].mua, p.regions[i].mus, p.regions[i].g = n, mua, mus, g
    p.num_depth_bins = num_bins
    p.coherence_length_source = SUBSTEPS * depth / num_bins
    p.target_depth_min = 0.0
    p.target_depth_max = depth
    p.max_filtered_optical_depth = depth
    p.optical_depth_shift = 0.0
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def read_opt(path, run_index=0, params=None):
    p = params if params is not None else default_params()
    n = C.c_int32(0)
    check(lib.octmps_read_opt(path.encode(), run_index, C.byref(p), C.byref(n)), "read_opt")
    return p, n.value


def read_mci(path, params=None):
    p = params if params is not None else default_params()
    check(lib.octmps_read_mci(path.encode(), C.byref(p)), "read_mci")
    return p


def rspecular(ni, nt):
    return lib.octmps_rspecular(ni, nt)


# ------------------------------------------------------------------ mesh
class Mesh:
    def __init__(self, handle):
        self.h = handle

    @classmethod
    def read(cls, path):
        h = C.c_void_p()
        check(lib.octmps_mesh_read(path.encode(), C.byref(h)), "mesh_read")
        return cls(h)

    @classmethod
    def build(cls, xyz, tets):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        tets = np.ascontiguousarray(tets, dtype=np.int32)
        h = C.c_void_p()
        check(lib.octmps_mesh_build(len(xyz), _dp(xyz), len(tets), _ip(tets), C.byref(h)), "mesh_build")
        return cls(h)

    @classmethod
    def generate(cls, kind, n, size, p=(), n_layers=0):
        s = GridSpec()
        s.kind = kind
        s.nx, s.ny, s.nz = n
        s.size_x, s.size_y, s.size_z = size
        s.n_layers = n_layers
        for i, v in enumerate(p):
            s.p[i] = v
        h = C.c_void_p()
        check(lib.octmps_mesh_generate(C.byref(s), C.byref(h)), "mesh_generate")
        return cls(h)

    def __del__(self):
        if getattr(self, "h", None):
            lib.octmps_mesh_free(self.h)
            self.h = None

    @property
    def num_tets(self):
        return lib.octmps_mesh_num_tets(self.h)

    @property
    def num_vertices(self):
        return lib.octmps_mesh_num_vertices(self.h)

    def soa(self):
        t = self.num_tets
        planes = np.empty((t, 4, 4), np.float64)
        links = np.empty((t, 4), np.int32)
        regions = np.empty(t, np.int32)
        check(lib.octmps_mesh_get_soa(self.h, _dp(planes), _ip(links), _ip(regions)), "mesh_get_soa")
        return planes, links, regions

    def find_root(self, x, y=0.0, z=0.0, project=False):
        r = C.c_int32(-1)
        zo = C.c_double(0)
        check(lib.octmps_mesh_find_root(self.h, x, y, z, 1 if project else 0, C.byref(r), C.byref(zo)), "find_root")
        return r.value, zo.value

    def write(self, path):
        check(lib.octmps_mesh_write(self.h, path.encode()), "mesh_write")


# ------------------------------------------------------------------ RNG
def safeprimes(n):
    out = np.empty(n, np.uint32)
    check(lib.octmps_safeprimes(n, out.ctypes.data_as(C.POINTER(C.c_uint32))), "safeprimes")
    return out


def photon_streams(seed, salt, photon):
    x = np.empty(3, np.uint64)
    a = np.empty(3, np.uint32)
    check(lib.octmps_photon_streams(seed, salt, photon, _up(x), a.ctypes.data_as(C.POINTER(C.c_uint32))), "streams")
    return x, a


# ------------------------------------------------------------------ tallies
def tally_words(bins):
    return int(lib.octmps_tally_words(bins))


TALLY_SUM_DIGITS, TALLY_SQ_DIGITS, TALLY_CLASS_WORDS = 9, 18, 29
TALLY_SUM_EXP, TALLY_SQ_EXP = -160, -320


def _digits_value(blk, exp0):
    """sum_k blk[k] * 2^(32k + exp0), accumulated in extended precision (exact for the
    statistics the tests need; octmps_tally_finalize converts exactly)."""
    acc = np.zeros(blk.shape[1], np.longdouble)
    for k in range(blk.shape[0] - 1, -1, -1):
        acc += np.ldexp(blk[k].astype(np.longdouble), 32 * k + exp0)
    return acc.astype(np.float64)


def decode_tally(t, bins):
    """Integer tally -> dict of float64 arrays [class, bin] (class 0 = Class I)."""
    t = np.asarray(t, dtype=np.uint64)
    out = {"num_filtered": int(t[0]), "num_class1": int(t[1]), "num_class2": int(t[2]), "bad": int(t[3])}
    s, sq, cnt, mx = [], [], [], []
    for c in range(2):
        base = 4 + c * TALLY_CLASS_WORDS * bins
        blk = t[base:base + TALLY_CLASS_WORDS * bins].reshape(TALLY_CLASS_WORDS, bins)
        s.append(_digits_value(blk[:TALLY_SUM_DIGITS], TALLY_SUM_EXP))
        sq.append(_digits_value(blk[TALLY_SUM_DIGITS:TALLY_SUM_DIGITS + TALLY_SQ_DIGITS], TALLY_SQ_EXP))
        cnt.append(blk[TALLY_CLASS_WORDS - 2].astype(np.float64))
        mx.append(blk[TALLY_CLASS_WORDS - 1].view(np.float64).copy())
    out["sum"] = np.array(s)
    out["sumsq"] = np.array(sq)
    out["count"] = np.array(cnt)
    out["max"] = np.array(mx)
    return out


def exact_tally_sums(t, bins):
    """Exact Sum and SumSq of every bin as Python Fractions: [class][bin] -> (Sum, SumSq)."""
    from fractions import Fraction
    t = np.asarray(t, dtype=np.uint64)
    out = []
    for c in range(2):
        base = 4 + c * TALLY_CLASS_WORDS * bins
        blk = t[base:base + TALLY_CLASS_WORDS * bins].reshape(TALLY_CLASS_WORDS, bins)
        row = []
        for b in range(bins):
            s = sum(int(blk[k, b]) << (32 * k) for k in range(TALLY_SUM_DIGITS))
            q = sum(int(blk[TALLY_SUM_DIGITS + k, b]) << (32 * k) for k in range(TALLY_SQ_DIGITS))
            row.append((Fraction(s, 2 ** -TALLY_SUM_EXP), Fraction(q, 2 ** -TALLY_SQ_EXP)))
        out.append(row)
    return out


def finalize(t, bins, n_photons, max_rel=0.01):
    out = np.empty((2, 2, 2, bins), np.float64)
    t = np.ascontiguousarray(t, dtype=np.uint64)
    check(lib.octmps_tally_finalize(_up(t), bins, n_photons, max_rel, _dp(out)), "tally_finalize")
    return out


# ------------------------------------------------------------------ device
class Context:
    def __init__(self, device=0):
        h = C.c_void_p()
        check(lib.octmps_ctx_create(device, C.byref(h)), "ctx_create")
        self.h = h
        self.meshes = []

    def close(self):
        if getattr(self, "h", None):
            lib.octmps_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def upload(self, mesh):
        mid = C.c_int32(-1)
        check(lib.octmps_ctx_upload_mesh(self.h, mesh.h, C.byref(mid)), "upload_mesh")
        return mid.value

    def register(self, mesh):
        """Registers a mesh that is made resident on demand (octmps_ctx_register_mesh)."""
        mid = C.c_int32(-1)
        check(lib.octmps_ctx_register_mesh(self.h, mesh.h, C.byref(mid)), "register_mesh")
        return mid.value

    def release(self, mesh_id):
        check(lib.octmps_ctx_release_mesh(self.h, mesh_id), "release_mesh")

    def mesh_info(self, mesh_id):
        """(resident, device bytes) of a mesh."""
        res, nb = C.c_int32(0), C.c_uint64(0)
        check(lib.octmps_ctx_mesh_info(self.h, mesh_id, C.byref(res), C.byref(nb)), "mesh_info")
        return bool(res.value), nb.value

    def run_ascan(self, mesh_id, params, scan, seed, opts=None):
        bins = params.num_depth_bins
        t = np.zeros(tally_words(bins), np.uint64)
        st = RunStats()
        check(lib.octmps_run_ascan(self.h, mesh_id, C.byref(params), C.byref(scan), seed,
                                   C.byref(opts) if opts is not None else None, _up(t), C.byref(st)), "run_ascan")
        return t, st

    def run_bscan(self, mesh_id, params, scans, seed, opts=None, log=False):
        bins = params.num_depth_bins
        arr = (AScan * len(scans))(*scans)
        t = np.zeros(tally_words(bins) * len(scans), np.uint64)
        st = RunStats()
        o = C.byref(opts) if opts is not None else None
        if log:
            total = sum(s.n_photons for s in scans)
            plog = (PhotonLogRec * total)()
            check(lib.octmps_run_bscan_logged(self.h, mesh_id, C.byref(params), len(scans), arr, seed, o, _up(t),
                                              C.byref(st), C.cast(plog, C.c_void_p), total), "run_bscan_logged")
            rec = np.frombuffer(plog, dtype=np.dtype([("steps", "<u4"), ("detections", "<u4"),
                                                      ("contrib", "<f4"), ("killed", "<u4")]))
            return t.reshape(len(scans), -1), st, rec.copy()
        check(lib.octmps_run_bscan(self.h, mesh_id, C.byref(params), len(scans), arr, seed, o, _up(t), C.byref(st)),
              "run_bscan")
        return t.reshape(len(scans), -1), st

    def autotune(self, mesh_id, params, scans, seed, calib=1 << 20):
        arr = (AScan * len(scans))(*scans)
        o = BatchOpts()
        rate = C.c_double(0)
        check(lib.octmps_autotune(self.h, mesh_id, C.byref(params), len(scans), arr, seed, calib, C.byref(o),
                                  C.byref(rate)), "autotune")
        return o, rate.value

    def probe_rng(self, seed, salt, photon0, n_photons, n_draws):
        out = np.empty((n_photons, 3, n_draws), np.float32)
        check(lib.octmps_probe_rng(self.h, seed, salt, photon0, n_photons, n_draws, _fp(out)), "probe_rng")
        return out

    def probe_fresnel(self, ni, nt, ca1):
        ni, nt, ca1 = (np.ascontiguousarray(v, np.float32) for v in (ni, nt, ca1))
        r = np.empty(len(ni), np.float32)
        check(lib.octmps_probe_fresnel(self.h, len(ni), _fp(ni), _fp(nt), _fp(ca1), _fp(r)), "probe_fresnel")
        return r

    def probe_tally(self, params, op, maxz, contrib):
        """Device detector tally of the given detections into one zeroed tally: (tally, bad)."""
        op, maxz, contrib = (np.ascontiguousarray(v, np.float32) for v in (op, maxz, contrib))
        t = np.zeros(tally_words(params.num_depth_bins), np.uint64)
        bad = C.c_uint64(0)
        check(lib.octmps_probe_tally(self.h, C.byref(params), len(op), _fp(op), _fp(maxz), _fp(contrib), _up(t),
                                     C.byref(bad)), "probe_tally")
        return t, bad.value

    def probe_copy(self, nbytes, repeats=3):
        ms = C.c_double(0)
        check(lib.octmps_probe_copy(self.h, nbytes, repeats, C.byref(ms)), "probe_copy")
        return ms.value

    def probe_gather(self, n_records=1 << 14, iters=4000):
        """16-byte lane-loads per second of cached divergent record gathers (5 per record)."""
        r = C.c_double(0)
        check(lib.octmps_probe_gather(self.h, n_records, iters, C.byref(r)), "probe_gather")
        return r.value

    def probe_l1(self, pattern, iters=2048):
        """(wave load instructions/s, CU cycles per load at the nominal clock) of an L1 access
        pattern: 0 divergent, 1 quad-coalesced, 2 half the lanes masked, 3 half out of range,
        4 quad-coalesced DMA into LDS."""
        r, cyc = C.c_double(0), C.c_double(0)
        check(lib.octmps_probe_l1(self.h, pattern, iters, C.byref(r), C.byref(cyc)), "probe_l1")
        return r.value, cyc.value

    def probe_valu(self, kind, iters=4096):
        """(wave VALU instructions/s, SIMD cycles per instruction at the nominal clock) of an
        instruction class: 0 FP32 fused multiply-add, 1 FP32 reciprocal, 2 64-bit integer
        multiply-add."""
        r, cyc = C.c_double(0), C.c_double(0)
        check(lib.octmps_probe_valu(self.h, kind, iters, C.byref(r), C.byref(cyc)), "probe_valu")
        return r.value, cyc.value

    def probe_sampler(self, kind, param, u):
        u = np.ascontiguousarray(u, np.float32)
        out = np.empty(len(u), np.float32)
        check(lib.octmps_probe_sampler(self.h, kind, param, len(u), _fp(u), _fp(out)), "probe_sampler")
        return out


class Server:
    def __init__(self, ctx, **opts):
        o = ServeOpts()
        o.fail_request_index = -1
        for k, v in opts.items():
            setattr(o, k, v)
        h = C.c_void_p()
        check(lib.octmps_serve_open(ctx.h, C.byref(o), C.byref(h)), "serve_open")
        self.h = h
        self._keep = {}

    def submit(self, mesh_id, params, scans, seed):
        arr = (AScan * len(scans))(*scans)
        out = np.zeros(tally_words(params.num_depth_bins) * len(scans), np.uint64)
        r = Request()
        r.mesh_id = mesh_id
        r.n_scans = len(scans)
        r.params = C.pointer(params)
        r.scans = arr
        r.seed = seed
        r.tallies_out = _up(out)
        rid = C.c_uint64(0)
        check(lib.octmps_serve_submit(self.h, C.byref(r), C.byref(rid)), "serve_submit")
        self._keep[rid.value] = (out, len(scans))
        return rid.value

    def wait(self, rid, timeout=-1.0):
        info = RequestInfo()
        check(lib.octmps_serve_wait(self.h, rid, timeout, C.byref(info)), "serve_wait")
        out, n = self._keep[rid]
        return info, out.reshape(n, -1)

    def poll(self, rid):
        info = RequestInfo()
        check(lib.octmps_serve_poll(self.h, rid, C.byref(info)), "serve_poll")
        return info

    def cancel(self, rid):
        check(lib.octmps_serve_cancel(self.h, rid), "serve_cancel")

    def release(self, rid):
        """Forget a finished request (octmps_serve_release) and drop its output buffer."""
        check(lib.octmps_serve_release(self.h, rid), "serve_release")
        self._keep.pop(rid, None)

    def residency(self):
        up, ev, used, budget = C.c_uint64(), C.c_uint64(), C.c_uint64(), C.c_uint64()
        check(lib.octmps_serve_residency(self.h, C.byref(up), C.byref(ev), C.byref(used), C.byref(budget)),
              "serve_residency")
        return {"uploads": up.value, "evictions": ev.value, "used": used.value, "budget": budget.value}

    def stats(self):
        ph, st, sl = C.c_uint64(), C.c_uint64(), C.c_uint64()
        km = C.c_double()
        check(lib.octmps_serve_stats(self.h, C.byref(ph), C.byref(st), C.byref(km), C.byref(sl)), "serve_stats")
        return {"photons": ph.value, "steps": st.value, "kernel_ms": km.value, "slices": sl.value}

    def close(self):
        if self.h:
            check(lib.octmps_serve_close(self.h), "serve_close")
            self.h = None


def ascan(x, y, z, root, n_photons, salt=0, offset=0):
    s = AScan()
    s.x, s.y, s.z = x, y, z
    s.root_tet = root
    s.n_photons = n_photons
    s.stream_salt = salt
    s.photon_offset = offset
    return s
