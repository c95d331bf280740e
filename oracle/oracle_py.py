"""ctypes wrapper of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the CPU-baseline
leg of bench.py, always as the checker, never as the measured or shipped path.
"""
import contextlib
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libref_host.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise ImportError(f"oracle not built ({LIB}); run __graft_entry__.build()")
        _lib = _configure(C.CDLL(LIB))
    return _lib


@contextlib.contextmanager
def variant(path):
    """Runs the block with another build of oracle.cpp (a planted-mistake variant that a pin
    test must reject), then restores the oracle."""
    global _lib
    saved = lib()
    _lib = _configure(C.CDLL(path))
    try:
        yield
    finally:
        _lib = saved


def _configure(L):
    dp, ip, up = C.POINTER(C.c_double), C.POINTER(C.c_int32), C.POINTER(C.c_uint64)
    u32p = C.POINTER(C.c_uint32)
    L.oracle_run_ascan.restype = C.c_int
    L.oracle_run_ascan.argtypes = [dp, ip, ip, C.c_int64, C.c_void_p, C.c_void_p, C.c_uint64, u32p, C.c_uint32,
                                   C.c_uint64, C.c_int, dp, up, dp]
    L.oracle_trace_photon.restype = C.c_int
    L.oracle_trace_photon.argtypes = [dp, ip, ip, C.c_int64, C.c_void_p, C.c_void_p, C.c_uint64, u32p,
                                      C.c_uint32, C.c_uint64, C.c_uint64, C.c_int, dp]
    L.oracle_div_rd.restype = C.c_double
    L.oracle_div_rd.argtypes = [C.c_double, C.c_double]
    L.oracle_rand_co.restype = None
    L.oracle_rand_co.argtypes = [up, C.c_uint32, C.c_int, dp]
    L.oracle_photon_streams.restype = None
    L.oracle_photon_streams.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, u32p, C.c_uint32, up, u32p]
    L.oracle_fresnel.restype = C.c_double
    L.oracle_fresnel.argtypes = [C.c_double, C.c_double, C.c_double]
    L.oracle_spin_theta.restype = C.c_double
    L.oracle_spin_theta.argtypes = [C.c_double, C.c_double]
    L.oracle_spin_theta_bias.restype = C.c_double
    L.oracle_spin_theta_bias.argtypes = [C.c_double, C.c_double]
    L.oracle_hit_boundary.restype = None
    L.oracle_hit_boundary.argtypes = [dp, dp, dp, C.c_double, C.c_double, dp]
    L.oracle_spin_bias_batch.restype = None
    L.oracle_spin_bias_batch.argtypes = [C.c_void_p, dp, C.c_double, C.c_double, C.c_double, C.c_uint64, u32p,
                                         C.c_uint32, C.c_int, dp]
    L.oracle_tally_detection.restype = None
    L.oracle_tally_detection.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double,
                                         C.c_double, dp, up, up]
    L.oracle_detect_filter.restype = C.c_int
    L.oracle_detect_filter.argtypes = [C.c_void_p, dp, C.c_double, C.c_double]
    L.oracle_roulette_batch.restype = None
    L.oracle_roulette_batch.argtypes = [C.c_void_p, dp, dp, C.c_double, C.c_uint64, u32p, C.c_uint32, C.c_int,
                                        dp]
    return L


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _up(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint64))


def _u32p(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


class Table:
    """MWC multiplier table as used by the product (3 seeding + 3 per triplet)."""

    def __init__(self, mult, n_triplets):
        self.mult = np.ascontiguousarray(mult, dtype=np.uint32)
        self.k = int(n_triplets)


def run_ascan(soa, params, scan, seed, table, threads=8, max_steps=0, log=False):
    planes, links, regions = soa
    planes = np.ascontiguousarray(planes, np.float64)
    links = np.ascontiguousarray(links, np.int32)
    regions = np.ascontiguousarray(regions, np.int32)
    B = params.num_depth_bins
    tally = np.zeros(2 * 4 * B, np.float64)
    cnt = np.zeros(6, np.uint64)
    plog = np.zeros(3 * scan.n_photons, np.float64) if log else None
    rc = lib().oracle_run_ascan(_dp(planes), _ip(links), _ip(regions), len(regions), C.addressof(params),
                                C.addressof(scan), seed, _u32p(table.mult), table.k, max_steps, threads,
                                _dp(tally), _up(cnt), _dp(plog) if log else None)
    if rc != 0:
        raise RuntimeError(f"oracle_run_ascan failed ({rc})")
    t = tally.reshape(2, 4, B)
    out = {"sum": t[:, 0].copy(), "sumsq": t[:, 1].copy(), "max": t[:, 2].copy(), "count": t[:, 3].copy(),
           "num_filtered": int(cnt[0]), "num_class1": int(cnt[1]), "num_class2": int(cnt[2]),
           "photons": int(cnt[3]), "steps": int(cnt[4]), "killed": int(cnt[5])}
    if log:
        out["log"] = plog.reshape(-1, 3)
    return out


def trace_photon(soa, params, scan, seed, table, photon, max_events=4096, max_steps=0):
    planes, links, regions = soa
    planes = np.ascontiguousarray(planes, np.float64)
    links = np.ascontiguousarray(links, np.int32)
    regions = np.ascontiguousarray(regions, np.int32)
    out = np.zeros((max_events, 16), np.float64)
    n = lib().oracle_trace_photon(_dp(planes), _ip(links), _ip(regions), len(regions), C.addressof(params),
                                  C.addressof(scan), seed, _u32p(table.mult), table.k, photon, max_steps,
                                  max_events, _dp(out))
    if n < 0:
        raise RuntimeError("oracle_trace_photon failed")
    return out[:min(n, max_events)]


def div_rd(a, b):
    return lib().oracle_div_rd(a, b)


def rand_co(x, a, n):
    xs = C.c_uint64(x)
    out = np.empty(n, np.float64)
    lib().oracle_rand_co(C.byref(xs), a, n, _dp(out))
    return out, xs.value


def photon_streams(seed, salt, photon, table):
    x = np.empty(3, np.uint64)
    a = np.empty(3, np.uint32)
    lib().oracle_photon_streams(seed, salt, photon, _u32p(table.mult), table.k, _up(x), _u32p(a))
    return x, a


def fresnel(ni, nt, ca1):
    return lib().oracle_fresnel(ni, nt, ca1)


def spin_theta(g, u):
    return lib().oracle_spin_theta(g, u)


def spin_theta_bias(a, u):
    return lib().oracle_spin_theta_bias(a, u)


def hit_boundary(planes16, pos, direction, s, mut):
    out = np.zeros(5, np.float64)
    lib().oracle_hit_boundary(_dp(np.ascontiguousarray(planes16, np.float64)),
                              _dp(np.ascontiguousarray(pos, np.float64)),
                              _dp(np.ascontiguousarray(direction, np.float64)), s, mut, _dp(out))
    return out


def tally_detection(params, op, maxz, w, r_fresnel, lr):
    """One detection into an empty tally: (sums[2, B], counts[2, B], counters[3])."""
    B = params.num_depth_bins
    sums = np.zeros(2 * B, np.float64)
    counts = np.zeros(2 * B, np.uint64)
    counters = np.zeros(3, np.uint64)
    lib().oracle_tally_detection(C.addressof(params), op, maxz, w, r_fresnel, lr, _dp(sums), _up(counts),
                                 _up(counters))
    return sums.reshape(2, B), counts.reshape(2, B), counters


def detect_filter(params, x, y, uz, fst_back, loc_fst, max_depth, nbs, probe=(0.0, 0.0)):
    st = np.array([x, y, uz, fst_back, loc_fst, max_depth, nbs], np.float64)
    return lib().oracle_detect_filter(C.addressof(params), _dp(st), probe[0], probe[1])


def roulette_batch(params, photon, cont, g, seed, table, n):
    """photon, cont: x,y,z,ux,uy,uz,w,LR,LR_afterFstBias,FstBack. Rows: fate, x,y,z,ux,uy,uz,w,LR,FstBack."""
    ph = np.ascontiguousarray(photon, np.float64)
    ct = np.ascontiguousarray(cont, np.float64)
    out = np.zeros((n, 10), np.float64)
    lib().oracle_roulette_batch(C.addressof(params), _dp(ph), _dp(ct), g, seed, _u32p(table.mult), table.k, n,
                                _dp(out))
    return out


def spin_bias_batch(params, state, probe, seed, table, n):
    st = np.ascontiguousarray(state, np.float64)
    out = np.zeros((n, 8), np.float64)
    lib().oracle_spin_bias_batch(C.addressof(params), _dp(st), probe[0], probe[1], probe[2], seed,
                                 _u32p(table.mult), table.k, n, _dp(out))
    return out
