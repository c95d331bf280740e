"""The allocation-free batched entry point (octmps_run_bscan_noalloc), called through the C
ABI, against octmps_run_bscan: the same tallies bit for bit on two calls in a row, the second
smaller than the first (the shared scratch array keeps the first call's contents)."""
import ctypes as C

import numpy as np
import pytest

import octmps as O
from octmps import configs

pytestmark = pytest.mark.gpu


def _noalloc(ctx, mid, params, scans, seed, opts=None):
    arr = (O.AScan * len(scans))(*scans)
    t = np.zeros(O.tally_words(params.num_depth_bins) * len(scans), np.uint64)
    st = O.RunStats()
    status = O.lib.octmps_run_bscan_noalloc(ctx.h, mid, C.byref(params), len(scans), arr, seed,
                                            C.byref(opts) if opts is not None else None,
                                            t.ctypes.data_as(C.POINTER(C.c_uint64)), C.byref(st))
    return status, t.reshape(len(scans), -1), st


def test_noalloc_equals_batched_on_two_calls_in_a_row(ctx):
    mesh, params, scans, seed = configs.slab4(cells=(16, 16, 8), n_scans=6, photons=4000, bins=128)
    mid = ctx.upload(mesh)
    small = [O.ascan(s.x, s.y, s.z, s.root_tet, 1500, s.stream_salt) for s in scans[:2]]
    ref_big, st_big = ctx.run_bscan(mid, params, scans, seed)
    ref_small, st_small = ctx.run_bscan(mid, params, small, seed)
    s1, got_big, g1 = _noalloc(ctx, mid, params, scans, seed)
    s2, got_small, g2 = _noalloc(ctx, mid, params, small, seed)
    assert s1 == 0 and s2 == 0
    assert np.array_equal(got_big, ref_big)
    assert np.array_equal(got_small, ref_small)
    assert (g1.photons, g1.photon_steps) == (st_big.photons, st_big.photon_steps)
    assert (g2.photons, g2.photon_steps) == (st_small.photons, st_small.photon_steps)
    assert g1.bad_contributions == 0 and g2.bad_contributions == 0


def test_noalloc_rejects_a_run_beyond_its_scratch(ctx):
    # 1000 A-scans at 1024 bins need ~475 MB of tallies: more than the 64 MiB array. The
    # check is made on the host, before anything is launched.
    mesh, params, scans, seed = configs.sphere(cells=8, photons=10, bins=1024)
    mid = ctx.upload(mesh)
    many = [O.ascan(scans[0].x, scans[0].y, scans[0].z, scans[0].root_tet, 10, i) for i in range(1000)]
    status, _, _ = _noalloc(ctx, mid, params, many, seed)
    assert status == 5  # OCTMPS_E_RANGE
    assert b"scratch" in O.lib.octmps_last_error()
