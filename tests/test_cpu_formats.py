"""Drop-in formats against the reference's own host code (oracle/_ref/libref_host.so, built
from src/mesh.cpp, src/bias.cpp and src/octmps_io.cpp by oracle/Makefile): .opt/.mci
parsing, mesh serialisation (planes, signs, adjacency, root) and the CSV/.mco writer."""
import ctypes as C
import filecmp
import os

import numpy as np
import pytest

import octmps as O
from tally_util import make_tally
import oracle_py as OR

INPUT = "/root/reference/input"
CASES = {
    "Sphere": ("Sphere/Sphere.opt", "Sphere/BiasStartup.mci", "Sphere/Sphere.mesh"),
    "FourLayers": ("FourLayers/FourLayers.opt", "FourLayers/BiasStartup.mci", "FourLayers/FourLayers.mesh"),
    "WangBump": ("WangBump/WangBump.opt", "WangBump/BiasStartup.mci", None),
    "Ellipsoid2S": ("Ellipsoidand2Spheres/ellipsoid_and_two_spheres_60_degree.opt",
                    "Ellipsoidand2Spheres/BiasStartup.mci", None),
}

needs_ref = pytest.mark.skipif(not (os.path.exists(OR.REF_LIB) and os.path.isdir(INPUT)),
                               reason="reference host build (oracle/_ref) or sample inputs absent")


def ref():
    L = C.CDLL(OR.REF_LIB)
    L.ref_read_inputs.restype = C.c_int
    L.ref_read_inputs.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(C.c_double), C.c_int]
    L.ref_serialize_mesh.restype = C.c_int
    L.ref_serialize_mesh.argtypes = [C.c_char_p, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_int),
                                     C.POINTER(C.c_int), C.POINTER(C.c_int)]
    dp, up = C.POINTER(C.c_double), C.POINTER(C.c_ulonglong)
    L.ref_write_result.restype = None
    L.ref_write_result.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_double, dp, dp, dp, up, dp, dp, dp, up,
                                   C.c_ulonglong, C.c_ulonglong, C.c_ulonglong]
    return L


@needs_ref
@pytest.mark.parametrize("case", sorted(CASES))
def test_opt_and_mci_parse_like_reference(case):
    opt, mci, _ = (os.path.join(INPUT, f) if f else None for f in CASES[case])
    buf = np.zeros(1024, np.float64)
    n = ref().ref_read_inputs(opt.encode(), mci.encode(), buf.ctypes.data_as(C.POINTER(C.c_double)), 1024)
    assert n > 0
    p, runs = O.read_opt(opt)
    O.read_mci(mci, p)
    k = 0
    assert buf[k] == p.n_regions
    k += 1
    for i in range(p.n_regions):
        r = p.regions[i]
        assert buf[k] == r.n
        if i:
            assert (buf[k + 1], buf[k + 2], buf[k + 3]) == (r.mua, r.mus, r.g)
        k += 5
    mine = [p.number_of_photons, p.probe_start_x, p.probe_step_x, p.probe_end_x, p.type_simulation, p.type_bias,
            p.backward_bias_coefficient, p.target_depth_min, p.target_depth_max, p.coherence_length_source,
            p.record_photons_flag, p.max_collecting_radius, p.max_collecting_angle_deg,
            p.probability_additional_bias, p.max_relative_contribution, p.max_filtered_optical_depth,
            p.num_depth_bins, p.optical_depth_shift]
    assert list(buf[k:k + len(mine)]) == [float(v) for v in mine]


def test_mci_float_parse_truncates_bins():
    # 6 * 0.19f / 0.0015f = 759.99998 in double -> 759 (src/bias.cpp:99-101).
    mci = os.path.join(os.path.dirname(__file__), "golden", "bias_sphere.mci")
    p = O.read_mci(mci)
    assert p.num_depth_bins == 759
    assert p.target_depth_max == float(np.float32(0.19))
    assert p.max_filtered_optical_depth == float(np.float32(0.19))


@needs_ref
@pytest.mark.parametrize("case", ["Sphere", "FourLayers"])
def test_mesh_serialisation_like_reference(case):
    opt, _, mesh_path = (os.path.join(INPUT, f) for f in CASES[case])
    p, _ = O.read_opt(opt)
    m = O.Mesh.read(mesh_path)
    T = m.num_tets
    planes = np.zeros((T, 16), np.float64)
    links = np.zeros((T, 4), np.int32)
    regions = np.zeros(T, np.int32)
    nt = C.c_int(0)
    root = ref().ref_serialize_mesh(mesh_path.encode(), p.probe_start_x, planes.ctypes.data_as(C.POINTER(C.c_double)),
                                    links.ctypes.data_as(C.POINTER(C.c_int)),
                                    regions.ctypes.data_as(C.POINTER(C.c_int)), C.byref(nt))
    assert nt.value == T
    assert root >= 0, "reference found no root"
    mp, ml, mr = m.soa()
    assert np.array_equal(mr, regions)
    assert np.array_equal(ml, links)
    assert np.array_equal(mp.reshape(T, 16), planes, equal_nan=True)  # bit-identical FP64 planes
    mine, _ = m.find_root(p.probe_start_x, 0.0, 0.0)
    assert mine == root


@needs_ref
def test_writer_byte_identical_to_reference(tmp_path):
    opt, mci, _ = (os.path.join(INPUT, f) if f else None for f in CASES["Sphere"])
    p, _ = O.read_opt(opt)
    O.read_mci(mci, p)
    B = p.num_depth_bins
    rng = np.random.default_rng(4)
    k = rng.integers(0, 2 ** 36, size=(2, B)) * (rng.random((2, B)) < 0.7)
    sums = k.astype(np.float64) * 2.0 ** -30
    sqs = np.rint(sums * sums * 0.3 * 2 ** 30) * 2.0 ** -30
    # Max is an FP32 contribution on the device, so the test's maxima are FP32 values.
    maxs = np.where(rng.random((2, B)) < 0.5, sums * 0.5, sums * 0.001).astype(np.float32).astype(np.float64)
    cnts = rng.integers(0, 1000, size=(2, B)).astype(np.uint64)
    nf = [12345, 678, 90]
    t = make_tally(sums, sqs, maxs, cnts, nf)
    dec = O.decode_tally(t, B)
    assert np.array_equal(dec["sum"], sums) and np.array_equal(dec["sumsq"], sqs)
    mine, theirs = tmp_path / "mine", tmp_path / "ref"
    mine.mkdir()
    theirs.mkdir()
    for x in (-0.01, -0.0098):
        O.check(O.lib.octmps_write_ascan_outputs(str(mine).encode(), C.byref(p), x,
                                                 t.ctypes.data_as(C.POINTER(C.c_uint64))), "write")
        dp = lambda a: np.ascontiguousarray(a).ctypes.data_as(C.POINTER(C.c_double))
        up = lambda a: np.ascontiguousarray(a).ctypes.data_as(C.POINTER(C.c_ulonglong))
        ref().ref_write_result(str(theirs).encode(), opt.encode(), mci.encode(), x, dp(sums[0]), dp(sqs[0]),
                               dp(maxs[0]), up(cnts[0]), dp(sums[1]), dp(sqs[1]), dp(maxs[1]), up(cnts[1]),
                               *nf)
    for f in ("BiasParam.mco", "ClassI_Scatt.out", "ClassII_Scatt.out", "ClassI_ScattFilt.out",
              "ClassII_ScattFilt.out", "FilteredPhotons.mco"):
        assert filecmp.cmp(mine / f, theirs / f, shallow=False), f


TWO_RUNS = ("1.0  # version\n2  # runs\n"
            "run0.mco A\n3000  # photons\n2  # regions\n1.0  # ambient n\n"
            "1.40 1.0 100 0.9\n1.33 0.4 60 0.9\n-0.002 0.002 0.003  # start step end\n"
            "run1.mco A\n1500\n3\n1.0\n1.38 0.5 80 0.85\n1.36 0.2 60 0.9\n1.33 0.1 50 0.9\n"
            "-0.001 0.001 0.001\n")


@needs_ref
def test_multi_run_opt_parses_like_reference(tmp_path):
    # Each run of a multi-run .opt parses to the reference's regions, photon count and probe
    # line for that run (read_simulation_data, src/octmps_io.cpp:504-645).
    opt = tmp_path / "two.opt"
    opt.write_text(TWO_RUNS)
    L = C.CDLL(OR.REF_LIB)
    L.ref_read_run.restype = C.c_int
    L.ref_read_run.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_double), C.c_int]
    for run in (0, 1):
        buf = np.zeros(256, np.float64)
        n = L.ref_read_run(str(opt).encode(), run, buf.ctypes.data_as(C.POINTER(C.c_double)), 256)
        assert n > 0
        p, runs = O.read_opt(str(opt), run)
        assert runs == 2
        mine = [p.n_regions]
        for i in range(p.n_regions):
            r = p.regions[i]
            mine += [r.n, r.mua if i else 0, r.mus if i else 0, r.g if i else 0,
                     (1.0 / (r.mua + r.mus)) if i else 0]
        mine += [p.number_of_photons, p.probe_start_x, p.probe_step_x, p.probe_end_x]
        assert len(mine) == n
        assert np.array_equal(np.array(mine), buf[:n]), run
    with pytest.raises(O.OctmpsError) as e:
        O.read_opt(str(opt), 2)
    assert e.value.status == 5
