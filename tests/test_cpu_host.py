"""Host-side logic: structured meshes, root finding, error codes, tally finalisation."""
import ctypes as C
import os

import numpy as np
import pytest

import octmps as O
from tally_util import make_tally
from octmps import configs


def test_structured_mesh_counts_and_closure():
    m = O.Mesh.generate(O.GEOM_HOMOGENEOUS, (4, 3, 2), (0.4, 0.3, 0.2))
    assert m.num_tets == 6 * 4 * 3 * 2 and m.num_vertices == 5 * 4 * 3
    planes, links, regions = m.soa()
    # every interior face is shared by exactly two tets, links are symmetric
    for t in range(m.num_tets):
        for j in range(4):
            n = links[t, j]
            if n >= 0:
                assert t in links[n]
    boundary = (links < 0).sum()
    assert boundary == 2 * 2 * (4 * 3 + 4 * 2 + 3 * 2)  # two triangles per boundary square
    # planes are unit normals, pre-signed so the centroid is inside
    assert np.allclose(np.linalg.norm(planes[:, :, :3], axis=2), 1)


def test_sphere_regions_by_centroid():
    mesh, params, scans, _ = configs.sphere(cells=16, photons=10)
    _, _, regions = mesh.soa()
    frac_inside = (regions == 1).mean()
    # sphere of r = 0.9 mm in a 2 mm cube: volume fraction 4/3 pi 0.9^3 / 8 = 0.382
    assert abs(frac_inside - 0.3817) < 0.03
    assert set(np.unique(regions)) == {1, 2}


def test_root_at_a_grid_vertex_is_a_tet_through_the_probe():
    mesh, params, scans, _ = configs.sphere(cells=16, photons=10)
    planes, links, _ = mesh.soa()
    r, z = mesh.find_root(0.0, 0.0, 0.0)
    assert z == 0.0
    d = planes[r, :, :3] @ np.zeros(3) + planes[r, :, 3]
    assert np.all(d >= -1e-12)
    on = np.abs(d) <= 1e-12
    assert np.any(on & (links[r] < 0))           # on a boundary face
    assert np.all(planes[r, on, 2] >= 0)          # +z does not leave the tet at once


def test_probe_off_the_surface_is_rejected():
    mesh = O.Mesh.generate(O.GEOM_HOMOGENEOUS, (4, 4, 4), (0.1, 0.1, 0.1))
    with pytest.raises(O.OctmpsError) as e:
        mesh.find_root(0.0, 0.0, 0.05)  # inside the tissue, not on the surface
    assert e.value.status == 4
    with pytest.raises(O.OctmpsError) as e:
        mesh.find_root(0.5, 0.0, 0.0)  # outside the mesh
    assert e.value.status == 4


def test_probe_projection_onto_a_curved_surface():
    m = O.Mesh.generate(O.GEOM_BUMP, (24, 24, 12), (0.6, 0.6, 0.3), p=(0.03, 0.05, 0.02, 0.07))
    planes, links, _ = m.soa()
    for x in (0.0, 0.011, 0.05, -0.2):
        r, z = m.find_root(x, 0.003, 0.0, project=True)
        zs = -0.03 * np.exp(-(x * x + 0.003 ** 2) / 0.05 ** 2)
        assert abs(z - zs) < 5e-3  # piecewise-planar surface of 250 um cells
        d = planes[r, :, :3] @ np.array([x, 0.003, z]) + planes[r, :, 3]
        assert np.all(d >= -1e-9) and np.any((np.abs(d) <= 1e-9) & (links[r] < 0))


def test_mesh_file_round_trip(tmp_path):
    m = O.Mesh.generate(O.GEOM_LAYERS, (6, 5, 4), (0.3, 0.25, 0.2), p=(0.05, 0.1), n_layers=3)
    p = str(tmp_path / "m.mesh")
    m.write(p)
    m2 = O.Mesh.read(p)
    a, b = m.soa(), m2.soa()
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_input_errors(tmp_path):
    with pytest.raises(O.OctmpsError) as e:
        O.Mesh.read(str(tmp_path / "missing.mesh"))
    assert e.value.status == 2
    bad = tmp_path / "bad.mesh"
    bad.write_text("3\n1\n0 0 0\n1 0 0\n")
    with pytest.raises(O.OctmpsError) as e:
        O.Mesh.read(str(bad))
    assert e.value.status == 3
    with pytest.raises(O.OctmpsError) as e:  # region 0 is the ambient medium
        O.Mesh.build(np.eye(4, 3), np.array([[0, 1, 2, 3, 0]], np.int32))
    assert e.value.status == 5
    with pytest.raises(O.OctmpsError) as e:
        O.read_opt(str(tmp_path / "none.opt"))
    assert e.value.status == 2
    opt = tmp_path / "x.opt"
    opt.write_text("1.0\n1\nout.mco A\n100\n")
    with pytest.raises(O.OctmpsError) as e:
        O.read_opt(str(opt))
    assert e.value.status == 3


def test_tally_finalize_formula():
    B = 8
    rng = np.random.default_rng(0)
    n = 1000
    s = rng.integers(0, 2 ** 40, (2, B)).astype(np.float64) * 2.0 ** -20
    q = np.floor(s ** 2 / n * rng.uniform(1, 3, (2, B)) * 2 ** 20) * 2.0 ** -20
    mx = (s * rng.uniform(0, 0.03, (2, B))).astype(np.float32).astype(np.float64)
    t = make_tally(s, q, mx, np.zeros((2, B), np.uint64), [0, 0, 0])
    out = O.finalize(t, B, n, 0.01)
    for c in range(2):
        mean = s[c] / n
        se = np.sqrt((q[c] / n - mean ** 2) / (n - 1))
        assert np.allclose(out[0, c, 0], mean, rtol=1e-15) and np.allclose(out[0, c, 1], se, rtol=1e-12, equal_nan=True)
        cut = mx[c] > 0.01 * s[c]
        s2, q2 = np.where(cut, s[c] - mx[c], s[c]), np.where(cut, q[c] - mx[c] * mx[c], q[c])
        assert np.allclose(out[1, c, 0], s2 / n, rtol=1e-15)
        assert np.allclose(out[1, c, 1], np.sqrt((q2 / n - (s2 / n) ** 2) / (n - 1)), rtol=1e-12, equal_nan=True)


def test_safeprimes_file_override(tmp_path):
    # A reference-format table (first column used, src/octmps_kernel.cuh:95-105).
    m = O.safeprimes(3 + 3 * 8)
    f = tmp_path / "safeprimes_base32.txt"
    f.write_text("".join(f"{a} 0 0\n" for a in m))
    try:
        O.check(O.lib.octmps_set_stream_triplets(8), "triplets")
        O.check(O.lib.octmps_set_safeprimes_file(str(f).encode()), "file")
        x, a = O.photon_streams(1, 0, 9)
        assert list(map(int, a)) == [int(v) for v in m[3 + 3 * 1:6 + 3 * 1]]
    finally:
        O.check(O.lib.octmps_set_safeprimes_file(None), "file")
        O.check(O.lib.octmps_set_stream_triplets(4096), "triplets")
    x, a = O.photon_streams(1, 0, 9)
    assert int(a[0]) == int(O.safeprimes(3 + 3 * 10)[3 + 3 * 9])


def test_benchmark_workloads_match_committed_traces():
    # The ellipsoid B-scan's per-position budgets and the served request traces that the
    # benchmark runs, against the committed fixture (tests/golden/make_workload_traces.py):
    # a change of generator or of numpy's random streams must not silently change them.
    import json

    with open(os.path.join(os.path.dirname(__file__), "golden", "workload_traces.json")) as f:
        d = json.load(f)
    b = configs.ellipsoid_budgets()
    assert [int(x) for x in b] == d["ellipsoid_budgets_seed7"]
    assert len(b) == 512 and b.min() >= 1e5 and b.max() <= 1e7
    # log-uniform on [1e5, 1e7]: mean (1e7 - 1e5) / ln 100 = 2.15e6, log-mean at 1e6
    assert abs(b.mean() / 2.15e6 - 1) < 0.1 and abs(np.log(b).mean() - np.log(1e6)) < 0.25
    for s in d["served_seed11"]:
        got = configs.served_stream(s["rate"], s["window"])
        assert len(got) == len(s["requests"])
        for (t, m, n, bud), r in zip(got, s["requests"]):
            assert t == r["t"] and m == r["mesh"] and n == r["n_scans"] and [int(x) for x in bud] == r["budgets"]
            assert 0 <= m < 4 and 1 <= n <= 256 and len(bud) == n


def test_single_contribution_bin_has_zero_filtered_se():
    # A bin holding one contribution c: the filtered outputs drop it (c > MaxRelContrib * c),
    # leaving Sum - Max = 0 and SumSq - Max^2 = 0 exactly, so the filtered SE is 0 (the
    # reference's FP64 arithmetic gives 0 there too), never NaN. Contributions span the FP32
    # range, subnormals included.
    cs = np.array([1.0, 0.3, 1e-16, 3.6e-15, 5e-20, 2e-40, 1e5, 4e9, 1e30], np.float32).astype(np.float64)
    B = len(cs)
    z = np.zeros((2, B))
    sums = np.stack([cs, z[0]])
    t = make_tally(sums, sums * sums, sums.copy(), np.stack([np.ones(B, np.uint64), np.zeros(B, np.uint64)]),
                   [B, B, 0])
    n = 1000
    out = O.finalize(t, B, n, 0.01)
    assert np.array_equal(out[1, 0, 0], np.zeros(B)) and np.array_equal(out[1, 0, 1], np.zeros(B))
    assert np.all(np.isfinite(out[0, 0, 1])) and np.all(out[0, 0, 1] > 0)
    assert np.allclose(out[0, 0, 0], cs / n, rtol=1e-15, atol=0)
    # SE of one contribution c in N photons: sqrt((c^2/N - c^2/N^2) / (N-1)) = c / N.
    assert np.allclose(out[0, 0, 1], cs / n, rtol=1e-12, atol=0)


def test_finalize_sums_over_the_whole_fp32_range():
    # Sums of many FP32 contributions from 1e-40 to 1e38 are held exactly; the FP64 mean is
    # the correctly rounded exact sum divided by N (to an ulp), whatever the spread.
    from fractions import Fraction
    rng = np.random.default_rng(5)
    B = 6
    cs = [np.float32(10.0) ** rng.uniform(lo, hi, 40).astype(np.float32)
          for lo, hi in ((-40, -30), (-20, -10), (-5, 0), (0, 5), (10, 20), (30, 38))]
    exact_s = [sum(Fraction(float(c)) for c in row) for row in cs]
    exact_q = [sum(Fraction(float(c)) ** 2 for c in row) for row in cs]
    maxs = np.array([max(float(c) for c in row) for row in cs])
    sums = np.array([[float(s) for s in exact_s], [0.0] * B])
    # make_tally needs exactly representable sums: encode the exact Fractions directly.
    t = make_tally(np.zeros((2, B)), np.zeros((2, B)), np.stack([maxs, np.zeros(B)]),
                   np.zeros((2, B), np.uint64), [0, 0, 0])
    W = O.TALLY_CLASS_WORDS
    for b in range(B):
        S = exact_s[b] * 2 ** -O.TALLY_SUM_EXP
        Q = exact_q[b] * 2 ** -O.TALLY_SQ_EXP
        assert S.denominator == 1 and Q.denominator == 1
        for k in range(O.TALLY_SUM_DIGITS):
            t[4 + k * B + b] = (int(S) >> (32 * k)) & 0xFFFFFFFF
        for k in range(O.TALLY_SQ_DIGITS):
            t[4 + (O.TALLY_SUM_DIGITS + k) * B + b] = (int(Q) >> (32 * k)) & 0xFFFFFFFF
    n = 10 ** 6
    out = O.finalize(t, B, n, 0.01)
    for b in range(B):
        mean = float(exact_s[b] / n)
        assert abs(out[0, 0, 0, b] - mean) <= 2 * np.spacing(mean)
        var = float((exact_q[b] / n - (exact_s[b] / n) ** 2) / (n - 1))
        assert abs(out[0, 0, 1, b] - np.sqrt(var)) <= 1e-12 * np.sqrt(var)
        # filtered: the largest contribution removed exactly (it exceeds 1 % of these sums)
        if maxs[b] > 0.01 * float(exact_s[b]):
            fs = exact_s[b] - Fraction(maxs[b])
            assert abs(out[1, 0, 0, b] - float(fs / n)) <= 2 * np.spacing(float(fs / n)) + 1e-300
    assert np.allclose(sums[0], [float(s) for s in exact_s])
