"""Host-side logic: structured meshes, root finding, error codes, tally finalisation."""
import ctypes as C
import os

import numpy as np
import pytest

import octmps as O
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
    t = np.zeros(O.tally_words(B), np.uint64)
    n = 1000
    vals = {}
    for c in range(2):
        base = 4 + c * 7 * B
        s = rng.integers(0, 2 ** 40, B).astype(np.uint64)
        q = (s.astype(np.float64) ** 2 / 2 ** 20 / n * rng.uniform(1, 3, B)).astype(np.uint64)
        # value s * 2^-20: high word (units 2^-32) = s * 2^12 - (low part), low word (2^-64 units)
        t[base + B:base + 2 * B] = (s >> np.uint64(4)) << np.uint64(16)
        t[base:base + B] = (s & np.uint64(15)) << np.uint64(44)
        t[base + 3 * B:base + 4 * B] = (q >> np.uint64(4)) << np.uint64(16)
        t[base + 2 * B:base + 3 * B] = (q & np.uint64(15)) << np.uint64(44)
        mx = (s.astype(np.float64) * 2.0 ** -20) * rng.uniform(0, 0.03, B)
        t[base + 5 * B:base + 6 * B] = mx.view(np.uint64)
        vals[c] = (s.astype(np.float64) * 2.0 ** -20, q.astype(np.float64) * 2.0 ** -20, mx)
    out = O.finalize(t, B, n, 0.01)
    for c in range(2):
        s, q, mx = vals[c]
        mean = s / n
        se = np.sqrt((q / n - mean ** 2) / (n - 1))
        assert np.allclose(out[0, c, 0], mean, rtol=1e-15) and np.allclose(out[0, c, 1], se, rtol=1e-12, equal_nan=True)
        cut = mx > 0.01 * s
        s2, q2 = np.where(cut, s - mx, s), np.where(cut, q - mx * mx, q)
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
