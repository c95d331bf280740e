"""GPU parity of the transport path against the FP64 CPU oracle (device 0)."""
import numpy as np
import pytest

import octmps as O
import oracle_py as OR
from octmps import configs
from stats_util import chi2_decimated, compare

pytestmark = pytest.mark.gpu


def test_rng_streams_bit_exact(ctx, table):
    # Per-photon stream seeding and MWC draws of the device equal the oracle's bit for bit.
    seed, salt, p0, n, d = 1, 5, 1000, 64, 16
    dev = ctx.probe_rng(seed, salt, p0, n, d)
    for i in range(n):
        x, a = OR.photon_streams(seed, salt, p0 + i, table)
        for s in range(3):
            ref, _ = OR.rand_co(int(x[s]), int(a[s]), d)
            assert np.array_equal(dev[i, s].astype(np.float64), ref), (i, s)


def test_fresnel_device_vs_oracle(ctx):
    rng = np.random.default_rng(3)
    ni = rng.uniform(1.0, 1.6, 4000).astype(np.float32)
    nt = rng.uniform(1.0, 1.6, 4000).astype(np.float32)
    ca1 = rng.uniform(0.0, 1.0, 4000).astype(np.float32)
    ca1[:50] = 1.0
    r = ctx.probe_fresnel(ni, nt, ca1)
    ref = np.array([OR.fresnel(float(a), float(b), float(c)) for a, b, c in zip(ni, nt, ca1)])
    # Away from the critical angle, where R jumps to 1, FP32 agrees to ~1e-5 absolute.
    sa1 = np.sqrt(np.maximum(1 - ca1.astype(np.float64) ** 2, 0))
    crit = np.abs(ni.astype(np.float64) / nt * sa1 - 1.0) < 1e-3
    assert np.max(np.abs(r[~crit] - ref[~crit])) < 2e-5


@pytest.mark.parametrize("kind,param", [(0, 0.9), (0, 0.0), (0, 0.95), (1, 0.925), (1, 0.5)])
def test_samplers_device_vs_oracle(ctx, kind, param):
    u = (np.arange(1, 4097, dtype=np.float64) / 4096.0).astype(np.float32)
    dev = ctx.probe_sampler(kind, param, u)
    f = OR.spin_theta if kind == 0 else OR.spin_theta_bias
    ref = np.array([f(param, float(v)) for v in u])
    assert np.max(np.abs(dev - ref)) < 5e-5


def _run_pair(ctx, table, mesh, params, scan, seed, threads=8):
    mid = ctx.upload(mesh)
    t, st, plog = ctx.run_bscan(mid, params, [scan], seed, log=True)
    gpu = O.decode_tally(t[0], params.num_depth_bins)
    ref = OR.run_ascan(mesh.soa(), params, scan, seed, table, threads=threads, log=True)
    return gpu, st, plog, ref


def test_sphere_parity_paired(ctx, table):
    n = 20000
    mesh, params, scans, seed = configs.sphere(cells=16, photons=n)
    gpu, st, plog, ref = _run_pair(ctx, table, mesh, params, scans[0], seed)
    assert st.photons == n and st.killed == 0 and ref["killed"] == 0
    print("steps gpu/ref", st.photon_steps, ref["steps"])
    same_steps = np.mean(plog["steps"].astype(np.float64) == ref["log"][:, 0])
    same_det = np.mean(plog["detections"].astype(np.float64) == ref["log"][:, 1])
    print("paired photons: same steps", same_steps, "same detections", same_det)
    c = compare(gpu, ref, n)
    print(c)
    # Paired streams: FP32 and FP64 photons start identical; a near-edge crossing decided
    # differently changes the step count without changing the tallies much.
    assert same_steps > 0.5 and same_det > 0.85
    assert c["frac_3sigma"] <= 0.01 and c["max_abs_z"] < 5
    assert c["R_rel"] < 0.005 and c["Rfilt_rel"] < 0.005


@pytest.mark.parametrize("name", ["slab4", "ellipsoid"])
def test_layered_and_inclusion_parity_paired(ctx, table, name):
    # Four index-mismatched layers (Fresnel at every layer crossing) and an ellipsoid with two
    # sphere inclusions (three regions, curved interfaces), on coarser meshes of the same
    # geometry so the FP64 oracle finishes in seconds. One A-scan at the centre.
    n = 12000 if name == "slab4" else 40000
    if name == "slab4":
        mesh, params, scans, seed = configs.slab4(cells=(32, 32, 16), n_scans=1, photons=n, bins=256)
    else:
        mesh, params, scans, seed = configs.ellipsoid(cells=(32, 32, 16), n_scans=1, budgets=[n], bins=256)
    gpu, st, plog, ref = _run_pair(ctx, table, mesh, params, scans[0], seed, threads=16)
    assert st.photons == n and st.killed == 0 and ref["killed"] == 0
    c = compare(gpu, ref, n)
    print(name, c)
    assert c["frac_3sigma"] <= 0.01 and c["max_abs_z"] < 5
    assert c["R_rel"] < 0.005 and c["Rfilt_rel"] < 0.005


def _layered_grid(n=(16, 16, 40), size=(0.2, 0.2, 0.1)):
    """Cells split into 6 tets around the cell diagonal (conforming), region = layer k + 1."""
    nx, ny, nz = n
    xs = np.linspace(-size[0] / 2, size[0] / 2, nx + 1)
    ys = np.linspace(-size[1] / 2, size[1] / 2, ny + 1)
    zs = np.linspace(0.0, size[2], nz + 1)
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    xyz = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)

    def vid(i, j, k):
        return (k * (ny + 1) + j) * (nx + 1) + i

    tets = []
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                for order in ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)):
                    c = [0, 0, 0]
                    vs = [vid(i, j, k)]
                    for a in order:
                        c[a] += 1
                        vs.append(vid(i + c[0], j + c[1], k + c[2]))
                    tets.append(vs + [k + 1])
    return xyz, np.array(tets, np.int32)


def test_many_regions_escaped_links_parity_paired(ctx, table):
    # 40 index-mismatched layers, 25 um thick (regions 1-40). A neighbour word has 5 bits
    # for the neighbour's region, so links into regions >= 31 carry the escape value: the
    # kernel then reads the region from the neighbour's record and fetches all four planes
    # of the tet it enters. Photons reach those layers (the bottom quarter of the mesh).
    n = 12000
    xyz, tets = _layered_grid()
    mesh = O.Mesh.build(xyz, tets)
    regs = [configs._reg(1.33 + 0.004 * (i % 7), 0.05, 8.0, 0.9) for i in range(40)]
    params = O.make_params(regs, num_bins=256, depth=0.06, **configs.BIAS)
    root, z = mesh.find_root(0.0, 0.0, 0.0)
    scan = O.ascan(0.0, 0.0, z, root, n)
    gpu, st, plog, ref = _run_pair(ctx, table, mesh, params, scan, 1, threads=16)
    assert st.photons == n and st.killed == 0 and ref["killed"] == 0
    same_steps = np.mean(plog["steps"].astype(np.float64) == ref["log"][:, 0])
    c = compare(gpu, ref, n)
    print("escaped links", c, "same steps", same_steps)
    assert same_steps > 0.5
    assert c["frac_3sigma"] <= 0.01 and c["max_abs_z"] < 5
    assert c["R_rel"] < 0.005 and c["Rfilt_rel"] < 0.005
    mid = ctx.upload(mesh)
    s2 = O.ascan(0.0, 0.0, z, root, 1500)
    tb, _ = ctx.run_bscan(mid, params, [s2, s2], 7)
    ta, _ = ctx.run_ascan(mid, params, s2, 7)
    assert np.array_equal(ta, tb[0])


def test_finite_next_step_parity_paired(ctx, table):
    # rndStepSizeInTissue = 0.05 (all other cases use 0, an infinite next step): a photon
    # biased before is turned towards the probe (F = 1) only when -ln(0.05)/mu_t (0.30 mm in
    # the sphere) reaches the probe and the probe is inside the cone, else it takes the
    # mixture (F = 2). In oracle traces of 300 photons, ~1,000 scatterings take F = 1 by
    # that test and ~600 more take F = 2 than with the infinite step: both sides run.
    n = 20000
    mesh, params, scans, seed = configs.sphere(cells=16, photons=n)
    params.rnd_step_size_in_tissue = 0.05
    gpu, st, plog, ref = _run_pair(ctx, table, mesh, params, scans[0], seed)
    assert st.photons == n and st.killed == 0 and ref["killed"] == 0
    assert ref["num_filtered"] > 100
    same_steps = np.mean(plog["steps"].astype(np.float64) == ref["log"][:, 0])
    c = compare(gpu, ref, n)
    print("finite next step", c, "same steps", same_steps)
    assert same_steps > 0.5
    assert c["frac_3sigma"] <= 0.01 and c["max_abs_z"] < 5
    assert c["R_rel"] < 0.005 and c["Rfilt_rel"] < 0.005


def test_bump_slope_parity_paired(ctx, table):
    # A probe on the bump's flank (x = 0.35 mm, surface tilted ~27 deg): the root tet comes
    # from projecting the probe onto the surface, photons enter and leave through tilted
    # boundary faces, and Fresnel and the detector's cone see non-axis-aligned planes. A
    # wide detector (0.5 mm, 30 deg) collects light leaving the slope. Its likelihood
    # ratios are heavy-tailed: ten oracle runs of 16,000 photons with different seeds give
    # total R with a 33 % relative spread, so a fixed 0.5 % bound on total R is far below
    # the noise of any affordable run. Total R is tested instead through the paired
    # per-photon contributions (GPU minus oracle, photon by photon), whose mean must lie
    # within 5 standard errors of zero; per-bin z scores as in the other cases.
    n = 16000
    mesh, params, _, seed = configs.bump(cells=(48, 48, 24), n_scans=1, photons=1, bins=256)
    params.max_collecting_radius = 0.05
    params.max_collecting_angle_deg = 30.0
    x = 0.35 * configs.MM
    root, z = mesh.find_root(x, 0.0, 0.0, project=True)
    assert z < -0.015  # on the flank, well above z = 0
    scan = O.ascan(x, 0.0, z, root, n)
    gpu, st, plog, ref = _run_pair(ctx, table, mesh, params, scan, seed, threads=16)
    assert st.photons == n and st.killed == 0 and ref["killed"] == 0
    assert ref["num_filtered"] > 1000
    same_steps = np.mean(plog["steps"].astype(np.float64) == ref["log"][:, 0])
    c = compare(gpu, ref, n)
    d = plog["contrib"].astype(np.float64) - ref["log"][:, 2]
    z_total = d.sum() / (d.std(ddof=1) * np.sqrt(n))
    print("bump slope", c, "same steps", same_steps, "paired total z", z_total)
    assert same_steps > 0.5
    assert c["frac_3sigma"] <= 0.01 and c["max_abs_z"] < 5
    assert abs(z_total) < 5 and c["R_rel"] < 0.05


@pytest.mark.parametrize("type_bias", [0, 3])
def test_unbiased_and_no_spin_parity_paired(ctx, table, type_bias):
    # TypeBias 0 scatters with Spin (Henyey-Greenstein only) and accepts every photon that
    # reaches the detector; TypeBias 3 does not turn photons at interaction sites (the
    # reference's switch has no case for it) and never restores a continuation photon.
    # A wide detector (0.5 mm, 30 deg) gives enough detections without the bias.
    n = 12000
    mesh, params, scans, seed = configs.sphere(cells=16, photons=n)
    params.type_bias = type_bias
    params.max_collecting_radius = 0.05
    params.max_collecting_angle_deg = 30.0
    gpu, st, plog, ref = _run_pair(ctx, table, mesh, params, scans[0], seed, threads=16)
    assert st.photons == n and st.killed == 0 and ref["killed"] == 0
    assert ref["num_filtered"] > 100
    c = compare(gpu, ref, n)
    print(type_bias, c)
    assert c["frac_3sigma"] <= 0.01 and c["max_abs_z"] < 5
    assert c["R_rel"] < 0.005 and c["Rfilt_rel"] < 0.005


def test_independent_runs_chi_square(ctx, table):
    # Unpaired: the GPU and the oracle draw different photons (different stream salts), so
    # per-bin differences are pure Monte Carlo noise and chi^2/ndf must sit inside the 99 %
    # interval of its distribution. Bins 6 apart are independent (a detection spans 6).
    # Unbiased transport (TypeBias 0, wide detector): with the backward bias the likelihood
    # ratios are so heavy-tailed that sample variances, and so chi^2, are dominated by rare
    # photons (two oracle runs of 40,000 photons differ by 2.6 of their own standard errors in
    # mean detected weight); the biased path is compared paired instead.
    from scipy.stats import chi2
    mesh, params, scans, seed = configs.sphere(cells=16, photons=1)
    params.type_bias = 0
    params.max_collecting_radius = 0.05
    params.max_collecting_angle_deg = 30.0
    s = scans[0]
    gscan = O.ascan(s.x, s.y, s.z, s.root_tet, 400000, 101)
    rscan = O.ascan(s.x, s.y, s.z, s.root_tet, 100000, 202)
    mid = ctx.upload(mesh)
    t, st = ctx.run_ascan(mid, params, gscan, seed)
    gpu = O.decode_tally(t, params.num_depth_bins)
    ref = OR.run_ascan(mesh.soa(), params, rscan, seed, table, threads=16)
    x2, ndf = chi2_decimated(gpu, ref, 400000, 100000)
    lo, hi = chi2.ppf(0.005, ndf), chi2.ppf(0.995, ndf)
    print("chi2", x2, "ndf", ndf, "99% interval", lo, hi)
    assert ndf > 40 and lo <= x2 <= hi


def test_glass_slab_known_answer(ctx):
    # One glass region (mus = 0: no scattering) with n = 1.5 under air. Every photon travels
    # straight down and up; at the first return it is detected with weight
    # c1 = (1-R)^2 R, R = ((1.5-1)/(1.5+1))^2, at optical depth d (Class I, six bins).
    d, h, cells = 0.1, 0.1 / 8, 8
    mesh = O.Mesh.generate(O.GEOM_HOMOGENEOUS, (cells, cells, cells), (cells * h, cells * h, d))
    params = O.make_params([(1.5, 0.0, 0.0, 0.9)], num_bins=900, depth=0.27)
    x0, y0 = 0.13 * h, 0.07 * h
    root, z = mesh.find_root(x0, y0, 0.0)
    n = 4000
    scan = O.ascan(x0, y0, z, root, n)
    mid = ctx.upload(mesh)
    t, st = ctx.run_ascan(mid, params, scan, 1)
    g = O.decode_tally(t, params.num_depth_bins)
    R = ((1.5 - 1.0) / 2.5) ** 2
    c1 = (1 - R) ** 2 * R
    w = params.coherence_length_source / 6
    vp = int(np.floor(d / w))
    assert g["num_class1"] == n
    cls1 = g["sum"][0] / n
    for b in range(vp - 3, vp + 3):
        assert g["count"][0][b] == n
        assert abs(cls1[b] - c1) < 1e-6 * c1
    # Second return (Class II): expected (1-R)^2 R^3 per photon after unbiased roulette.
    vp2 = int(np.floor(2 * d / w))
    m2 = g["sum"][1][vp2] / n
    e2 = (1 - R) ** 2 * R ** 3
    se = np.sqrt((g["sumsq"][1][vp2] / n - m2 * m2) / (n - 1))
    assert abs(m2 - e2) < 5 * se + 1e-12


def test_batched_equals_plain_bitwise(ctx):
    mesh, params, scans, seed = configs.slab4(cells=(16, 16, 8), n_scans=3, photons=3000, bins=128)
    mid = ctx.upload(mesh)
    tb, stb = ctx.run_bscan(mid, params, scans, seed)
    for i, s in enumerate(scans):
        ta, sta = ctx.run_ascan(mid, params, s, seed)
        assert np.array_equal(ta, tb[i]), i


def test_graph_replay_matches_eager(ctx):
    mesh, params, scans, seed = configs.slab4(cells=(16, 16, 8), n_scans=4, photons=2000, bins=128)
    mid = ctx.upload(mesh)
    te, _ = ctx.run_bscan(mid, params, scans, seed)
    o = O.BatchOpts()
    o.use_graph = 1
    for _ in range(3):
        tg, _ = ctx.run_bscan(mid, params, scans, seed, opts=o)
        assert np.array_equal(te, tg)
        ta, _ = ctx.run_ascan(mid, params, scans[0], seed)  # eager work between replays
        assert np.array_equal(ta, te[0])


@pytest.mark.parametrize("sq,q,block", [(1, 1, 64), (16, 8, 128), (64, 32, 256), (48, 4, 192)])
def test_launch_shape_never_changes_results(ctx, sq, q, block):
    # Scatter quantum, refill quantum and block size only reorder work between lanes;
    # every photon runs the same operations on the same streams.
    mesh, params, scans, seed = configs.sphere(cells=12, photons=3000, bins=96)
    mid = ctx.upload(mesh)
    ref, _ = ctx.run_bscan(mid, params, scans, seed)
    o = O.BatchOpts()
    o.scatter_quantum, o.refill_quantum, o.block_size = sq, q, block
    got, st = ctx.run_bscan(mid, params, scans, seed, opts=o)
    assert st.block_size == block and st.refill_quantum == q
    assert np.array_equal(ref, got)


@pytest.mark.parametrize("name", ["sphere_full", "slab4_full", "ellipsoid_full", "bump_full"])
def test_full_size_benchmark_meshes_parity_paired(ctx, table, name):
    # The benchmark meshes themselves, not coarsened: the sphere (32^3 cells, 196,608 tets),
    # the four-layer slab and the ellipsoid with two spheres (64x64x32 cells, 786,432 tets,
    # 101 MB of records, kernel with the entry-plane skip) and the bump (128x128x64 cells,
    # 6,291,456 tets, 805 MB: larger than the Infinity Cache, so the launch takes the variant
    # without the skip). One A-scan at the centre, 10,000 paired photons (20,000 for the
    # sphere and 40,000 for the ellipsoid, as on their coarser meshes above).
    n = {"sphere_full": 20000, "ellipsoid_full": 40000}.get(name, 10000)
    if name == "sphere_full":
        mesh, params, scans, seed = configs.sphere(photons=n)
        assert mesh.num_tets == 196608
    elif name == "slab4_full":
        mesh, params, scans, seed = configs.slab4(n_scans=1, photons=n)
        assert mesh.num_tets == 786432
    elif name == "ellipsoid_full":
        mesh, params, scans, seed = configs.ellipsoid(n_scans=1, budgets=[n])
        assert mesh.num_tets == 786432
    else:
        mesh, params, scans, seed = configs.bump(n_scans=1, photons=n)
        assert mesh.num_tets == 6291456
    gpu, st, plog, ref = _run_pair(ctx, table, mesh, params, scans[0], seed, threads=16)
    assert st.photons == n and st.killed == 0 and ref["killed"] == 0 and st.bad_contributions == 0
    assert ref["num_filtered"] > 100
    same_steps = np.mean(plog["steps"].astype(np.float64) == ref["log"][:, 0])
    c = compare(gpu, ref, n)
    print(name, c, "same steps", same_steps, "steps/photon", st.photon_steps / n)
    # Pairing check: two unpaired photons have the same step count with probability ~1e-4
    # (counts spread over ~1e4 values). On these fine meshes a photon takes ~1e4-3e4 steps, so
    # its FP32 and FP64 paths split at some near-edge crossing more often than on the coarse
    # meshes above: 0.45 of photons keep the same count on the sphere, 0.35 on the slab and
    # 0.16 on the ellipsoid, whose curved index-mismatched interfaces reflect more often.
    assert same_steps > 0.1
    assert c["frac_3sigma"] <= 0.01 and c["max_abs_z"] < 5
    assert c["R_rel"] < 0.005 and c["Rfilt_rel"] < 0.005
