"""GPU tests of the batched and served paths, the CLI and whole-path properties."""
import os
import subprocess
import time

import numpy as np
import pytest

import octmps as O
from octmps import configs
from stats_util import bin_stats

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_meshes(ctx):
    out = {}
    m, p, s, seed = configs.slab4(cells=(16, 16, 8), n_scans=4, photons=1500, bins=128)
    out["slab"] = (m, ctx.upload(m), p, s, seed)
    m, p, s, seed = configs.sphere(cells=12, photons=1200, bins=96)
    out["sphere"] = (m, ctx.upload(m), p, s, seed)
    m, p, s, seed = configs.ellipsoid(cells=(16, 16, 8), n_scans=6, budgets=[300, 2000, 700, 50, 1200, 900],
                                      bins=128)
    out["ellipsoid"] = (m, ctx.upload(m), p, s, seed)
    return out


def test_uneven_budgets_batched_equals_plain(ctx, small_meshes):
    m, mid, p, scans, seed = small_meshes["ellipsoid"]
    tb, st = ctx.run_bscan(mid, p, scans, seed)
    assert st.photons == sum(s.n_photons for s in scans) and st.killed == 0
    for i, s in enumerate(scans):
        ta, _ = ctx.run_ascan(mid, p, s, seed)
        assert np.array_equal(ta, tb[i])


@pytest.mark.parametrize("copies,graph", [(2, 0), (5, 1), (16, 1)])
def test_tally_privatisation_depth_never_changes_results(ctx, small_meshes, copies, graph):
    # k private tally copies, folded after the launch: integer sums and maxima are exact,
    # so every depth gives the same words as one copy.
    m, mid, p, scans, seed = small_meshes["ellipsoid"]
    one = O.BatchOpts()
    one.tally_copies = 1
    ref, _ = ctx.run_bscan(mid, p, scans, seed, opts=one)
    o = O.BatchOpts()
    o.tally_copies, o.use_graph = copies, graph
    for _ in range(2):
        got, _ = ctx.run_bscan(mid, p, scans, seed, opts=o)
        assert np.array_equal(ref, got)


@pytest.mark.parametrize("entry_skip", [1, 2])
def test_entry_plane_skip_never_changes_results(ctx, small_meshes, entry_skip):
    # Skipping the plane of the face a photon entered through (per-lane gathers) removes a
    # face that cannot be its exit: the same bits with the skip forced on or off, as with the
    # default.
    for name in ("slab", "ellipsoid"):
        m, mid, p, scans, seed = small_meshes[name]
        ref, _ = ctx.run_bscan(mid, p, scans, seed)
        o = O.BatchOpts()
        o.record_fetch, o.entry_skip = 1, entry_skip
        got, _ = ctx.run_bscan(mid, p, scans, seed, opts=o)
        assert np.array_equal(ref, got), name


@pytest.mark.parametrize("name", ["slab", "sphere", "ellipsoid"])
def test_record_fetch_paths_give_the_same_bits(ctx, small_meshes, name):
    # The cooperative fetch (four lanes read one record's planes into LDS, and a record stays
    # staged while its photon stays in the tet) reads the same floats as the per-lane gathers:
    # the same tallies, step counts and per-photon summaries, bit for bit, in plain and logged
    # launches and at every block size the launch accepts.
    m, mid, p, scans, seed = small_meshes[name]
    runs = []
    for fetch, skip, block in ((1, 2, 256), (1, 1, 256), (2, 0, 256), (2, 0, 64), (2, 0, 128), (2, 0, 192),
                               (0, 0, 0)):
        o = O.BatchOpts()
        o.record_fetch, o.entry_skip, o.block_size = fetch, skip, block
        t, st = ctx.run_bscan(mid, p, scans, seed, opts=o)
        tl, stl, rec = ctx.run_bscan(mid, p, scans, seed, opts=o, log=True)
        assert np.array_equal(t, tl) and st.photon_steps == stl.photon_steps
        runs.append((t, st.photon_steps, rec))
    for t, steps, rec in runs[1:]:
        assert np.array_equal(t, runs[0][0]) and steps == runs[0][1]
        assert np.array_equal(rec, runs[0][2])


@pytest.mark.parametrize("field,value", [("tally_copies", 17), ("entry_skip", 3), ("entry_skip", -1),
                                         ("record_fetch", 3), ("record_fetch", -1),
                                         ("block_size", 96), ("block_size", 512)])
def test_invalid_launch_options_are_rejected(ctx, small_meshes, field, value):
    m, mid, p, scans, seed = small_meshes["ellipsoid"]
    o = O.BatchOpts()
    setattr(o, field, value)
    with pytest.raises(O.OctmpsError) as e:
        ctx.run_bscan(mid, p, scans, seed, opts=o)
    assert e.value.status == 1


def test_served_results_equal_batched_alone_and_in_a_mix(ctx, small_meshes):
    ref = {k: ctx.run_bscan(v[1], v[2], v[3], v[4])[0] for k, v in small_meshes.items()}
    srv = O.Server(ctx, slice_photons=4096)
    try:
        # alone
        _, mid, p, s, seed = small_meshes["slab"]
        info, out = srv.wait(srv.submit(mid, p, s, seed), 120)
        assert info.status == 0 and np.array_equal(out, ref["slab"])
        # in a mix of meshes, parameter sets and budgets, submitted together
        ids = {k: srv.submit(v[1], v[2], v[3], v[4]) for k, v in small_meshes.items()}
        for k, rid in ids.items():
            info, out = srv.wait(rid, 120)
            assert info.status == 0 and info.done
            assert np.array_equal(out, ref[k]), k
            assert info.t_submit_ms <= info.t_admit_ms <= info.t_done_ms
        st = srv.stats()
        assert st["slices"] > 1 and st["photons"] > 0
    finally:
        srv.close()


@pytest.mark.parametrize("record_fetch", [0, 1])
def test_served_sets_beyond_the_lds_cache_equal_batched(ctx, small_meshes, record_fetch):
    # A slice keeps the parameter sets in LDS only while at most 16 requests are in flight
    # and their meshes' regions fit the LDS table; otherwise the kernel reads each lane's set
    # from memory. Both give the batched bits, with the cooperative record fetch (0) and the
    # per-lane gathers (1): 20 requests with 20 different parameter sets submitted at once
    # (slices switch between the two as requests end), then a 40-region mesh beside a small
    # one.
    from test_gpu_core import _layered_grid

    _, mid, p, s, seed = small_meshes["sphere"]
    sets = []
    for i in range(20):
        q = type(p).from_buffer_copy(p)
        q.regions[1].g = 0.80 + 0.005 * i
        q.regions[2].mus = p.regions[2].mus * (1.0 + 0.02 * i)
        sets.append(q)
    refs = [ctx.run_bscan(mid, q, s, seed + i)[0] for i, q in enumerate(sets)]
    xyz, tets = _layered_grid(n=(8, 8, 40))
    mesh40 = O.Mesh.build(xyz, tets)
    mid40 = ctx.upload(mesh40)
    regs = [configs._reg(1.33 + 0.004 * (i % 7), 0.05, 8.0, 0.9) for i in range(40)]
    p40 = O.make_params(regs, num_bins=128, depth=0.06, **configs.BIAS)
    root, z = mesh40.find_root(0.0, 0.0, 0.0)
    s40 = [O.ascan(0.0, 0.0, z, root, 900), O.ascan(0.02, 0.0, z, mesh40.find_root(0.02, 0.0, 0.0)[0], 700, 1)]
    ref40 = ctx.run_bscan(mid40, p40, s40, 5)[0]
    _, mids, ps, ss, seeds = small_meshes["slab"]
    ref_slab = ctx.run_bscan(mids, ps, ss, seeds)[0]
    srv = O.Server(ctx, slice_photons=4096, record_fetch=record_fetch)
    try:
        ids = [srv.submit(mid, q, s, seed + i) for i, q in enumerate(sets)]
        for i, rid in enumerate(ids):
            info, out = srv.wait(rid, 120)
            assert info.status == 0 and np.array_equal(out, refs[i]), i
        r40, rslab = srv.submit(mid40, p40, s40, 5), srv.submit(mids, ps, ss, seeds)
        info, out = srv.wait(r40, 120)
        assert info.status == 0 and np.array_equal(out, ref40)
        info, out = srv.wait(rslab, 120)
        assert info.status == 0 and np.array_equal(out, ref_slab)
    finally:
        srv.close()


def test_served_slices_switch_kernel_variants_and_equal_batched(ctx, small_meshes):
    # The LDS set table beside the cooperative fetch's stage holds 16 sets of up to 8
    # regions or 8 sets of up to 16; sets that fit neither but fit the per-lane variant's
    # table (16 of 16) run the per-lane gathers. Eleven requests in flight, one of them on an
    # 11-region mesh, start on the per-lane variant; as requests end the slices move to the
    # cooperative 8 x 16 table and, once the 11-region request is done, to 16 x 8 (a long
    # request outlives it). Photons saved by one slice resume in the next whatever variant
    # runs it (a restored photon's tet is fetched again): every request gives the batched bits.
    from test_gpu_core import _layered_grid

    xyz, tets = _layered_grid(n=(8, 8, 10))
    mesh10 = O.Mesh.build(xyz, tets)
    mid10 = ctx.upload(mesh10)
    regs = [configs._reg(1.33 + 0.01 * (i % 5), 0.05, 8.0, 0.9) for i in range(10)]
    p10 = O.make_params(regs, num_bins=128, depth=0.06, **configs.BIAS)
    root, z = mesh10.find_root(0.0, 0.0, 0.0)
    s10 = [O.ascan(0.0, 0.0, z, root, 3000), O.ascan(0.02, 0.0, z, mesh10.find_root(0.02, 0.0, 0.0)[0], 2000, 1)]
    ref10 = ctx.run_bscan(mid10, p10, s10, 9)[0]
    reqs = {(k, j): (v[1], v[2], v[3], v[4] + j) for k, v in small_meshes.items() for j in range(3)}
    _, mids, ps, ss, seeds = small_meshes["slab"]
    reqs[("slab_long", 0)] = (mids, ps, [O.ascan(x.x, x.y, x.z, x.root_tet, 30000, x.stream_salt) for x in ss],
                              seeds + 7)
    refs = {k: ctx.run_bscan(*v)[0] for k, v in reqs.items()}
    srv = O.Server(ctx, slice_photons=2048)
    try:
        ids = {k: srv.submit(*v) for k, v in reqs.items()}
        r10 = srv.submit(mid10, p10, s10, 9)
        info, out = srv.wait(r10, 120)
        assert info.status == 0 and np.array_equal(out, ref10)
        for k, rid in ids.items():
            info, out = srv.wait(rid, 120)
            assert info.status == 0 and np.array_equal(out, refs[k]), k
        assert srv.stats()["slices"] > 2
    finally:
        srv.close()
        ctx.release(mid10)


@pytest.mark.parametrize("block", [64, 128, 192])
def test_served_block_sizes_equal_batched(ctx, small_meshes, block):
    # Blocks of fewer than 256 lanes use fewer of the cooperative fetch's per-wave stages (and
    # a smaller grid per CU): the batched bits at every block size the server accepts.
    refs = {k: ctx.run_bscan(v[1], v[2], v[3], v[4])[0] for k, v in small_meshes.items()}
    srv = O.Server(ctx, slice_photons=4096, block_size=block)
    try:
        ids = {k: srv.submit(v[1], v[2], v[3], v[4]) for k, v in small_meshes.items()}
        for k, rid in ids.items():
            info, out = srv.wait(rid, 120)
            assert info.status == 0 and np.array_equal(out, refs[k]), k
    finally:
        srv.close()


def test_served_per_lane_record_path_equals_batched(ctx, small_meshes):
    # record_fetch 1 keeps every slice on the per-lane gathers, even where the cooperative
    # fetch fits; small slices make photons cross slice boundaries (save and restore).
    refs = {k: ctx.run_bscan(v[1], v[2], v[3], v[4])[0] for k, v in small_meshes.items()}
    srv = O.Server(ctx, slice_photons=4096, record_fetch=1)
    try:
        ids = {k: srv.submit(v[1], v[2], v[3], v[4]) for k, v in small_meshes.items()}
        for k, rid in ids.items():
            info, out = srv.wait(rid, 120)
            assert info.status == 0 and np.array_equal(out, refs[k]), k
        assert srv.stats()["slices"] > 2
    finally:
        srv.close()


def test_hand_out_policies_give_the_same_bits_and_order_service(ctx, small_meshes):
    # A large request (4 A-scans x 4e6 photons, ~3 s) and a small one (1 A-scan x 1e5)
    # submitted together. Least-remaining-first hands out the small request's photons first,
    # so it ends after about its slowest photon's lifetime and well before the large one;
    # equal shares give it a fifth of each slice's photons, so it ends later. Both policies
    # give the batched bits for both requests.
    _, mid, p, s, seed = small_meshes["slab"]
    big = [O.ascan(x.x, x.y, x.z, x.root_tet, 4_000_000, x.stream_salt) for x in s]
    small = [O.ascan(s[0].x, s[0].y, s[0].z, s[0].root_tet, 100_000, 9)]
    ref_big, _ = ctx.run_bscan(mid, p, big, seed)
    ref_small, _ = ctx.run_bscan(mid, p, small, seed)
    lat = {}
    for policy in (0, 1):
        srv = O.Server(ctx, policy=policy)
        try:
            rb = srv.submit(mid, p, big, seed)
            rs = srv.submit(mid, p, small, seed)
            i_s, out_s = srv.wait(rs, 120)
            i_b, out_b = srv.wait(rb, 120)
        finally:
            srv.close()
        assert i_s.status == 0 and i_b.status == 0
        assert np.array_equal(out_s, ref_small) and np.array_equal(out_b, ref_big)
        lat[policy] = (i_s.t_done_ms - i_s.t_submit_ms, i_b.t_done_ms - i_b.t_submit_ms)
    print("latency small/large (ms): least remaining first", lat[0], "equal shares", lat[1])
    assert lat[0][0] < 0.5 * lat[0][1]
    assert lat[0][0] < lat[1][0]


@pytest.mark.parametrize("field,value", [("policy", 2), ("policy", -1), ("record_fetch", 3), ("record_fetch", -1), ("block_size", 320)])
def test_invalid_serve_options_are_rejected(ctx, field, value):
    with pytest.raises(O.OctmpsError) as e:
        O.Server(ctx, **{field: value})
    assert e.value.status == 1


def test_serving_failure_is_retried_deterministically(ctx, small_meshes):
    _, mid, p, s, seed = small_meshes["sphere"]
    ref = ctx.run_bscan(mid, p, s, seed)[0]
    srv = O.Server(ctx, fail_request_index=0, slice_photons=2048)
    try:
        info, out = srv.wait(srv.submit(mid, p, s, seed), 120)
        assert info.status == 0 and np.array_equal(out, ref)
    finally:
        srv.close()


def test_serving_cancel_and_admission(ctx, small_meshes):
    _, mid, p, s, seed = small_meshes["slab"]
    big = [O.ascan(x.x, x.y, x.z, x.root_tet, 4_000_000, x.stream_salt) for x in s]
    words = O.tally_words(p.num_depth_bins)
    # The budget fits the mesh and one request's tallies and state at a time: the second
    # request waits for admission.
    mesh_bytes = ctx.mesh_info(mid)[1]
    srv = O.Server(ctx, hbm_budget_bytes=mesh_bytes + len(s) * words * 8 + 8192, slice_photons=8192)
    try:
        with pytest.raises(O.OctmpsError) as e:
            srv.submit(mid, p, s + s, seed)  # twice the budget: rejected at submit
        assert e.value.status == 7
        r1 = srv.submit(mid, p, big, seed)
        r2 = srv.submit(mid, p, s, seed)
        time.sleep(0.05)
        i2 = srv.poll(r2)
        assert not i2.done
        srv.cancel(r1)
        info1, _ = srv.wait(r1, 120)
        assert info1.status == 10
        info2, out2 = srv.wait(r2, 120)
        assert info2.status == 0
        assert np.array_equal(out2, ctx.run_bscan(mid, p, s, seed)[0])
        assert info2.t_admit_ms >= info1.t_done_ms - 1e-6
    finally:
        srv.close()


def test_completion_copies_hold_resources_until_done(ctx, small_meshes):
    # A finished request's tallies are copied to the caller beside the next slice, and the
    # request keeps its scan slots, parameter set and tally range until that copy is done.
    # With slots for one request at a time, each of six requests is admitted only after the
    # previous one's copy has ended, and every one gives the batched bits.
    _, mid, p, s, seed = small_meshes["ellipsoid"]
    refs = [ctx.run_bscan(mid, p, s, seed + k)[0] for k in range(6)]
    srv = O.Server(ctx, max_inflight_scans=len(s), slice_photons=4096)
    try:
        ids = [srv.submit(mid, p, s, seed + k) for k in range(6)]
        infos = []
        for k, rid in enumerate(ids):
            info, out = srv.wait(rid, 120)
            assert info.status == 0 and np.array_equal(out, refs[k]), k
            infos.append(info)
        for a, b in zip(infos, infos[1:]):
            assert b.t_admit_ms >= a.t_done_ms - 1e-6
    finally:
        srv.close()
    # Requests arriving while others finish make the tally pool grow (it starts at 8 MiB;
    # each of these holds 16 x 4096 bins, 7 MiB) beside the copies: the same bits.
    mb, pb, sb, seedb = configs.slab4(cells=(16, 16, 8), n_scans=16, photons=3000, bins=4096)
    midb = ctx.upload(mb)
    refs = [ctx.run_bscan(midb, pb, sb, seedb + k)[0] for k in range(4)]
    srv = O.Server(ctx, slice_photons=8192)
    try:
        ids = []
        for k in range(4):
            ids.append(srv.submit(midb, pb, sb, seedb + k))
            time.sleep(0.02)
        for k, rid in enumerate(ids):
            info, out = srv.wait(rid, 120)
            assert info.status == 0 and np.array_equal(out, refs[k]), k
    finally:
        srv.close()
        ctx.release(midb)


def test_close_finishes_queued_running_and_copying_requests(ctx, small_meshes):
    # octmps_serve_close runs every submitted request to its end before it returns: requests
    # still queued (slots for one request at a time), running, or having their tallies copied
    # by the completion thread. Their output buffers then hold the batched bits.
    _, mid, p, s, seed = small_meshes["ellipsoid"]
    refs = [ctx.run_bscan(mid, p, s, seed + 20 + k)[0] for k in range(5)]
    srv = O.Server(ctx, max_inflight_scans=2 * len(s), slice_photons=4096)
    ids = [srv.submit(mid, p, s, seed + 20 + k) for k in range(5)]
    outs = [srv._keep[rid] for rid in ids]  # the buffers the server writes (binding-owned)
    srv.close()
    for k, (out, n) in enumerate(outs):
        assert np.array_equal(out.reshape(n, -1), refs[k]), k


def test_random_submit_cancel_release_interleaving(ctx, small_meshes):
    # Requests of random meshes, seeds and A-scan subsets arrive while others finish, are
    # copied back and are released; some are cancelled on the way. Every request ends: a
    # cancelled one as cancelled (or done, if it finished first, then with the batched bits),
    # every other with the batched bits; released ids are unknown afterwards.
    rng = np.random.default_rng(5)
    names = sorted(small_meshes)
    srv = O.Server(ctx, slice_photons=2048, max_inflight_scans=24)
    plan, results = [], {}

    def take(rid):
        info, out = srv.wait(rid, 120)
        assert info.done
        results[rid] = (info.status, out.copy())
        srv.release(rid)
        with pytest.raises(O.OctmpsError):
            srv.poll(rid)

    try:
        live = []
        for i in range(24):
            name = names[int(rng.integers(len(names)))]
            _, mid, p, s, seed = small_meshes[name]
            pick = sorted(rng.choice(len(s), size=int(rng.integers(1, len(s) + 1)), replace=False))
            scans = [s[j] for j in pick]
            rid = srv.submit(mid, p, scans, seed + 100 + i)
            cancel = bool(rng.random() < 0.25)
            if cancel:
                srv.cancel(rid)
            plan.append((rid, mid, p, scans, seed + 100 + i, cancel))
            live.append(rid)
            if rng.random() < 0.3 and len(live) > 3:  # an earlier one ends and is released
                take(live.pop(0))
        for rid in live:
            take(rid)
    finally:
        srv.close()
    for rid, mid, p, scans, seed, cancel in plan:
        status, out = results[rid]
        if cancel and status == 10:
            continue
        assert status == 0, (rid, status)
        assert np.array_equal(out, ctx.run_bscan(mid, p, scans, seed)[0]), rid


def test_translation_invariance_in_layers(ctx):
    # In a laterally uniform layered medium, A-scans away from the sides are statistically
    # the same (independent streams per A-scan).
    m, p, scans, seed = configs.slab4(cells=(32, 32, 16), n_scans=2, photons=60000, bins=128, pitch=0.08)
    mid = ctx.upload(m)
    t, st = ctx.run_bscan(mid, p, scans, seed)
    assert st.killed == 0
    a, b = (O.decode_tally(x, p.num_depth_bins) for x in t)
    n = scans[0].n_photons
    ma, sa = bin_stats(a["sum"], a["sumsq"], n)
    mb, sb = bin_stats(b["sum"], b["sumsq"], n)
    sig = np.sqrt(sa ** 2 + sb ** 2)
    pop = sig > 0
    z = (ma[pop] - mb[pop]) / sig[pop]
    assert np.mean(np.abs(z) > 3) <= 0.02 and np.abs(z).max() < 6
    ra, rb = ma.sum(), mb.sum()
    sr = np.sqrt((sa ** 2).sum() + (sb ** 2).sum())
    assert abs(ra - rb) < 5 * sr


def test_zero_budget_a_scan_in_a_batch(ctx, small_meshes):
    # An A-scan with no photons gets an all-zero tally and changes nothing for the others
    # (every photon's streams depend only on its own A-scan's seed, salt and index).
    m, mid, p, scans, seed = small_meshes["slab"]
    ref, _ = ctx.run_bscan(mid, p, scans, seed)
    s2 = scans[2]
    mixed = list(scans[:2]) + [O.ascan(s2.x, s2.y, s2.z, s2.root_tet, 0, s2.stream_salt)] + list(scans[3:])
    got, st = ctx.run_bscan(mid, p, mixed, seed)
    assert st.photons == sum(s.n_photons for s in mixed)
    assert not got[2].any()
    for i in (0, 1, 3):
        assert np.array_equal(got[i], ref[i]), i


def test_step_guard_kills_the_photons_the_oracle_kills(ctx, table):
    # max_steps_per_photon ends a photon after that many steps, as the oracle's max_steps
    # does; with paired streams the same photons are cut short (a few may differ where an
    # FP32 and an FP64 step decide a near-edge crossing differently).
    import oracle_py as OR
    n, cap = 3000, 3000
    mesh, params, scans, seed = configs.sphere(cells=12, photons=n, bins=96)
    mid = ctx.upload(mesh)
    o = O.BatchOpts()
    o.max_steps_per_photon = cap
    t, st, plog = ctx.run_bscan(mid, params, scans, seed, opts=o, log=True)
    ref = OR.run_ascan(mesh.soa(), params, scans[0], seed, table, threads=8, max_steps=cap, log=True)
    assert st.photons == n and ref["killed"] > n // 2
    gpu_killed = plog["killed"].astype(bool)
    # The oracle's log has no kill flag; a photon that reached the cap was killed, except
    # one whose last step (a roulette death) was exactly the cap-th.
    ref_killed = ref["log"][:, 0] >= cap
    assert st.killed == int(gpu_killed.sum()) and 0 <= int(ref_killed.sum()) - ref["killed"] <= 5
    assert np.mean(gpu_killed == ref_killed) > 0.95
    assert abs(st.killed - ref["killed"]) <= 0.03 * ref["killed"]


def test_gather_probe_measures_a_sane_ceiling(ctx):
    # 256 CUs issue at most a few 16-byte lane-loads per cycle between them at ~2.4 GHz.
    r = ctx.probe_gather(1 << 12, 200)
    assert 1e10 < r < 1e13
    with pytest.raises(O.OctmpsError):
        ctx.probe_gather(0, 10)


def test_l1_probe_shows_the_lookup_cost_of_each_pattern(ctx):
    # The cost model behind the cooperative record fetch: the L1 looks up one line per cycle
    # and serves four lanes of a 16-byte load per cycle, so a divergent wave load takes ~64
    # CU cycles, a quad-coalesced one ~16 (also as DMA into LDS), and masked-off or
    # out-of-range lanes cost nothing.
    cyc = [ctx.probe_l1(p)[1] for p in range(5)]
    assert 40 < cyc[0] < 90
    assert cyc[1] < cyc[0] / 3 and cyc[4] < cyc[0] / 3
    assert cyc[2] < 0.7 * cyc[0] and cyc[3] < 0.7 * cyc[0]
    with pytest.raises(O.OctmpsError):
        ctx.probe_l1(5)


def test_valu_probe_rates_by_instruction_class(ctx):
    # Issue cost of the VALU classes the transport loop uses, at 6 waves per SIMD: a wave64
    # FP32 FMA takes a few SIMD cycles, a transcendental at least as many, an MWC step (two
    # instructions) more than one FMA.
    rates = [ctx.probe_valu(k) for k in range(3)]
    for r, cyc in rates:
        assert r > 0 and 0.5 < cyc < 64
    assert rates[1][1] >= 0.9 * rates[0][1]
    assert rates[2][1] > rates[0][1]
    with pytest.raises(O.OctmpsError):
        ctx.probe_valu(3)


def test_autotune_picks_a_valid_configuration(ctx, small_meshes):
    _, mid, p, s, seed = small_meshes["slab"]
    o, rate = ctx.autotune(mid, p, s, seed, calib=4000)
    assert o.block_size in (128, 192, 256) and o.refill_quantum in (1, 2, 4, 8, 16) and rate > 0
    assert 16 <= o.scatter_quantum <= 44 and o.tally_copies in (1, 2, 4)
    assert o.entry_skip in (0, 1, 2) and o.record_fetch in (0, 1)
    t1 = ctx.run_bscan(mid, p, s, seed)[0]
    t2 = ctx.run_bscan(mid, p, s, seed, opts=o)[0]
    assert np.array_equal(t1, t2)  # launch shape never changes results


def test_bump_projected_probes_run(ctx):
    m, p, scans, seed = configs.bump(cells=(32, 32, 16), n_scans=3, photons=2000, bins=256)
    mid = ctx.upload(m)
    assert scans[1].z < 0  # on the bump, above z = 0
    t, st = ctx.run_bscan(mid, p, scans, seed)
    assert st.killed == 0 and st.photons == 6000
    assert all(O.decode_tally(x, 256)["num_filtered"] > 0 for x in t)


def _write_case(d, mesh):
    (d / "case.opt").write_text(
        "1.0  # version\n1  # runs\nout.mco A\n3000  # photons\n2  # regions\n1.0  # ambient n\n"
        "1.40 1.0 100 0.9\n1.33 0.4 60 0.9\n-0.002 0.002 0.003  # start step end\n")
    (d / "case.mci").write_text("1\n37\n0.925\n0.0\n0.06\n0.0015\n0\n10e-4\n5\n0.5\n0.01\n")
    mesh.write(str(d / "case.mesh"))


def test_cli_writes_reference_outputs(tmp_path, ctx):
    mesh = O.Mesh.generate(O.GEOM_LAYERS, (16, 16, 8), (0.2, 0.2, 0.1), p=(0.02,), n_layers=2)
    _write_case(tmp_path, mesh)
    r = subprocess.run([O.CLI_PATH, "-S985456376", "-G1", "case.opt", "case.mesh", "case.mci"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    p, _ = O.read_opt(str(tmp_path / "case.opt"))
    O.read_mci(str(tmp_path / "case.mci"), p)
    B = p.num_depth_bins
    for f in ("ClassI_Scatt.out", "ClassII_Scatt.out", "ClassI_ScattFilt.out", "ClassII_ScattFilt.out"):
        rows = (tmp_path / f).read_text().splitlines()
        assert len(rows) == 3 * B  # x = -0.002, 0, 0.002 (x < end, src/main.cu:948-950)
    # the CLI's first A-scan equals the library's plain run with salt 0
    m2 = O.Mesh.read(str(tmp_path / "case.mesh"))
    mid = ctx.upload(m2)
    root, _ = m2.find_root(-0.002, 0.0, 0.0)
    t, _ = ctx.run_ascan(mid, p, O.ascan(-0.002, 0.0, 0.0, root, 3000), 985456376)
    out = O.finalize(t, B, 3000, p.max_relative_contribution)
    rows = [r.split(",") for r in (tmp_path / "ClassI_Scatt.out").read_text().splitlines()[:B]]
    got = np.array([float(r[2]) for r in rows])
    assert np.allclose(got, out[0, 0, 0], rtol=1e-5, atol=0)


def test_cli_runs_every_run_of_a_multi_run_opt(tmp_path, ctx):
    # A 2-run .opt: at each probe position of run 0's line, run 0 then run 1 are simulated with
    # their own regions and photon counts and appended in that order (src/main.cu:948-1036).
    from test_cpu_formats import TWO_RUNS
    mesh = O.Mesh.generate(O.GEOM_LAYERS, (16, 16, 8), (0.2, 0.2, 0.1), p=(0.02,), n_layers=2)
    _write_case(tmp_path, mesh)
    (tmp_path / "case.opt").write_text(TWO_RUNS)
    r = subprocess.run([O.CLI_PATH, "-S985456376", "case.opt", "case.mesh", "case.mci"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    p0, runs = O.read_opt(str(tmp_path / "case.opt"), 0)
    assert runs == 2
    O.read_mci(str(tmp_path / "case.mci"), p0)
    p1 = type(p0).from_buffer_copy(p0)
    O.read_opt(str(tmp_path / "case.opt"), 1, p1)
    B = p0.num_depth_bins
    rows = [x.split(",") for x in (tmp_path / "ClassI_Scatt.out").read_text().splitlines()]
    xs = [-0.002, 0.0, 0.002]  # run 0's line; run 1's own line is not the B-scan's
    assert len(rows) == len(xs) * 2 * B
    m2 = O.Mesh.read(str(tmp_path / "case.mesh"))
    mid = ctx.upload(m2)
    for i, x in enumerate(xs):
        for k, (pk, nk) in enumerate(((p0, 3000), (p1, 1500))):
            block = rows[(2 * i + k) * B:(2 * i + k + 1) * B]
            assert all(abs(float(row[0]) - x) < 1e-12 for row in block)
            if i == 0:
                root, _ = m2.find_root(x, 0.0, 0.0)
                t, st = ctx.run_ascan(mid, pk, O.ascan(x, 0.0, 0.0, root, nk), 985456376)
                assert st.photons == nk
                want = O.finalize(t, B, nk, pk.max_relative_contribution)
                got = np.array([float(row[2]) for row in block])
                assert np.allclose(got, want[0, 0, 0], rtol=1e-5, atol=0), k
    assert "1500" in (tmp_path / "BiasParam.mco").read_text()  # rewritten by the last run


def test_integration_python_example_runs(tmp_path, ctx, monkeypatch):
    # The Python example of the integration guide, run as written on a small two-layer case: its
    # A-scan equals the library's own plain run of the same inputs.
    import re

    mesh = O.Mesh.generate(O.GEOM_LAYERS, (16, 16, 8), (0.2, 0.2, 0.1), p=(0.02,), n_layers=2)
    _write_case(tmp_path, mesh)
    root_dir = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    text = open(os.path.join(root_dir, "INTEGRATION.md")).read()
    block = re.search(r"```python\n(.*?)```", text, re.S).group(1)
    monkeypatch.chdir(tmp_path)
    ns = {}
    exec(compile(block, "INTEGRATION.md", "exec"), ns)
    p = ns["p"]
    assert ns["stats"].photons == p.number_of_photons == 3000
    out = ns["mean_se"]  # [filtered][class][mean, SE][bin]
    assert out.shape == (2, 2, 2, p.num_depth_bins) and np.all(np.isfinite(out[:, :, 0]))
    # SE is sqrt of (SumSq/N - mean^2)/(N-1) unclamped, as the reference prints it
    # (src/octmps_io.cpp:385-388, 408-411). The tally holds every contribution and its square
    # exactly, so no bin's variance is negative: no SE is NaN, raw or filtered.
    assert np.all(np.isfinite(out))
    ns["ctx"].close()
    mid = ctx.upload(ns["mesh"])
    t, _ = ctx.run_ascan(mid, p, O.ascan(0.0, 0.0, ns["z"], ns["root"], 3000), 985456376)
    assert np.array_equal(t, ns["tally"])
