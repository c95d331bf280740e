"""The request server on the served-stream workload and its HBM-footprint admission.

The served stream (Poisson arrivals, seed 11) replaces the reference's host B-scan loop
(src/main.cu:948-1051); the footprint admission replaces its per-A-scan allocation of mesh,
RNG and state (src/main.cu:100-465)."""
import json
import os
import time

import numpy as np
import pytest

import octmps as O
from octmps import configs

pytestmark = pytest.mark.gpu

TRACES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "workload_traces.json")


def _state_bound(n_scans):
    # upper bound of a request's footprint besides tallies and mesh (slots and parameter set)
    return n_scans * 72 + 80 + 3200 + 4096


def test_served_stream_c3_replay_equals_batched(ctx):
    # The 400 req/s seed-11 trace (tests/golden/workload_traces.json), first 8 requests, with
    # their arrival times, over the four meshes of the served stream registered with the
    # context: the 32^3-cell sphere (196,608 tets), the four-layer slab and the ellipsoid
    # (786,432 tets each) and the 128^3-cell layered slab (12,582,912 tets, 1.6 GB of records).
    # Budgets are capped at 2e4 photons per A-scan. The server's budget holds the four meshes
    # and about one request's tallies, so later requests wait for admission. Every request
    # must equal octmps_run_bscan of the same inputs bit for bit.
    trace = json.load(open(TRACES))["served_seed11"]
    stream = next(s for s in trace if s["rate"] == 400.0)["requests"][:8]
    built = [configs.sphere(photons=1), configs.slab4(n_scans=1, photons=1),
             configs.ellipsoid(n_scans=1, budgets=[1])]
    meshes = [(m, p, s[0]) for m, p, s, _ in built]
    m128, p128 = configs.slab128()
    assert m128.num_tets == 12582912
    root, z = m128.find_root(0.0, 0.0, 0.0)
    meshes.append((m128, p128, O.ascan(0.0, 0.0, z, root, 1)))
    ids = [ctx.register(m) for m, _, _ in meshes]
    mesh_bytes = sum(ctx.mesh_info(i)[1] for i in ids)
    reqs = []
    for k, r in enumerate(stream):
        _, p, probe = meshes[r["mesh"]]
        scans = [O.ascan(probe.x, probe.y, probe.z, probe.root_tet, min(int(b), 20000), salt=1000 * k + j)
                 for j, b in enumerate(r["budgets"])]
        reqs.append((r["t"], r["mesh"], p, scans))
    biggest = max(O.tally_words(p.num_depth_bins) * 8 * len(sc) + _state_bound(len(sc)) for _, _, p, sc in reqs)
    srv = O.Server(ctx, hbm_budget_bytes=mesh_bytes + biggest)
    try:
        rids = []
        start = time.perf_counter()
        for t, mi, p, scans in reqs:
            delay = start + t - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            rids.append(srv.submit(ids[mi], p, scans, 11))
        infos, outs = zip(*[srv.wait(rid, 300) for rid in rids])
        res = srv.residency()
    finally:
        srv.close()
    assert all(i.status == 0 and i.done for i in infos)
    # later requests waited for earlier ones to leave: admission after another's completion
    waited = [k for k in range(1, len(infos)) if infos[k].t_admit_ms >= min(i.t_done_ms for i in infos[:k]) - 1e-6]
    print("admitted after an earlier request ended:", waited, "residency", res)
    assert waited
    assert res["uploads"] >= len({mi for _, mi, _, _ in reqs}) and res["used"] <= res["budget"]
    for k, (t, mi, p, scans) in enumerate(reqs):
        ref, st = ctx.run_bscan(ids[mi], p, scans, 11)
        assert st.bad_contributions == 0
        assert np.array_equal(outs[k], ref), k


@pytest.fixture(scope="module")
def two_meshes():
    a = configs.slab4(cells=(24, 24, 12), n_scans=3, photons=60000, bins=128)
    b = configs.sphere(cells=20, photons=4000, bins=96)
    return a, b


def test_admission_by_mesh_footprint_evicts_idle_meshes(ctx, two_meshes):
    # Two registered meshes whose requests fit the budget one at a time but not together: the
    # second request waits until the first is done, then its mesh is uploaded after the
    # first's is evicted. A request whose mesh alone exceeds the budget is refused at submit.
    (ma, pa, sa, seed_a), (mb, pb, sb, seed_b) = two_meshes
    # references from pinned uploads of the same meshes (results do not depend on placement)
    ref_a = ctx.run_bscan(ctx.upload(ma), pa, sa, seed_a)[0]
    ref_b = ctx.run_bscan(ctx.upload(mb), pb, sb, seed_b)[0]
    ida, idb = ctx.register(ma), ctx.register(mb)
    assert not ctx.mesh_info(ida)[0] and not ctx.mesh_info(idb)[0]
    bytes_a, bytes_b = ctx.mesh_info(ida)[1], ctx.mesh_info(idb)[1]
    fa = bytes_a + O.tally_words(pa.num_depth_bins) * 8 * len(sa) + _state_bound(len(sa))
    fb = bytes_b + O.tally_words(pb.num_depth_bins) * 8 * len(sb) + _state_bound(len(sb))
    budget = max(fa, fb) + min(bytes_a, bytes_b) // 2
    assert budget < fa + fb - 8192
    srv = O.Server(ctx, hbm_budget_bytes=budget)
    try:
        m_big, p_big, s_big, _ = configs.slab4(cells=(48, 48, 24), n_scans=1, photons=10, bins=128)
        id_big = ctx.register(m_big)
        assert ctx.mesh_info(id_big)[1] > budget
        with pytest.raises(O.OctmpsError) as e:
            srv.submit(id_big, p_big, s_big, 1)
        assert e.value.status == 7
        r1 = srv.submit(ida, pa, sa, seed_a)
        r2 = srv.submit(idb, pb, sb, seed_b)
        i1, o1 = srv.wait(r1, 300)
        i2, o2 = srv.wait(r2, 300)
        res = srv.residency()
    finally:
        srv.close()
    assert i1.status == 0 and i2.status == 0
    assert np.array_equal(o1, ref_a) and np.array_equal(o2, ref_b)
    assert i2.t_admit_ms >= i1.t_done_ms - 1e-6
    print("residency", res)
    assert res["uploads"] == 2 and res["evictions"] == 1
    assert not ctx.mesh_info(ida)[0] and ctx.mesh_info(idb)[0]
    # an evicted registered mesh is uploaded again by a plain run, with the same bits
    assert np.array_equal(ctx.run_bscan(ida, pa, sa, seed_a)[0], ref_a)
    assert ctx.mesh_info(ida)[0]


def test_release_is_refused_while_a_request_uses_the_mesh(ctx, two_meshes):
    # A mesh with a submitted request (queued behind a long one) cannot be released until the
    # request is done; releasing it under the request would leave the request without a mesh.
    (ma, pa, sa, seed_a), (mb, pb, sb, seed_b) = two_meshes
    ida, idb = ctx.upload(ma), ctx.upload(mb)
    long_scans = [O.ascan(s.x, s.y, s.z, s.root_tet, 2_000_000, s.stream_salt) for s in sa]
    srv = O.Server(ctx, max_inflight_scans=len(long_scans))
    try:
        r_long = srv.submit(ida, pa, long_scans, seed_a)
        r_b = srv.submit(idb, pb, sb, seed_b)
        assert not srv.poll(r_b).done
        with pytest.raises(O.OctmpsError) as e:
            ctx.release(idb)
        assert e.value.status == 9  # OCTMPS_E_BUSY
        info, out = srv.wait(r_b, 300)
        srv.wait(r_long, 300)
        assert info.status == 0
    finally:
        srv.close()
    ctx.release(idb)
    with pytest.raises(O.OctmpsError) as e:
        ctx.mesh_info(idb)
    assert e.value.status == 11


def test_finished_requests_can_be_released(ctx, two_meshes):
    (ma, pa, sa, seed_a), _ = two_meshes
    mid = ctx.upload(ma)
    small = [O.ascan(s.x, s.y, s.z, s.root_tet, 2000, s.stream_salt) for s in sa]
    srv = O.Server(ctx)
    try:
        rid = srv.submit(mid, pa, small, seed_a)
        srv.wait(rid, 120)
        srv.release(rid)
        with pytest.raises(O.OctmpsError) as e:
            srv.poll(rid)
        assert e.value.status == 11
    finally:
        srv.close()
