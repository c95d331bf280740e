"""Benchmark: B-scan photons/s of the batched persistent kernel vs the plain one-A-scan-per-
launch path, on the four-layer slab B-scan (786,432 tets, 128 A-scans x 1e6 photons, 512
bins to 1.5 mm), with the roofline of the transport kernel and the CPU oracle's rate on the
host cores beside it.

A step is one full B-scan (all A-scans of the workload) through the batched path, launched
as one hipGraph. The whole run keeps to a wall-clock budget (--budget-s): the fixed-cost
legs run first (autotune on a 1/64 slice of at least 4e6 photons, warmup steps on a 1/64 slice, the plain path on
every 8th A-scan, the CPU oracle on a bounded sample, the served stream and its load sweep,
the cache and HBM probes), then as many of the requested timed steps as fit in what is left
(at least one). `steps` reports the number timed; `steps_requested` the number asked for.
Prints ONE JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

T_START = time.time()
ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "oct-mps_amd"))

ALG_BYTES_PER_STEP = 84  # one tet record per photon-step: 4 planes x 16 B + 4 links x 4 B + region 4 B
HBM_PEAK_GBS = 8000.0    # MI355X HBM3E nominal
LAYOUT_BYTES_PER_STEP = 80  # this build's packed tet record: 4 planes x 16 B + 4 packed neighbours x 4 B

# Counter-derived figures of the batched slab4 launch (128 A-scans x 1e6 photons) for each record
# path, each from the rocprofv3 summaries named beside it; per photon-step = counter / the
# launch's photon-steps. HBM bytes = FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 (the factors by which
# those counters read a copy kernel of known size). Wave-cycle shares are SQ_ACTIVE_INST_ANY
# (issuing), SQ_WAIT_INST_ANY (ready, waiting for an issue slot) and SQ_WAIT_ANY (waiting on a
# dependency) over SQ_WAVE_CYCLES.
PROFILE = {
    # Cooperative record fetch, quanta 28/1.
    "coop": {"l1_lookups_per_step": 1.72, "ta_busy": 0.65, "tcp_pending_l2": 0.38, "l2_hit": 0.717,
             "l2_latency_cycles": 324, "wave_wait_dependency": 0.27, "wave_wait_issue": 0.38, "wave_issuing": 0.35,
             "valu_per_step": 4.98, "salu_per_step": 1.74, "salu_per_cu_cycle": 0.472, "hbm_bytes_per_step": 24.99,
             "source": "profiles/transport_slab4_memory.md, transport_slab4_sq.md, transport_slab4_pmc.md"},
    # Per-lane gathers, quanta 24/20, entry-plane skip on / off.
    "lane_skip_on": {"l1_lookups_per_step": 4.16, "ta_busy": 0.84, "tcp_pending_l2": 0.18, "valu_per_step": 5.59,
                     "salu_per_step": 2.56, "salu_per_cu_cycle": 0.475, "hbm_bytes_per_step": 25.53,
                     "source": "profiles/transport_slab4_perlane_skip_memory.md, transport_slab4_perlane_skip_sq.md, "
                               "transport_slab4_perlane_skip_pmc.md"},
    "lane_skip_off": {"l1_lookups_per_step": 5.01, "ta_busy": 0.87, "tcp_pending_l2": 0.01, "valu_per_step": 5.32,
                      "salu_per_step": 2.53, "salu_per_cu_cycle": 0.485, "hbm_bytes_per_step": 25.56,
                      "source": "profiles/transport_slab4_perlane_memory.md, transport_slab4_perlane_sq.md, "
                                "transport_slab4_perlane_pmc.md"},
    # The bump sample (64 of 1024 A-scans x 1e6 photons, cooperative fetch, quanta 28/1, 4 blocks
    # per CU).
    "bump_hbm_bytes_per_step": 43.99, "bump_hbm_source": "profiles/transport_bump_pmc.md",
}


def log(msg):
    print(f"[bench {time.time() - T_START:6.1f}s] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--budget-s", type=float, default=530.0,
                    help="wall-clock budget of the whole run; timed steps stop when the next would pass it")
    ap.add_argument("--scans", type=int, default=128, help="A-scans in the B-scan (workload: 128)")
    ap.add_argument("--photons", type=int, default=1_000_000, help="photons per A-scan (workload: 1e6)")
    ap.add_argument("--slice-div", type=int, default=64,
                    help="warmup and autotune calibration use 1/k of every A-scan's photons (autotune: at least 4e6 "
                         "photons in all)")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-photons", type=int, default=3000, help="oracle photons per sampled A-scan")
    ap.add_argument("--cpu-scans", type=int, default=8)
    ap.add_argument("--no-autotune", action="store_true",
                    help="skip the autotuner (context defaults: scatter/refill quanta 28/1, block 256)")
    ap.add_argument("--quanta", default="",
                    help="scatter,refill quanta to launch with instead of autotuning (e.g. 36,16: the same "
                         "launch a profiler run can then attribute to one kernel dispatch)")
    ap.add_argument("--entry-skip", type=int, default=-1, choices=(-1, 0, 1, 2),
                    help="with --quanta: entry-plane skip of the per-lane record path (0 auto, 1 on, 2 off)")
    ap.add_argument("--record-fetch", type=int, default=-1, choices=(-1, 0, 1, 2),
                    help="with --quanta: record path (0 auto, 1 per-lane gathers, 2 cooperative)")
    ap.add_argument("--blocks-per-cu", type=int, default=0,
                    help="with --quanta: resident blocks per CU (0: the occupancy maximum)")
    ap.add_argument("--bump-quanta", default="28,1,4",
                    help="scatter quantum, refill quantum and blocks per CU of the bump sample (the autotuner's "
                         "choice for that mesh)")
    ap.add_argument("--serve-rate", type=float, default=50.0, help="Poisson arrival rate of the C3 stream, req/s")
    ap.add_argument("--serve-window", type=float, default=0.1, help="seconds of arrivals in the C3 stream")
    ap.add_argument("--sweep-rates", default="10,25,50,100,200",
                    help="arrival rates (req/s) of the load sweep on the budget-scaled request mix")
    ap.add_argument("--sweep-window", type=float, default=1.0, help="seconds of arrivals per sweep point")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-policy", type=int, default=0, choices=(0, 1),
                    help="request server hand-out order: 0 least remaining photons first, 1 equal shares")
    ap.add_argument("--serve-same", action="store_true",
                    help="submit the B-scan as one request to the server and compare its tallies with the "
                         "batched launch's (on by default, off with --no-serve unless this flag is given)")
    ap.add_argument("--no-serve-same", action="store_true")
    ap.add_argument("--slab128-cells", type=int, default=128)
    ap.add_argument("--calibrate-copy-mb", type=int, default=0,
                    help="run a plain copy kernel of this size first (byte-counter calibration)")
    ap.add_argument("--plain-every", type=int, default=8,
                    help="time the plain path on every k-th A-scan of the B-scan")
    ap.add_argument("--workload", choices=("slab4", "ellipsoid", "bump"), default="slab4",
                    help="slab4: the headline B-scan; ellipsoid: 512 positions with log-uniform budgets "
                         "in [1e5, 1e7] (seed 7); bump: 6.3M-tet bump mesh, 1024 A-scans x 1e6 photons")
    ap.add_argument("--stride", type=int, default=1,
                    help="keep every k-th A-scan of the workload (a bounded sample of a long B-scan)")
    ap.add_argument("--no-probes", action="store_true", help="skip the gather and HBM copy probes")
    ap.add_argument("--bump-stride", type=int, default=16,
                    help="slab4 runs also time the bump mesh on every k-th of its 1024 A-scans (0: skip)")
    return ap.parse_args()


def resident_meshes(O, configs, ctx, a):
    """The four meshes of the served stream (C3), uploaded once: sphere, slab4, ellipsoid and
    the 128^3-cell layered slab."""
    meshes = []
    m, p, s, _ = configs.sphere(photons=1)
    meshes.append((m, ctx.upload(m), p, s[0]))
    m, p, s, _ = configs.slab4(n_scans=1, photons=1)
    meshes.append((m, ctx.upload(m), p, s[0]))
    m, p, s, _ = configs.ellipsoid(n_scans=1, budgets=[1])
    meshes.append((m, ctx.upload(m), p, s[0]))
    c = a.slab128_cells
    m, p = configs.slab128(cells=(c, c, c))
    root, z = m.find_root(0.0, 0.0, 0.0)
    meshes.append((m, ctx.upload(m), p, O.ascan(0.0, 0.0, z, root, 1)))
    return meshes


def serve_trace(O, srv, meshes, trace, salt0):
    """Submits a Poisson trace at its arrival times and waits for every request. Returns the
    per-request latencies (submit -> result, server clock, ms), the photons and the span."""
    import numpy as np
    ids, photons = [], 0
    start = time.perf_counter()
    for i, (t_arr, mi, n, budgets) in enumerate(trace):
        mesh, mid, p, probe = meshes[mi]
        scans = [O.ascan(probe.x, probe.y, probe.z, probe.root_tet, int(b), salt=salt0 + 1000 * i + k)
                 for k, b in enumerate(budgets)]
        photons += int(sum(budgets))
        delay = start + t_arr - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        ids.append(srv.submit(mid, p, scans, 1))
    lat, first, last = [], None, None
    for rid in ids:
        info, _ = srv.wait(rid, -1.0)
        assert info.status == 0, info.status
        srv.release(rid)
        lat.append(info.t_done_ms - info.t_submit_ms)
        first = info.t_submit_ms if first is None else min(first, info.t_submit_ms)
        last = info.t_done_ms if last is None else max(last, info.t_done_ms)
    return np.array(lat), photons, (last - first) * 1e-3


def run_serving(O, configs, ctx, a):
    """Served stream (C3): Poisson arrivals (seed 11) over four resident meshes; each request
    is 1-256 A-scans with log-uniform budgets in [1e4, 1e6] (~2.8e7 photons per request, so
    one GPU serves ~0.2 request/s of it and every C3 rate of 50-400 req/s is far past
    saturation: latency there measures the queue). Beside it, a load sweep below saturation:
    the same generator with budgets scaled by 1e-3 ([10, 1000] photons per A-scan)."""
    import numpy as np
    t0 = time.time()
    meshes = resident_meshes(O, configs, ctx, a)
    log(f"served meshes resident ({sum(x[0].num_tets for x in meshes)} tets) in {time.time() - t0:.1f}s")
    srv = O.Server(ctx, policy=a.serve_policy)
    try:
        trace = configs.served_stream(a.serve_rate, a.serve_window, seed=11, n_meshes=len(meshes))
        s0 = srv.stats()
        lat, photons, span = serve_trace(O, srv, meshes, trace, 0)
        s1 = srv.stats()
        c3 = {"requests": len(lat), "rate_req_s": a.serve_rate, "arrival_window_s": a.serve_window,
              "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
              "max_ms": float(lat.max()), "photons": photons, "photons_per_s": photons / span, "span_s": span,
              "kernel_busy_frac": (s1["kernel_ms"] - s0["kernel_ms"]) * 1e-3 / span,
              "note": f"p99 of {len(lat)} requests is their maximum; the mix averages ~2.8e7 photons per "
                      "request, so every C3 rate (50-400 req/s) saturates one GPU"}
        log(f"C3 stream: {len(lat)} requests, p50 {c3['p50_ms']:.0f} ms, p99 {c3['p99_ms']:.0f} ms, "
            f"{c3['photons_per_s']:.3e} photons/s")
        # Latency floor: one small request (1 A-scan x 1e4 photons, sphere) on the idle server.
        mesh, mid, p, probe = meshes[0]
        floor = []
        for k in range(5):
            small = [O.ascan(probe.x, probe.y, probe.z, probe.root_tet, 10000, salt=999000 + k)]
            rid = srv.submit(mid, p, small, 1)
            info, _ = srv.wait(rid, -1.0)
            srv.release(rid)
            floor.append(info.t_done_ms - info.t_submit_ms)
        sweep = []
        for k, rate in enumerate(float(r) for r in a.sweep_rates.split(",") if r):
            tr = configs.served_stream(rate, a.sweep_window, seed=11, n_meshes=len(meshes), lo=10.0, hi=1000.0)
            if not tr:
                continue
            s0 = srv.stats()
            lat, photons, span = serve_trace(O, srv, meshes, tr, 10_000_000 * (k + 1))
            s1 = srv.stats()
            offered = photons / a.sweep_window
            sweep.append({"rate_req_s": rate, "requests": len(lat), "p50_ms": float(np.percentile(lat, 50)),
                          "p99_ms": float(np.percentile(lat, 99)), "photons_per_s": photons / span,
                          "offered_photons_per_s": offered,
                          "kernel_busy_frac": (s1["kernel_ms"] - s0["kernel_ms"]) * 1e-3 / span})
            log(f"sweep {rate:g} req/s: {len(lat)} requests, p50 {sweep[-1]['p50_ms']:.0f} ms, "
                f"p99 {sweep[-1]['p99_ms']:.0f} ms, busy {sweep[-1]['kernel_busy_frac']:.2f}")
        st = srv.stats()
    finally:
        srv.close()
    for m in meshes:
        ctx.release(m[1])
    return {"c3": c3, "policy": ("least remaining photons first, aged", "equal shares")[a.serve_policy],
            "idle_small_request_ms": float(np.median(floor)), "slices": st["slices"],
            "load_sweep": {"mix": "C3 generator (seed 11, 1-256 A-scans, 4 resident meshes) with per-A-scan "
                                  "budgets log-uniform in [10, 1000] photons", "points": sweep},
            "meshes": "sphere 196,608 + slab4 786,432 + ellipsoid 786,432 + "
                      f"slab128 {meshes[3][0].num_tets} tets, all resident"}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); the photon path shards by "
              f"A-scan with no cross-GPU traffic and is measured at N=1.", file=sys.stderr)
        sys.exit(2)
    import numpy as np

    import octmps as O
    from octmps import configs

    deadline = T_START + a.budget_s
    t0 = time.time()
    if a.workload == "slab4":
        mesh, params, scans, seed = configs.slab4(n_scans=a.scans, photons=a.photons)
        desc = (f"slab4 B-scan: 4x4x2 mm four-layer slab, 64x64x32 cells x 6 tets ({mesh.num_tets} tets), "
                f"{{n}} A-scans x {a.photons} photons at 20 um, 512 bins to 1.5 mm, TypeBias 37")
    elif a.workload == "ellipsoid":
        mesh, params, scans, seed = configs.ellipsoid()
        desc = (f"ellipsoid + two spheres: 4x4x2 mm, 64x64x32 cells x 6 tets ({mesh.num_tets} tets), "
                f"{{n}} of 512 A-scan positions with log-uniform budgets in [1e5, 1e7] (seed 7), 512 bins")
    else:
        mesh, params, scans, seed = configs.bump()
        desc = (f"bump: 6x6x3 mm, 128x128x64 cells x 6 tets ({mesh.num_tets} tets), {{n}} of 1024 A-scans x "
                f"1e6 photons, probes projected onto the bump, 1024 bins to 2 mm")
    scans = scans[::max(1, a.stride)]
    total_photons = sum(s.n_photons for s in scans)
    log(f"mesh {mesh.num_tets} tets built in {time.time() - t0:.1f}s")
    ctx = O.Context(0)
    mid = ctx.upload(mesh)
    copy = None
    if a.calibrate_copy_mb:
        nbytes = a.calibrate_copy_mb << 20
        ms = ctx.probe_copy(nbytes, 10)
        copy = {"bytes_read": nbytes, "bytes_written": nbytes, "ms": ms, "GB_s": 2 * nbytes / (ms * 1e-3) / 1e9}
        log(f"copy calibration: {nbytes} B in {ms:.3f} ms ({copy['GB_s']:.0f} GB/s r+w)")
    opts = O.BatchOpts()
    opts.use_graph = 1
    tuned = None
    div = max(1, a.slice_div)
    if a.quanta:
        opts.scatter_quantum, opts.refill_quantum = (int(v) for v in a.quanta.split(","))
        if a.entry_skip >= 0:
            opts.entry_skip = a.entry_skip
        if a.record_fetch >= 0:
            opts.record_fetch = a.record_fetch
        opts.blocks_per_cu = a.blocks_per_cu
        tuned = {"scatter_quantum": opts.scatter_quantum, "refill_quantum": opts.refill_quantum,
                 "entry_skip": opts.entry_skip, "record_fetch": opts.record_fetch,
                 "blocks_per_cu": opts.blocks_per_cu, "source": "--quanta"}
    elif not a.no_autotune:
        # The autotuner (coordinate search over launch options on a calibration slice of the
        # same B-scan: the first 1/div of every A-scan's photons), before the timed region;
        # results are the same bits under any options.
        # At least 4e6 photons (~10 per lane of the grid): a smaller slice is mostly the launch's
        # tail, and options tuned on a tail lose on the full B-scan.
        calib = min(total_photons, max(total_photons // div, 4_000_000))
        ts = time.perf_counter()
        o, rate = ctx.autotune(mid, params, scans, seed, calib=calib)
        opts.block_size, opts.refill_quantum, opts.scatter_quantum = o.block_size, o.refill_quantum, o.scatter_quantum
        opts.tally_copies, opts.entry_skip, opts.record_fetch = o.tally_copies, o.entry_skip, o.record_fetch
        opts.blocks_per_cu = o.blocks_per_cu
        tuned = {"block_size": o.block_size, "refill_quantum": o.refill_quantum,
                 "scatter_quantum": o.scatter_quantum, "tally_copies": o.tally_copies, "entry_skip": o.entry_skip,
                 "record_fetch": o.record_fetch, "blocks_per_cu": o.blocks_per_cu, "calibration_photons": calib,
                 "seconds": time.perf_counter() - ts}
        log(f"autotune ({tuned['seconds']:.1f} s): block {o.block_size} blocks/CU {o.blocks_per_cu or 'max'} "
            f"scatter/refill quanta {o.scatter_quantum}/{o.refill_quantum} tally copies {o.tally_copies} "
            f"entry skip {o.entry_skip} record fetch {o.record_fetch} ({rate:.3e} photons/s on the calibration slice)")
    # Warmup: the logged kernel variant once (a different kernel symbol, so a profiler's
    # per-kernel average of the unlogged single-set transport_kernel covers the timed launches
    # only), then W steps of the unlogged graph launch on a 1/div slice of the B-scan.
    small = [O.ascan(s.x, s.y, s.z, s.root_tet, min(s.n_photons, 500), s.stream_salt) for s in scans]
    ctx.run_bscan(mid, params, small, seed, opts=opts, log=True)
    wslice = [O.ascan(s.x, s.y, s.z, s.root_tet, max(1, s.n_photons // div), s.stream_salt) for s in scans]
    for i in range(a.warmup):
        _, st = ctx.run_bscan(mid, params, wslice, seed, opts=opts)
        log(f"warmup step {i} (1/{div} slice): kernel {st.kernel_ms:.1f} ms")

    plain = None
    if not a.no_plain:
        # One A-scan per launch, synchronous, tallies copied back each time. Every A-scan of
        # the B-scan is the same amount of work, so the rate over every k-th A-scan is the
        # plain B-scan rate.
        sub = scans[::max(1, a.plain_every)]
        ts = time.perf_counter()
        pk, pst, nph = 0.0, 0, 0
        for s in sub:
            tp, stp = ctx.run_ascan(mid, params, s, seed, opts=opts)
            pk += stp.kernel_ms
            pst += stp.photon_steps
            nph += stp.photons
        pw = time.perf_counter() - ts
        plain = {"photons_per_s": nph / pw, "wall_s": pw, "kernel_ms": pk, "photon_steps": pst,
                 "a_scans_timed": len(sub), "photons": nph}
        log(f"plain: {len(sub)} A-scans in {pw:.3f}s ({nph / pw:.3e} photons/s)")

    cpu = None
    if not a.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py as OR
        K = 4096
        table = OR.Table(O.safeprimes(3 + 3 * K), K)
        soa = mesh.soa()
        threads = min(16, os.cpu_count() or 1)
        idx = np.linspace(0, len(scans) - 1, a.cpu_scans).round().astype(int)
        ts = time.perf_counter()
        n = 0
        subs, oracle_sum = [], 0.0
        for i in idx:
            s = scans[i]
            sub = O.ascan(s.x, s.y, s.z, s.root_tet, a.cpu_photons, s.stream_salt)
            subs.append(sub)
            r = OR.run_ascan(soa, params, sub, seed, table, threads=threads)
            n += r["photons"]
            oracle_sum += float(r["sum"].sum())
        cw = time.perf_counter() - ts
        # Single-thread rate on the first sampled A-scan (a quarter of the photons).
        s0 = scans[idx[0]]
        one = O.ascan(s0.x, s0.y, s0.z, s0.root_tet, max(1, a.cpu_photons // 4), s0.stream_salt)
        ts = time.perf_counter()
        r1 = OR.run_ascan(soa, params, one, seed, table, threads=1)
        single = r1["photons"] / (time.perf_counter() - ts)
        model = ""
        try:
            with open("/proc/cpuinfo") as f:
                model = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), "")
        except OSError:
            pass
        # The same photons (same streams) through the GPU kernel: error against the FP64 oracle.
        tg, _ = ctx.run_bscan(mid, params, subs, seed)
        gpu_sum = float(sum(O.decode_tally(x, params.num_depth_bins)["sum"].sum() for x in tg))
        cpu = {"value": n / cw, "unit": "photons/s", "cores": threads, "kind": "port",
               "sample": f"{a.cpu_scans} A-scans x {a.cpu_photons} photons of the same B-scan "
                         f"(FP64 oracle, {threads} threads, {cw:.1f} s)",
               "single_thread_value": single, "cpu_model": model, "nproc": os.cpu_count(),
               "gpu_vs_oracle": {"photons": n, "reflectance_sum_gpu": gpu_sum, "reflectance_sum_oracle": oracle_sum,
                                 "rel_err": abs(gpu_sum - oracle_sum) / oracle_sum if oracle_sum else None}}
        log(f"cpu oracle: {n} photons in {cw:.1f}s; same photons on the GPU: summed reflectance "
            f"{gpu_sum:.6e} vs oracle {oracle_sum:.6e}")

    serving = None
    if not a.no_serve:
        serving = run_serving(O, configs, ctx, a)

    served_same = None
    if a.serve_same or not (a.no_serve or a.no_serve_same):
        # The same B-scan as one request to the request server (the multi-set kernel in
        # wall-clock slices): its rate beside the batched launch and its tallies against the
        # batched launch's, bit for bit (the batched tallies come from the timed steps below).
        srv = O.Server(ctx, record_fetch=1 if a.record_fetch == 1 else 0)
        try:
            ts = time.perf_counter()
            rid = srv.submit(mid, params, scans, seed)
            info, served_out = srv.wait(rid, -1.0)
            sw = time.perf_counter() - ts
            sst = srv.stats()
        finally:
            srv.close()
        served_same = {"photons_per_s": total_photons / sw, "wall_s": sw, "slices": sst["slices"],
                       "status": info.status, "record_fetch": "per-lane" if a.record_fetch == 1 else "auto"}
        log(f"B-scan as one served request: {total_photons / sw:.3e} photons/s")

    # Ceiling of the vector-memory path for divergent record gathers (cache-resident record
    # set), and the measured HBM copy rate beside the nominal peak the roofline uses: 4 GiB
    # read + 4 GiB written per copy (16x the Infinity Cache).
    gather_peak = hbm_copy = l1_probe = valu_probe = None
    if not a.no_probes:
        gather_peak = ctx.probe_gather()
        l1_probe = {name: ctx.probe_l1(i)[1] for i, name in enumerate(
            ("divergent", "quad_coalesced", "half_masked", "half_out_of_range", "quad_coalesced_lds_dma"))}
        valu_probe = {name: ctx.probe_valu(i) for i, name in enumerate(("fma_f32", "rcp_f32", "mwc_step"))}
        log("VALU probe, SIMD cycles per wave instruction: " + ", ".join(f"{k} {v[1]:.2f}" for k, v in valu_probe.items()))
        log("L1 probe, CU cycles per 16-byte wave load: " + ", ".join(f"{k} {v:.1f}" for k, v in l1_probe.items()))
        nb = 4 << 30
        ms = ctx.probe_copy(nb, 5)
        hbm_copy = 2 * nb / (ms * 1e-3) / 1e9
        log(f"gather ceiling {gather_peak:.3e} lane-loads/s; HBM copy {hbm_copy:.0f} GB/s (read + write)")

    # The bump mesh (6.29 M tets, 805 MB of records: larger than the Infinity Cache, the one
    # workload where HBM can bind) on a sample of its B-scan, with the profiled launch options.
    bump = None
    if a.workload == "slab4" and a.bump_stride > 0:
        ts = time.time()
        bm, bpar, bscans, bseed = configs.bump()
        bscans = bscans[::a.bump_stride]
        bmid = ctx.upload(bm)
        bo = O.BatchOpts()
        # The autotuner's choice for this mesh (octmps_autotune on a 4e6-photon slice of the
        # sample): 4 blocks per CU, fewer photons in flight than the occupancy maximum, since
        # the records do not fit the caches; the profiled options.
        bq = [int(v) for v in a.bump_quanta.split(",")]
        bo.use_graph, bo.record_fetch = 1, 2
        bo.scatter_quantum, bo.refill_quantum, bo.blocks_per_cu = bq[0], bq[1], bq[2]
        _, bst = ctx.run_bscan(bmid, bpar, bscans, bseed, opts=bo)
        ctx.release(bmid)
        del bm
        bsteps, bms = bst.photon_steps, bst.kernel_ms
        b_alg = bsteps * ALG_BYTES_PER_STEP / (bms * 1e-3) / 1e9
        b_hbm = bsteps * PROFILE["bump_hbm_bytes_per_step"] / (bms * 1e-3) / 1e9
        bump = {"workload": f"bump: 6x6x3 mm, 128x128x64 cells x 6 tets (6,291,456 tets), {len(bscans)} of 1024 "
                            f"A-scans x 1e6 photons, 1024 bins to 2 mm, quanta {bq[0]}/{bq[1]}, {bq[2]} blocks per "
                            f"CU, cooperative record fetch",
                "photons": bst.photons, "kernel_ms": bms, "photons_per_s": bst.photons / (bms * 1e-3),
                "photon_steps": bsteps, "steps_per_photon": bsteps / bst.photons,
                "roofline": {"bound": "hbm", "achieved": b_alg, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": b_alg / HBM_PEAK_GBS, "per_unit": f"{ALG_BYTES_PER_STEP} B per photon-step",
                             "measured_hbm": {"bytes_per_step": PROFILE["bump_hbm_bytes_per_step"], "GB_s": b_hbm,
                                              "frac": b_hbm / HBM_PEAK_GBS,
                                              "source": PROFILE["bump_hbm_source"]}}}
        log(f"bump sample: {bst.photons} photons in {bms:.0f} ms kernel ({bump['photons_per_s']:.3e} photons/s), "
            f"HBM {b_hbm:.0f} GB/s measured-bytes basis ({time.time() - ts:.0f} s with the mesh build)")

    # Timed steps: full B-scans, as many of the requested ones as fit in the budget.
    walls, kms, steps, iters, bad = [], [], [], [], 0
    t = None
    for i in range(a.steps):
        if walls:
            est = max(walls) * 1.05 + 2.0
            if time.time() + est > deadline:
                log(f"budget: {i} of {a.steps} steps timed ({deadline - time.time():.0f} s left, a step takes "
                    f"{max(walls):.1f} s)")
                break
        ts = time.perf_counter()
        t, st = ctx.run_bscan(mid, params, scans, seed, opts=opts)
        walls.append(time.perf_counter() - ts)
        path = "coop" if st.record_fetch == 2 else ("lane_skip_on" if st.entry_skip else "lane_skip_off")
        kms.append(st.kernel_ms)
        steps.append(st.photon_steps)
        iters.append(st.wave_iterations)
        bad += st.bad_contributions
        log(f"step {i}: {walls[-1]:.3f}s kernel {st.kernel_ms:.1f} ms, {st.photon_steps} photon-steps, "
            f"killed {st.killed}")
    n_steps = len(walls)
    wall = sum(walls)
    value = total_photons * n_steps / wall
    kernel_s = sum(kms) * 1e-3
    steps_per_launch = sum(steps) / n_steps
    launch_s = kernel_s / n_steps
    prof = PROFILE[path]
    achieved = steps_per_launch * ALG_BYTES_PER_STEP / launch_s / 1e9
    hbm_measured = steps_per_launch * prof["hbm_bytes_per_step"] / launch_s / 1e9
    lane_util = sum(steps) / (64.0 * max(1, sum(iters)))
    if served_same is not None:
        served_same["equal_to_batched"] = bool(np.array_equal(served_out, t))

    gather = None
    if gather_peak:
        # The L1 looks up one line per lane of a divergent vector load per cycle (four lanes
        # reading one 64-byte piece share a lookup): the probe's rate of divergent 16-byte
        # loads is the ceiling of L1 lookups.
        got = sum(steps) / kernel_s * prof["l1_lookups_per_step"]
        gather = {"bound": "L1 lookups", "achieved": got, "peak": gather_peak,
                  "unit": "L1 lookups/s", "frac": got / gather_peak,
                  "l1_lookups_per_step": prof["l1_lookups_per_step"], "record_path": path,
                  "peak_source": "octmps_probe_gather: 5 x 16 B from pseudo-random 128-byte records of a "
                                 "2 MiB set, 6 waves per SIMD",
                  "l1_probe_cu_cycles_per_wave_load": l1_probe}
    ctx.close()
    frac = achieved / HBM_PEAK_GBS
    if path == "coop":
        # Records are loaded ahead of use, so the loop is bound by issuing its instructions: the
        # share of each wave's cycles spent issuing or ready and waiting for an issue slot,
        # against waiting on data (SQ counters of the profiled launch).
        bound_by = {"resource": "instruction issue (VALU and SALU instructions and L1 lookups together; waiting on "
                                "memory is the smaller share)",
                    "frac": prof["wave_issuing"] + prof["wave_wait_issue"],
                    "frac_of": "share of each wave's cycles spent issuing or ready and waiting for an issue slot "
                               "(SQ_ACTIVE_INST_ANY + SQ_WAIT_INST_ANY over SQ_WAVE_CYCLES); the rest, "
                               f"{prof['wave_wait_dependency']:.2f}, waits on a dependency"}
    else:
        bound_by = {"resource": "L1 lookups of the per-lane record gathers (one per lane per 16-byte load)",
                    "frac": gather["frac"] if gather else None,
                    "frac_of": "L1 lookups per second over the divergent-gather ceiling of octmps_probe_gather"}
    bound_by.update({"record_path": path, "counters": {k: v for k, v in prof.items() if k != "source"},
                     "source": prof["source"], "lane_utilisation": lane_util, "gather_frac": gather["frac"] if gather else None})
    if valu_probe:
        # VALU instructions the launch retired per second, against a stream of independent FP32
        # FMAs (octmps_probe_valu): above 1 means most of its VALU instructions issue in fewer
        # SIMD cycles than an FMA; the probe's cycle costs per class are listed beside it.
        valu_rate = prof["valu_per_step"] * sum(steps) / kernel_s
        bound_by["valu"] = {"insts_per_s": valu_rate, "fma_stream_insts_per_s": valu_probe["fma_f32"][0],
                            "ratio_to_fma_stream": valu_rate / valu_probe["fma_f32"][0],
                            "probe_simd_cycles_per_inst": {k: v[1] for k, v in valu_probe.items()}}
    line = {
        "metric": "photons/s per B-scan, batched vs plain kernel; p50/p99 A-scan request latency",
        "value": value,
        "unit": "photons/s",
        "n_gpus": 1,
        "steps": n_steps,
        "steps_requested": a.steps,
        "warmup": a.warmup,
        "ms_per_step": wall / n_steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": desc.format(n=len(scans)),
                   "photons": total_photons, "path": "batched persistent kernel, hipGraph launch",
                   "launch_options": tuned or "context defaults",
                   "warmup_slice": f"1/{div} of every A-scan's photons"},
        "batched": {"photons_per_s": value, "kernel_ms": kms, "photon_steps": steps,
                    "steps_per_photon": steps_per_launch / total_photons, "lane_utilisation": lane_util,
                    "bad_contributions": bad},
        "plain": plain,
        "copy_calibration": copy,
        "batched_over_plain": (value / plain["photons_per_s"]) if plain else None,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": frac,
                     "traffic": steps_per_launch * prof["hbm_bytes_per_step"],
                     "per_unit": f"{ALG_BYTES_PER_STEP} B per photon-step",
                     "binding": frac <= 1.0,
                     "note": ("algorithmic bytes (one tet record per photon-step) over the nominal HBM peak; "
                              "above 1 because the mesh is served mostly from L2 and the Infinity Cache, so HBM "
                              "does not bind" if frac > 1.0 else
                              "algorithmic bytes (one tet record per photon-step) over the nominal HBM peak"),
                     "peak_kind": "nominal MI355X HBM3E", "measured_copy_GB_s": hbm_copy,
                     "layout_bytes_per_step": LAYOUT_BYTES_PER_STEP,
                     "measured_hbm": {"bytes_per_step": prof["hbm_bytes_per_step"], "GB_s": hbm_measured,
                                      "frac": hbm_measured / HBM_PEAK_GBS, "source": prof["source"]}},
        "bound_by": bound_by,
        "gather_roofline": gather,
        "bump_sample": bump,
        "cpu_baseline": cpu,
        "serving": serving,
        "served_same_bscan": served_same,
        "wall_s": time.time() - T_START,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
