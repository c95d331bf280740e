"""Benchmark: B-scan photons/s of the batched persistent kernel vs the plain one-A-scan-per-
launch path, on the four-layer slab B-scan (786,432 tets, 128 A-scans x 1e6 photons, 512
bins to 1.5 mm), with the roofline of the transport kernel and the CPU oracle's rate on the
host cores beside it.

A step is one full B-scan (all A-scans of the workload) through the batched path, launched
as one hipGraph. The plain path (one A-scan per launch, synchronous) is timed after the steps
on every 8th A-scan of the same B-scan. Then the CPU oracle runs a bounded sample of the
same B-scan on the host cores, and a short served stream (Poisson arrivals over four
resident meshes) gives p50/p99 request latency. Prints ONE JSON line on stdout; progress
goes to stderr.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "oct-mps_amd"))

ALG_BYTES_PER_STEP = 84  # one tet record per photon-step: 4 planes x 16 B + 4 links x 4 B + region 4 B
HBM_PEAK_GBS = 8000.0    # MI355X HBM3E nominal
LAYOUT_BYTES_PER_STEP = 80  # this build's packed tet record: 4 planes x 16 B + 4 packed neighbours x 4 B
# Measured HBM bytes per photon-step of the batched slab4 launch: rocprofv3 --pmc FETCH_SIZE
# (x2.0, the factor by which that counter under-reads a copy kernel of known size) plus
# WRITE_SIZE (x1.0), divided by the launch's photon-steps. profiles/transport_slab4_pmc.md
# holds the counter values and commands.
TRAFFIC_B_PER_STEP = 26.09
# Measured L1 lookups per photon-step of the batched slab4 launch (rocprofv3 --pmc
# TCP_TOTAL_CACHE_ACCESSES_sum / photon-steps, profiles/transport_slab4_memory.md): the
# 16-byte lane-loads of the tet record (5 per visit, 4 after entering through a face) plus
# the other vector accesses. With octmps_probe_gather's ceiling for cached divergent record
# gathers this gives the fraction of the gather path the kernel uses.
LANE_LOADS_PER_STEP = 4.16


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--scans", type=int, default=128, help="A-scans in the B-scan (workload: 128)")
    ap.add_argument("--photons", type=int, default=1_000_000, help="photons per A-scan (workload: 1e6)")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-photons", type=int, default=3000, help="oracle photons per sampled A-scan")
    ap.add_argument("--cpu-scans", type=int, default=8)
    ap.add_argument("--no-autotune", action="store_true",
                    help="skip the autotuner (context defaults: scatter/refill quanta 40/16, block 256)")
    ap.add_argument("--quanta", default="",
                    help="scatter,refill quanta to launch with instead of autotuning (e.g. 36,16: the same "
                         "launch a profiler run can then attribute to one kernel dispatch)")
    ap.add_argument("--serve-rate", type=float, default=50.0, help="Poisson arrival rate, requests/s")
    ap.add_argument("--serve-window", type=float, default=0.1, help="seconds of arrivals in the served stream")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-policy", type=int, default=0, choices=(0, 1),
                    help="request server hand-out order: 0 least remaining photons first, 1 equal shares")
    ap.add_argument("--serve-same", action="store_true",
                    help="also submit the B-scan as one request to the server and compare its tallies")
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
    return ap.parse_args()


def run_serving(O, configs, ctx, a, log):
    """Served stream: Poisson arrivals (seed 11) over four resident meshes; each request is
    1-256 A-scans with log-uniform budgets in [1e4, 1e6]. Latency = submit -> result."""
    import threading

    import numpy as np

    meshes = []
    t0 = time.time()
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
    log(f"served meshes resident ({sum(x[0].num_tets for x in meshes)} tets) in {time.time() - t0:.1f}s")
    trace = configs.served_stream(a.serve_rate, a.serve_window, seed=11, n_meshes=len(meshes))
    srv = O.Server(ctx, policy=a.serve_policy)
    ids = []
    photons = 0
    start = time.perf_counter()
    for i, (t_arr, mi, n, budgets) in enumerate(trace):
        mesh, mid, p, probe = meshes[mi]
        scans = [O.ascan(probe.x, probe.y, probe.z, probe.root_tet, int(b), salt=1000 * i + k)
                 for k, b in enumerate(budgets)]
        photons += int(sum(budgets))
        delay = start + t_arr - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        ids.append(srv.submit(mid, p, scans, 1))
    lat = []
    first = last = None
    for rid in ids:
        info, _ = srv.wait(rid, -1.0)
        assert info.status == 0
        lat.append(info.t_done_ms - info.t_submit_ms)
        first = info.t_submit_ms if first is None else min(first, info.t_submit_ms)
        last = info.t_done_ms if last is None else max(last, info.t_done_ms)
    st = srv.stats()
    # Latency floor: one small request (1 A-scan x 1e4 photons, sphere) on the idle server,
    # five times.
    mesh, mid, p, probe = meshes[0]
    floor = []
    for k in range(5):
        small = [O.ascan(probe.x, probe.y, probe.z, probe.root_tet, 10000, salt=999000 + k)]
        info, _ = srv.wait(srv.submit(mid, p, small, 1), -1.0)
        floor.append(info.t_done_ms - info.t_submit_ms)
    srv.close()
    lat = np.array(lat)
    span_s = (last - first) * 1e-3
    return {"requests": len(ids), "rate_req_s": a.serve_rate, "arrival_window_s": a.serve_window,
            "policy": ("least remaining photons first, aged", "equal shares")[a.serve_policy],
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "idle_small_request_ms": float(np.median(floor)),
            "photons": photons, "photons_per_s": photons / span_s, "span_s": span_s,
            "slices": st["slices"], "kernel_busy_frac": st["kernel_ms"] * 1e-3 / span_s,
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
    print(f"[bench] mesh {mesh.num_tets} tets built in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    ctx = O.Context(0)
    mid = ctx.upload(mesh)
    copy = None
    if a.calibrate_copy_mb:
        nbytes = a.calibrate_copy_mb << 20
        ms = ctx.probe_copy(nbytes, 10)
        copy = {"bytes_read": nbytes, "bytes_written": nbytes, "ms": ms, "GB_s": 2 * nbytes / (ms * 1e-3) / 1e9}
        print(f"[bench] copy calibration: {nbytes} B in {ms:.3f} ms ({copy['GB_s']:.0f} GB/s r+w)",
              file=sys.stderr, flush=True)
    opts = O.BatchOpts()
    opts.use_graph = 1
    tuned = None
    if a.quanta:
        opts.scatter_quantum, opts.refill_quantum = (int(v) for v in a.quanta.split(","))
        tuned = {"scatter_quantum": opts.scatter_quantum, "refill_quantum": opts.refill_quantum, "source": "--quanta"}
    elif not a.no_autotune:
        # The autotuner (coordinate search over launch options on a calibration slice of the
        # same B-scan: the first 1/16 of every A-scan's photons at the default size), before
        # the timed region; results are the same bits under any options.
        calib = max(1, sum(s.n_photons for s in scans) // 16)
        ts = time.perf_counter()
        o, rate = ctx.autotune(mid, params, scans, seed, calib=calib)
        opts.block_size, opts.refill_quantum, opts.scatter_quantum = o.block_size, o.refill_quantum, o.scatter_quantum
        opts.tally_copies, opts.entry_skip = o.tally_copies, o.entry_skip
        tuned = {"block_size": o.block_size, "refill_quantum": o.refill_quantum,
                 "scatter_quantum": o.scatter_quantum, "tally_copies": o.tally_copies, "entry_skip": o.entry_skip,
                 "calibration_photons": calib, "seconds": time.perf_counter() - ts}
        print(f"[bench] autotune ({tuned['seconds']:.1f} s): block {o.block_size} scatter/refill quanta "
              f"{o.scatter_quantum}/{o.refill_quantum} tally copies {o.tally_copies} entry skip {o.entry_skip} "
              f"({rate:.3e} photons/s on the calibration slice)", file=sys.stderr, flush=True)
    # Warmup on a small slice of the same B-scan (code object load, buffer allocation), then W
    # full steps. The slice runs through the logged variant of the kernel, a different kernel
    # symbol, so a profiler's per-kernel average of the unlogged single-set transport_kernel
    # covers the timed launches only.
    small = [O.ascan(s.x, s.y, s.z, s.root_tet, min(s.n_photons, 500), s.stream_salt) for s in scans]
    ctx.run_bscan(mid, params, small, seed, opts=opts, log=True)
    for i in range(a.warmup):
        ctx.run_bscan(mid, params, scans, seed, opts=opts)
        print(f"[bench] warmup step {i} done", file=sys.stderr, flush=True)
    total_photons = sum(s.n_photons for s in scans)
    walls, kms, steps, iters = [], [], [], []
    for i in range(a.steps):
        ts = time.perf_counter()
        t, st = ctx.run_bscan(mid, params, scans, seed, opts=opts)
        walls.append(time.perf_counter() - ts)
        kms.append(st.kernel_ms)
        steps.append(st.photon_steps)
        iters.append(st.wave_iterations)
        print(f"[bench] step {i}: {walls[-1]:.3f}s kernel {st.kernel_ms:.1f} ms, {st.photon_steps} photon-steps, "
              f"killed {st.killed}", file=sys.stderr, flush=True)
    wall = sum(walls)
    value = total_photons * a.steps / wall
    kernel_s = sum(kms) * 1e-3
    steps_per_launch = sum(steps) / len(steps)
    achieved = steps_per_launch * ALG_BYTES_PER_STEP / (kernel_s / a.steps) / 1e9
    batched_tally = t

    plain = None
    if not a.no_plain:
        # One A-scan per launch, synchronous, tallies copied back each time. Every A-scan of
        # the B-scan is the same amount of work, so the rate over every k-th A-scan is the
        # plain B-scan rate.
        sub = scans[::max(1, a.plain_every)]
        ts = time.perf_counter()
        pk, pst, nph = 0.0, 0, 0
        for s in sub:
            tp, stp = ctx.run_ascan(mid, params, s, seed)
            pk += stp.kernel_ms
            pst += stp.photon_steps
            nph += stp.photons
        pw = time.perf_counter() - ts
        plain = {"photons_per_s": nph / pw, "wall_s": pw, "kernel_ms": pk, "photon_steps": pst,
                 "a_scans_timed": len(sub), "photons": nph}
        print(f"[bench] plain: {len(sub)} A-scans in {pw:.3f}s", file=sys.stderr, flush=True)

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
        print(f"[bench] cpu oracle: {n} photons in {cw:.1f}s; same photons on the GPU: summed reflectance "
              f"{gpu_sum:.6e} vs oracle {oracle_sum:.6e}", file=sys.stderr, flush=True)

    served_same = None
    if a.serve_same:
        # The same B-scan as one request to the request server (the multi-set kernel in
        # wall-clock slices), after the timed region: its rate beside the batched launch and
        # its tallies against the batched launch's, bit for bit.
        srv = O.Server(ctx)
        try:
            ts = time.perf_counter()
            info, out = srv.wait(srv.submit(mid, params, scans, seed), -1.0)
            sw = time.perf_counter() - ts
            sst = srv.stats()
        finally:
            srv.close()
        served_same = {"photons_per_s": total_photons / sw, "wall_s": sw, "slices": sst["slices"],
                       "status": info.status, "equal_to_batched": bool(np.array_equal(out, batched_tally))}
        print(f"[bench] B-scan as one served request: {total_photons / sw:.3e} photons/s, "
              f"equal to batched: {served_same['equal_to_batched']}", file=sys.stderr, flush=True)

    serving = None
    if not a.no_serve:
        serving = run_serving(O, configs, ctx, a, lambda msg: print(f"[bench] {msg}", file=sys.stderr, flush=True))
        print(f"[bench] served stream: {serving['requests']} requests, p50 {serving['p50_ms']:.1f} ms, "
              f"p99 {serving['p99_ms']:.1f} ms", file=sys.stderr, flush=True)

    # Ceiling of the vector-memory path for divergent record gathers (measured after the
    # timed region, on a cache-resident record set).
    gather = None
    if a.workload == "slab4":
        peak = ctx.probe_gather()
        got = sum(steps) / (sum(kms) * 1e-3) * LANE_LOADS_PER_STEP
        gather = {"bound": "vector-memory gathers", "achieved": got, "peak": peak, "unit": "16-byte lane-loads/s",
                  "frac": got / peak, "lane_loads_per_step": LANE_LOADS_PER_STEP,
                  "peak_source": "octmps_probe_gather: 5 x 16 B from pseudo-random 128-byte records of a "
                                 "2 MiB set, 6 waves per SIMD"}
        print(f"[bench] gather ceiling {peak:.3e} lane-loads/s; kernel {got:.3e} ({got / peak:.2f})",
              file=sys.stderr, flush=True)
    # Measured HBM copy rate on this GPU, beside the nominal peak the roofline uses: 4 GiB
    # read + 4 GiB written per copy (16x the Infinity Cache), after the timed region.
    hbm_copy = None
    if a.workload == "slab4":
        nb = 4 << 30
        ms = ctx.probe_copy(nb, 5)
        hbm_copy = 2 * nb / (ms * 1e-3) / 1e9
        print(f"[bench] HBM copy: {hbm_copy:.0f} GB/s (read + write)", file=sys.stderr, flush=True)
    ctx.close()
    line = {
        "metric": "photons/s per B-scan, batched vs plain kernel; p50/p99 A-scan request latency",
        "value": value,
        "unit": "photons/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": wall / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": desc.format(n=len(scans)),
                   "photons": total_photons, "path": "batched persistent kernel, hipGraph launch",
                   "launch_options": tuned or "context defaults"},
        "batched": {"photons_per_s": value, "kernel_ms": kms, "photon_steps": steps,
                    "steps_per_photon": steps_per_launch / total_photons,
                    "lane_utilisation": sum(steps) / (64.0 * max(1, sum(iters)))},
        "plain": plain,
        "copy_calibration": copy,
        "batched_over_plain": (value / plain["photons_per_s"]) if plain else None,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": (steps_per_launch * TRAFFIC_B_PER_STEP) if TRAFFIC_B_PER_STEP else None,
                     "per_unit": f"{ALG_BYTES_PER_STEP} B per photon-step",
                     "peak_kind": "nominal MI355X HBM3E", "measured_copy_GB_s": hbm_copy,
                     "layout_bytes_per_step": LAYOUT_BYTES_PER_STEP},
        "gather_roofline": gather,
        "cpu_baseline": cpu,
        "serving": serving,
        "served_same_bscan": served_same,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
