"""Writes workload_traces.json: the per-position photon budgets of the ellipsoid B-scan
(512 positions, log-uniform in [1e5, 1e7], seed 7) and the served request traces (Poisson
arrivals, seed 11) that the benchmark runs, so that a change of generator or of numpy's
random streams shows up as a test failure instead of a silently different workload.

Run from the repository root: python tests/golden/make_workload_traces.py"""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oct-mps_amd"))

from octmps import configs  # noqa: E402

STREAMS = [(50.0, 0.1), (100.0, 0.1), (200.0, 0.05), (400.0, 0.025)]


def traces():
    out = {"ellipsoid_budgets_seed7": [int(b) for b in configs.ellipsoid_budgets()], "served_seed11": []}
    for rate, window in STREAMS:
        out["served_seed11"].append({"rate": rate, "window": window, "requests": [
            {"t": t, "mesh": m, "n_scans": n, "budgets": [int(b) for b in bud]}
            for t, m, n, bud in configs.served_stream(rate, window)]})
    return out


if __name__ == "__main__":
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "workload_traces.json"), "w") as f:
        json.dump(traces(), f, separators=(",", ":"))
