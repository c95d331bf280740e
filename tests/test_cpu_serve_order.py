"""The request server's hand-out order (oct-mps_amd/csrc/serve_order.hpp), compiled with
the host compiler and driven with scenarios: shortest remaining first for a burst, aging
that lets a long-waiting large request overtake newer small ones, the slice budget, and
equal shares reaching every open A-scan in one round."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

DRIVER = r"""
#include <cstdio>
#include "serve_order.hpp"
int main() {
    int policy; double now; unsigned long long budget, round; int n;
    if (std::scanf("%d %lf %llu %llu %d", &policy, &now, &budget, &round, &n) != 5) return 2;
    std::vector<octmps::OpenScan> open(n);
    for (auto &o : open) {
        unsigned long long id, from, rem; unsigned slot;
        if (std::scanf("%llu %lf %llu %llu %u", &id, &o.t_submit_ms, &from, &rem, &slot) != 5) return 2;
        o.req_id = id; o.from = from; o.rem = rem; o.slot = slot;
    }
    std::vector<octmps::Segment> segs;
    unsigned long long total = policy == 0 ? octmps::order_least_remaining(open, now, budget, segs)
                                           : octmps::order_equal_shares(open, round, budget, segs);
    std::printf("%llu\n", total);
    for (auto &g : segs) std::printf("%llu %llu %u %zu\n", (unsigned long long)g.start, (unsigned long long)g.poff, g.slot, g.open);
    return 0;
}
"""


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    d = tmp_path_factory.mktemp("serve_order")
    (d / "driver.cpp").write_text(DRIVER)
    exe = d / "driver"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "oct-mps_amd", "csrc"),
                    str(d / "driver.cpp"), "-o", str(exe)], check=True)

    def run(policy, now, budget, open_scans, round_=0):
        lines = [f"{policy} {now} {budget} {round_} {len(open_scans)}"]
        lines += [" ".join(str(v) for v in o) for o in open_scans]
        out = subprocess.run([str(exe)], input="\n".join(lines) + "\n", capture_output=True, text=True,
                             check=True).stdout.split("\n")
        total = int(out[0])
        segs = [tuple(int(v) for v in l.split()) for l in out[1:] if l.strip()]
        return total, segs

    return run


def lengths(total, segs):
    """(open index, ids listed) per segment."""
    starts = [s[0] for s in segs] + [total]
    return [(segs[i][3], starts[i + 1] - starts[i]) for i in range(len(segs))]


def test_burst_goes_shortest_request_first(driver):
    # three requests submitted together: A 2 scans x 500, B 1 scan x 100, C 1 scan x 300
    open_scans = [(1, 0.0, 0, 500, 10), (1, 0.0, 0, 500, 11), (2, 0.0, 0, 100, 12), (3, 0.0, 7, 300, 13)]
    total, segs = driver(0, 5.0, 10_000, open_scans)
    assert total == 1400
    assert [s[2] for s in segs] == [12, 13, 10, 11]          # B, C, then A's two A-scans
    assert segs[1][1] == 7                                    # C resumes at its first unclaimed photon
    assert lengths(total, segs) == [(2, 100), (3, 300), (0, 500), (1, 500)]


def test_ties_go_by_submission_order(driver):
    total, segs = driver(0, 0.0, 10_000, [(5, 0.0, 0, 200, 1), (4, 0.0, 0, 200, 2)])
    assert [s[2] for s in segs] == [2, 1]


def test_aging_lets_a_long_waiting_large_request_overtake(driver):
    # A: 1e6 photons submitted at t = 0; B: 2.1e5 photons submitted at t = 40 s. At 40 s,
    # A counts 1e6 / (1 + 4) = 2e5 < 2.1e5: A first. Ten seconds earlier (now = 30 s, B
    # just submitted) A still counts 2.5e5 > 2.1e5: B first.
    open_scans = [(1, 0.0, 0, 1_000_000, 1), (2, 40_000.0, 0, 210_000, 2)]
    _, segs = driver(0, 40_000.0, 10_000_000, open_scans)
    assert [s[2] for s in segs] == [1, 2]
    open_scans = [(1, 0.0, 0, 1_000_000, 1), (2, 30_000.0, 0, 210_000, 2)]
    _, segs = driver(0, 30_000.0, 10_000_000, open_scans)
    assert [s[2] for s in segs] == [2, 1]


def test_budget_truncates_the_listing(driver):
    open_scans = [(1, 0.0, 0, 700, 1), (2, 0.0, 0, 900, 2), (3, 0.0, 0, 5000, 3)]
    total, segs = driver(0, 0.0, 1000, open_scans)
    assert total == 1000
    assert lengths(total, segs) == [(0, 700), (1, 300)]
    total, segs = driver(1, 0.0, 1000, open_scans, round_=768)
    assert total == 1000 and lengths(total, segs) == [(0, 256), (1, 256), (2, 256), (0, 232)]


def test_equal_shares_reach_every_open_scan_in_one_round(driver):
    open_scans = [(i, 0.0, 100 * i, 100_000, i) for i in range(1, 6)]
    total, segs = driver(1, 0.0, 1_000_000, open_scans, round_=50_000)
    first = lengths(total, segs)[:5]
    assert [k for k, _ in first] == [0, 1, 2, 3, 4] and all(n == 10_000 for _, n in first)
    assert total == 500_000                                   # everything listed within the budget
    assert segs[5][1] == 100 + 10_000                         # second round continues each A-scan
    # chunks never fall below 256 ids
    total, segs = driver(1, 0.0, 1_000_000, open_scans, round_=100)
    assert all(n == 256 for _, n in lengths(total, segs)[:5])


def test_nothing_open_lists_nothing(driver):
    assert driver(0, 0.0, 1000, []) == (0, [])
    assert driver(1, 0.0, 1000, [], round_=1000) == (0, [])
