"""The device's detector tally on chosen contributions, checked against exact arithmetic."""
from fractions import Fraction

import numpy as np
import pytest

import octmps as O

pytestmark = pytest.mark.gpu


def _params(B=64):
    return O.make_params([(1.4, 1.0, 100.0, 0.9)], num_bins=B, depth=0.1, type_bias=37)


def _op_for_bin(p, v):
    # optical path whose depth op/2 falls in the middle of bin v (bin width Lc/6)
    w = p.coherence_length_source / 6
    return 2.0 * (v + 0.5) * w


def test_device_tally_is_exact_over_the_fp32_range(ctx):
    # Contributions from a subnormal up to 3e38 in the same and in different bins: the
    # device's integer tally equals the exact rational Sum and SumSq of what it received, so
    # each bin's value does not depend on the order of the adds or on the spread of sizes.
    p = _params()
    B = p.num_depth_bins
    rng = np.random.default_rng(11)
    cs = np.concatenate([[1e-16, 1e5, 3e-40, 3e38, 1.0, 0.0, 7.5e-20, 4.1e9, 6.3e4],
                         np.float32(10.0) ** rng.uniform(-30, 30, 500).astype(np.float32)]).astype(np.float32)
    bins = rng.integers(0, B, len(cs))
    bins[:9] = [10, 10, 10, 30, 30, 30, 50, 50, 50]
    op = np.array([_op_for_bin(p, v) for v in bins], np.float32)
    maxz = np.where(rng.random(len(cs)) < 0.5, 1.0, 0.0).astype(np.float32)  # Class I / Class II
    t, bad = ctx.probe_tally(p, op, maxz, cs)
    assert bad == 0 and t[3] == 0
    exp_s = [[Fraction(0)] * B for _ in range(2)]
    exp_q = [[Fraction(0)] * B for _ in range(2)]
    exp_n = np.zeros((2, B), np.uint64)
    exp_max = np.zeros((2, B))
    for c, v, mz, o in zip(cs, bins, maxz, op):
        depth = float(np.float32(o) * np.float32(0.5))
        cls = 0 if mz > depth - p.coherence_length_source / 2 else 1
        for b in range(v - 3, v + 3):
            if 0 <= b < B:
                exp_s[cls][b] += Fraction(float(c))
                exp_q[cls][b] += Fraction(float(c)) ** 2
                exp_n[cls, b] += 1
                exp_max[cls, b] = max(exp_max[cls, b], float(c))
    got = O.exact_tally_sums(t, B)
    dec = O.decode_tally(t, B)
    for cls in range(2):
        for b in range(B):
            assert got[cls][b] == (exp_s[cls][b], exp_q[cls][b]), (cls, b)
    assert np.array_equal(dec["count"], exp_n.astype(np.float64))
    assert np.array_equal(dec["max"], exp_max)
    assert dec["num_filtered"] == len(cs)


def test_device_tally_counts_what_it_cannot_hold(ctx):
    # An infinite or NaN contribution cannot be held exactly: it is counted (tally word 3 and
    # the launch's bad-contribution counter) and the bins still count the detection.
    p = _params()
    cs = np.array([np.inf, np.nan, 0.5], np.float32)
    op = np.full(3, _op_for_bin(p, 20), np.float32)
    t, bad = ctx.probe_tally(p, op, np.ones(3, np.float32), cs)
    assert bad == 2 and t[3] == 2
    dec = O.decode_tally(t, p.num_depth_bins)
    assert dec["count"][0, 20] == 3 and dec["sum"][0, 20] == 0.5 and dec["max"][0, 20] == 0.5


def test_single_detection_bins_finalize_to_zero_filtered_se(ctx):
    # One device contribution per bin: the filtered outputs drop it exactly (SE 0, not NaN).
    p = _params(B=96)
    cs = np.array([1e-16, 3e-12, 0.25, 1e5], np.float32)
    op = np.array([_op_for_bin(p, v) for v in (10, 30, 50, 70)], np.float32)
    t, _ = ctx.probe_tally(p, op, np.ones(4, np.float32), cs)
    out = O.finalize(t, p.num_depth_bins, 1000, 0.01)
    assert np.all(np.isfinite(out))
    hit = O.decode_tally(t, p.num_depth_bins)["count"][0] > 0
    assert np.all(out[1, 0, 1][hit] == 0.0) and np.all(out[1, 0, 0][hit] == 0.0)
