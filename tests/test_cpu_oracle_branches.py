"""Known-answer pins of the oracle's branch rules, case by case: SpinBias's choice between
the first bias, the additional-bias mixture and plain HG (src/octmps_kernel.cuh:918-968),
the detector's spatial filter (:667-681) and the roulette with the continuation-photon
restore (:1376-1391). Each state is built to sit on one side of one condition of the
reference's text, and the expected outcome is read off that condition; a flipped
comparison, a non-strict bound or a dropped clause fails a case."""
import math

import numpy as np
import pytest

import octmps as O
import oracle_py as OR


def params(type_bias=37, rnd_step=0.0, radius=0.01, angle=5.0):
    # target depth (0, 0.1), probe at the origin, region 1: mua + mus = 101 per length unit
    return O.make_params([(1.4, 1.0, 100.0, 0.9)], num_bins=64, depth=0.1, type_bias=type_bias,
                         backward_bias_coefficient=0.925, probability_additional_bias=0.5,
                         max_collecting_angle_deg=angle, max_collecting_radius=radius,
                         rnd_step_size_in_tissue=rnd_step)


def unit(v):
    v = np.asarray(v, float)
    return v / np.linalg.norm(v)


def spin_bias(p, pos, u, fst, nbs, table, n=2000, lr=0.8, seed=3):
    state = [*pos, *unit(u), 0.5, lr, fst, nbs, 0.9, 101.0]
    return OR.spin_bias_batch(p, state, (0.0, 0.0, 0.0), seed, table, n)


FWD, BACK = (0.2, 0.1, 0.9), (0.3, 0.0, -0.95)
OFF_AXIS = (0.03, -0.02, 0.02)       # the probe is 56 deg off the -z axis: outside the 5 deg cone
BELOW = (0.0, 0.0, 0.05)             # straight below the probe: angle 0


@pytest.mark.parametrize("pos,u,fst,nbs,flag,fst_after", [
    # first bias: inside (min, max), going forward, not yet biased; sets the flag
    (BELOW, FWD, 0, 0, 1, 1),
    (OFF_AXIS, FWD, 0, 0, 1, 1),
    # plain HG: going backward, or outside the open interval, with no earlier bias or
    # backward specular reflection
    (BELOW, BACK, 0, 0, 0, 0),
    ((0.0, 0.0, 0.1), FWD, 0, 0, 0, 0),     # z == TargetDepthMax: strict bound
    ((0.0, 0.0, 0.15), FWD, 0, 0, 0, 0),
    ((0.01, 0.0, 0.0), FWD, 0, 0, 0, 0),    # z == TargetDepthMin: strict bound
    # biased before, or reflected backwards before: mixture (2), or biased towards the
    # probe (1) when the next step reaches it (rnd_step 0: infinite step) inside the cone;
    # neither sets nor clears the flag
    (OFF_AXIS, FWD, 1, 0, 2, 1),
    (BELOW, FWD, 1, 0, 1, 1),
    (OFF_AXIS, BACK, 0, 1, 2, 0),
    (BELOW, BACK, 0, 3, 1, 0),
])
def test_spin_bias_branch(table, pos, u, fst, nbs, flag, fst_after):
    out = spin_bias(params(), pos, u, fst, nbs, table)
    assert np.all(out[:, 4] == flag)
    assert np.all(out[:, 7] == fst_after)
    if flag == 0:
        # plain HG: no weight change, mean cosine to the old direction = g
        assert np.all(out[:, 3] == 1.0) and np.all(out[:, 6] == 0.8)
        c = out[:, :3] @ unit(u)
        assert abs(c.mean() - 0.9) < 5 * c.std() / math.sqrt(len(c))
    if flag == 1:
        # every draw is around v, the direction towards the probe: within 90 deg of it
        v = unit(-np.asarray(pos, float))
        assert np.all(out[:, :3] @ v >= -1e-12)


def test_spin_bias_next_step_against_distance(table):
    # Biased before, straight below the probe at distance 0.05: the bias towards the probe is
    # forced (1) when -log(rnd_step)/mut >= distance, the mixture (2) when it is shorter.
    mut, d = 101.0, 0.05
    for ratio, flag in ((2.0, 1), (1.0001, 1), (0.9999, 2), (0.5, 2)):
        p = params(rnd_step=math.exp(-mut * d * ratio))
        out = spin_bias(p, BELOW, FWD, 1, 0, table, n=200)
        assert np.all(out[:, 4] == flag), ratio
    out = spin_bias(params(rnd_step=1.0), BELOW, FWD, 1, 0, table, n=200)  # step 0
    assert np.all(out[:, 4] == 2)


def test_spin_bias_collection_cone_bound(table):
    # The probe seen at 4.9 deg from -z is inside the 5 deg cone (<=), at 5.1 deg outside.
    for deg, flag in ((4.9, 1), (5.1, 2)):
        t = math.radians(deg)
        pos = (0.05 * math.sin(t), 0.0, 0.05 * math.cos(t))
        out = spin_bias(params(), pos, FWD, 1, 0, table, n=200)
        assert np.all(out[:, 4] == flag), deg


R, ANG = 0.01, 5.0
IN = dict(x=0.006, y=0.0, uz=-1.0)                       # inside radius and cone
CASES = [
    # type_bias, position/direction, FstBack, LocationFstBias, MaxDepth, NumBackSpec, tallied
    (0, IN, 0, -1.0, 0.3, 0, 1),
    (3, IN, 0, -1.0, 0.3, 0, 1),
    (37, IN, 0, -1.0, 0.3, 0, 0),      # TypeBias 37 needs a backward bias or reflection
    (37, IN, 1, 0.2, 0.3, 0, 1),
    (37, IN, 0, -1.0, 0.3, 2, 1),
    (37, IN, 0, 0.3, 0.3, 0, 0),       # first bias at max depth counts only for TypeBias != 37
    (1, IN, 0, 0.3, 0.3, 0, 1),
    (1, IN, 0, 0.2, 0.3, 0, 0),
    (1, IN, 1, 0.2, 0.3, 0, 0),        # FstBack counts only for TypeBias 37
    (1, IN, 0, -1.0, 0.3, 1, 1),
    (0, dict(x=0.01, y=0.0, uz=-1.0), 1, 0.3, 0.3, 1, 0),     # radius == R: strict
    (37, dict(x=0.006, y=0.008, uz=-1.0), 1, 0.3, 0.3, 1, 0),  # radius 0.01 again, via x and y
    (0, dict(x=0.0, y=0.0, uz=-math.cos(math.radians(5.2))), 1, 0.3, 0.3, 1, 0),  # outside the cone
    (37, dict(x=0.0, y=0.0, uz=-math.cos(math.radians(4.8))), 1, 0.3, 0.3, 1, 1),
]


@pytest.mark.parametrize("tb,pos,fst,loc,maxd,nbs,tallied", CASES)
def test_detector_filter(tb, pos, fst, loc, maxd, nbs, tallied):
    p = params(type_bias=tb, radius=R, angle=ANG)
    assert OR.detect_filter(p, pos["x"], pos["y"], pos["uz"], fst, loc, maxd, nbs) == tallied


def test_detector_filter_is_relative_to_the_probe():
    p = params(type_bias=0, radius=R, angle=ANG)
    assert OR.detect_filter(p, 0.506, 0.2, -1.0, 0, -1.0, 0.3, 0, probe=(0.5, 0.2)) == 1
    assert OR.detect_filter(p, 0.006, 0.0, -1.0, 0, -1.0, 0.3, 0, probe=(0.5, 0.2)) == 0


PHOTON = [0.01, 0.02, 0.3, *unit((0.1, 0.2, -0.9)), 5e-5, 0.25, 0.3, 1]
CONT = [0.0, 0.001, 0.04, *unit((0.0, 0.3, 0.95)), 0.61, 0.9, 0.7, 0]


def roulette(table, tb, photon, cont=CONT, n=20000, g=0.8):
    return OR.roulette_batch(params(type_bias=tb), photon, cont, g, 5, table, n)


def test_roulette_leaves_heavy_photons_alone(table):
    ph = list(PHOTON)
    ph[6] = 1e-4                       # WEIGHT itself: not below it
    out = roulette(table, 37, ph, n=50)
    assert np.all(out[:, 0] == 0) and np.all(out[:, 7] == 1e-4)


@pytest.mark.parametrize("lr_after,lr_expect", [(0.3, 0.7), (1.0, 1.0), (1.7, 1.0)])
def test_roulette_survival_then_restore(table, lr_after, lr_expect):
    ph = list(PHOTON)
    ph[8] = lr_after
    out = roulette(table, 37, ph)
    fate = out[:, 0]
    assert set(np.unique(fate)) <= {1, 2}
    k = (fate == 1).mean()
    assert abs(k - 0.1) < 5 * math.sqrt(0.1 * 0.9 / len(fate))          # CHANCE
    surv = out[fate == 1]
    assert np.allclose(surv[:, 7], 5e-4, rtol=1e-12) and np.all(surv[:, 8] == 0.25)
    assert np.allclose(surv[:, 1:4], PHOTON[:3]) and np.all(surv[:, 9] == 1)
    rest = out[fate == 2]
    # the continuation photon: its position, weight and flag (taken before the first bias),
    # LR from the photon's LR after its first bias, direction HG-spun from the
    # continuation's direction (mean cosine g)
    assert np.all(rest[:, 1:4] == CONT[:3]) and np.all(rest[:, 7] == CONT[6])
    assert np.all(rest[:, 9] == 0)
    assert np.allclose(rest[:, 8], lr_expect, rtol=0, atol=1e-15)
    c = rest[:, 4:7] @ np.asarray(CONT[3:6])
    assert abs(c.mean() - 0.8) < 5 * c.std() / math.sqrt(len(c))


@pytest.mark.parametrize("tb,fst,w,fates", [
    (3, 1, 5e-5, {1, 3}),      # TypeBias 3 never restores
    (37, 0, 5e-5, {1, 3}),     # not biased backwards: dies
    (0, 0, 5e-5, {1, 3}),
    (37, 0, 0.0, {3}),         # zero weight never survives
    (37, 1, 0.0, {2}),         # ... but is restored when biased backwards
])
def test_roulette_death_and_restore_rules(table, tb, fst, w, fates):
    ph = list(PHOTON)
    ph[9], ph[6] = fst, w
    out = roulette(table, tb, ph, n=4000)
    assert set(np.unique(out[:, 0])) == fates
