"""Pins of the oracle's ComputeStepSize (src/octmps_kernel.cuh:505-515, with the leftover of
HitBoundary, :559-566) and Drop (:584-588) through per-photon event traces.

* Free path. A photon's path between two scatterings is one exponential draw in optical
  depth: tau = sum_i mu_t,i * l_i over the regions it crosses follows Exp(1), whatever faces
  it crosses, because a step cut short at a face carries its leftover in optical units and
  the next step converts it with the new region's mu_t. Paths cut by the mesh boundary never
  end in a scattering (index-matched ambient: the photon is lost there), so a complete path
  follows Exp(1) truncated at the ray's optical distance tau_D to the boundary, and
  (1 - exp(-tau)) / (1 - exp(-tau_D)) is uniform on (0, 1): a KS test over every complete
  path of a stack of layers whose mu_t alternates 50 / 200 per cm.
* Drop. Between two scatterings that only cross faces (no reflection, no roulette), the
  weight changes by the factor mus / (mua + mus) of the region of the second scattering,
  to rounding.

Each pin must reject a planted mistake, built from oracle.cpp on the CPU with a switch:
the leftover converted with the old region's mu_t, and Drop with mua / mus."""
import os
import subprocess

import numpy as np
import pytest
from scipy.stats import kstest

import octmps as O
import oracle_py as OR

ORACLE_SRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "oracle.cpp")
EV_LAUNCH, EV_SCATTER, EV_TRANSMIT, EV_REFLECT, EV_BOUNDARY, EV_SURVIVE = range(6)

N_LAYERS, DZ, WIDTH = 12, 0.01, 0.3          # cm
MU = ((2.0, 48.0), (10.0, 190.0))            # (mua, mus) of odd / even layers: mu_t 50 and 200 per cm


@pytest.fixture(scope="module")
def layered():
    H = N_LAYERS * DZ
    bottoms = [DZ * (k + 1) for k in range(N_LAYERS - 1)]
    mesh = O.Mesh.generate(O.GEOM_LAYERS, (15, 15, 2 * N_LAYERS), (WIDTH, WIDTH, H), p=tuple(bottoms),
                           n_layers=N_LAYERS)
    regs = [(1.4, MU[k % 2][0], MU[k % 2][1], 0.9) for k in range(N_LAYERS)]
    # Ambient index = tissue index: no Fresnel anywhere, the mesh boundary absorbs.
    params = O.make_params(regs, n_ambient=1.4, num_bins=64, depth=0.1, type_bias=0)
    root, z = mesh.find_root(0.0, 0.0, 0.0)
    zb = np.array([0.0] + bottoms + [H])
    mut = np.array([sum(MU[k % 2]) for k in range(N_LAYERS)])
    keep = np.array([MU[k % 2][1] / sum(MU[k % 2]) for k in range(N_LAYERS)])
    return mesh, mesh.soa(), params, O.ascan(0.0, 0.0, z, root, 1), zb, mut, keep, H


def _build_variant(tmp_path_factory, macro):
    out = tmp_path_factory.mktemp("oracle_plant") / f"liboracle_{macro}.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-pthread", f"-D{macro}", ORACLE_SRC, "-o",
                    str(out)], check=True)
    return str(out)


def _traces(layered, table, n=1500):
    _, soa, params, scan, *_ = layered
    return [OR.trace_photon(soa, params, scan, 1, table, i, max_events=20000) for i in range(n)]


def _paths_and_ratios(traces):
    """Complete free paths (start, direction, length) and weight ratios between scatterings
    joined only by face crossings (with the tet of the second scattering)."""
    starts, dirs, lens, ratios = [], [], [], []
    for tr in traces:
        seg, w_prev, clean = None, None, False
        for r in tr:
            ev = int(r[0])
            if ev == EV_LAUNCH:
                seg, w_prev, clean = (r[2:5].copy(), np.array([0.0, 0.0, 1.0]), r[10]), r[8], True
            elif ev == EV_TRANSMIT:
                continue
            elif ev == EV_SCATTER:
                if seg is not None:
                    starts.append(seg[0])
                    dirs.append(seg[1])
                    lens.append(r[10] - seg[2])  # optical path accumulates geometric length
                if clean:
                    ratios.append((r[8] / w_prev, int(r[1])))
                seg, w_prev, clean = (r[2:5].copy(), r[5:8].copy(), r[10]), r[8], True
            elif ev in (EV_REFLECT, EV_BOUNDARY):
                seg, clean = None, False  # direction changed, or cut by the mesh boundary
            elif ev == EV_SURVIVE:
                clean = False
            else:
                seg = None
    return np.array(starts), np.array(dirs), np.array(lens), ratios


def _optical_depth(p0, u, length, zb, mut):
    """Sum of mu_t over the straight path p0 + t u, 0 <= t <= length, through z-layers."""
    tau = np.zeros(len(length))
    z0, uz = p0[:, 2], u[:, 2]
    inside = lambda j: (z0 >= zb[j]) & (z0 < zb[j + 1])
    for j in range(len(mut)):
        with np.errstate(divide="ignore", invalid="ignore"):
            t1, t2 = (zb[j] - z0) / uz, (zb[j + 1] - z0) / uz
        ta = np.where(uz != 0, np.minimum(t1, t2), np.where(inside(j), 0.0, np.inf))
        tb = np.where(uz != 0, np.maximum(t1, t2), np.where(inside(j), np.inf, -np.inf))
        tau += mut[j] * np.clip(np.minimum(tb, length) - np.maximum(ta, 0.0), 0.0, None)
    return tau


def _exit_distance(p0, u, H):
    d = np.full(len(p0), np.inf)
    for ax, lo, hi in ((0, -WIDTH / 2, WIDTH / 2), (1, -WIDTH / 2, WIDTH / 2), (2, 0.0, H)):
        with np.errstate(divide="ignore", invalid="ignore"):
            t = np.where(u[:, ax] > 0, (hi - p0[:, ax]) / u[:, ax],
                         np.where(u[:, ax] < 0, (lo - p0[:, ax]) / u[:, ax], np.inf))
        d = np.minimum(d, t)
    return d


def _free_path_ks(layered, traces):
    _, _, _, _, zb, mut, _, H = layered
    p0, u, length, _ = _paths_and_ratios(traces)
    tau = _optical_depth(p0, u, length, zb, mut)
    tau_d = _optical_depth(p0, u, _exit_distance(p0, u, H), zb, mut)
    crossing = np.searchsorted(zb, p0[:, 2], side="right") != np.searchsorted(zb, p0[:, 2] + u[:, 2] * length,
                                                                              side="right")
    return kstest(-np.expm1(-tau) / -np.expm1(-tau_d), "uniform"), len(length), int(crossing.sum())


def _weight_ratio_error(layered, traces):
    _, soa, *_, keep, _ = layered
    ratios = _paths_and_ratios(traces)[3]
    got = np.array([r for r, _ in ratios])
    want = keep[soa[2][[t for _, t in ratios]] - 1]
    return len(got), float(np.max(np.abs(got / want - 1.0)))


def test_free_paths_follow_exp1_in_optical_depth_across_mu_t_changes(layered, table):
    ks, n, crossing = _free_path_ks(layered, _traces(layered, table))
    print("free paths", n, "crossing a layer", crossing, ks)
    assert n > 20000 and crossing > 5000
    assert ks.pvalue > 1e-3


def test_drop_keeps_mus_over_mut_of_the_weight(layered, table):
    n, err = _weight_ratio_error(layered, _traces(layered, table))
    assert n > 20000 and err < 1e-12


def test_free_path_pin_rejects_the_old_regions_mu_t_in_the_leftover(layered, table, tmp_path_factory):
    lib = _build_variant(tmp_path_factory, "ORACLE_PLANT_LEFTOVER_OLD_REGION")
    with OR.variant(lib):
        traces = _traces(layered, table)
    ks, n, _ = _free_path_ks(layered, traces)
    print("planted leftover", n, ks)
    assert ks.pvalue < 1e-20


def test_drop_pin_rejects_mua_over_mus(layered, table, tmp_path_factory):
    lib = _build_variant(tmp_path_factory, "ORACLE_PLANT_DROP_MUA_OVER_MUS")
    with OR.variant(lib):
        traces = _traces(layered, table)
    n, err = _weight_ratio_error(layered, traces)
    assert n > 20000 and err > 1e-4
    # the planted Drop leaves the free paths alone
    assert _free_path_ks(layered, traces)[0].pvalue > 1e-3
