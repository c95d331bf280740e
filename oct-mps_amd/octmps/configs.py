"""Synthetic workloads: meshes, optical parameters and A-scan lists.

Lengths are converted to the reference's units (cm, 1/cm; input/*/*.opt). Every mesh is a
structured grid with 6 tets per cell, centred on x = y = 0 with the tissue surface at z = 0.
Detector and bias settings follow the sample BiasStartup.mci files (TypeBias 37, a = 0.925,
p = 0.5, radius 10 um, 5 deg); bins are given explicitly and Lc = 6 * depth / bins.

Workloads (sizes in mm in the names):
  sphere    2x2x2 cube, 32^3 cells, sphere r 0.9 at depth 1.0; 256 bins over 0-1; 1 A-scan
  slab4     4x4x2, 64x64x32 cells, layers at 0.1/0.3/0.6; 128 A-scans at 20 um; 512 bins to 1.5
  ellipsoid 4x4x2, 64x64x32 cells, tilted ellipsoid + 2 spheres; 512 A-scans at 7.5 um with
            log-uniform budgets in [1e5, 1e7] (seed 7); 512 bins to 1.5
  slab128   4x4x4, 128^3 cells, the slab4 layers (resident mesh of the served stream)
  bump      6x6x3, 128x128x64 cells, Gaussian surface bump; 1024 A-scans at 5.5 um;
            1024 bins to 2
"""
import numpy as np

from . import (GEOM_BUMP, GEOM_ELLIPSOID2S, GEOM_LAYERS, GEOM_SPHERE, Mesh, ascan, make_params)

MM = 0.1  # cm per mm
INV_MM = 10.0  # 1/cm per 1/mm

BIAS = dict(type_bias=37, type_simulation=1, backward_bias_coefficient=0.925, probability_additional_bias=0.5,
            max_collecting_radius=10e-4, max_collecting_angle_deg=5.0, max_relative_contribution=0.01,
            rnd_step_size_in_tissue=0.0)


def _reg(n, mua_mm, mus_mm, g):
    return (n, mua_mm * INV_MM, mus_mm * INV_MM, g)


def sphere(cells=32, photons=1_000_000, bins=256):
    mesh = Mesh.generate(GEOM_SPHERE, (cells, cells, cells), (2 * MM, 2 * MM, 2 * MM),
                         p=(0.0, 0.0, 1.0 * MM, 0.9 * MM))
    params = make_params([_reg(1.40, 0.05, 10, 0.9), _reg(1.33, 0.02, 5, 0.9)], num_bins=bins, depth=1.0 * MM,
                         **BIAS)
    root, z = mesh.find_root(0.0, 0.0, 0.0)
    return mesh, params, [ascan(0.0, 0.0, z, root, photons)], 1


SLAB4_LAYERS = [_reg(1.40, 0.1, 10, 0.9), _reg(1.38, 0.05, 8, 0.85), _reg(1.36, 0.02, 6, 0.9),
                _reg(1.33, 0.01, 5, 0.9)]


def slab4(cells=(64, 64, 32), n_scans=128, photons=1_000_000, bins=512, pitch=0.02 * MM):
    mesh = Mesh.generate(GEOM_LAYERS, cells, (4 * MM, 4 * MM, 2 * MM), p=(0.1 * MM, 0.3 * MM, 0.6 * MM),
                         n_layers=4)
    params = make_params(SLAB4_LAYERS, num_bins=bins, depth=1.5 * MM, **BIAS)
    scans = []
    for i in range(n_scans):
        x = (i - (n_scans - 1) / 2.0) * pitch
        root, z = mesh.find_root(x, 0.0, 0.0)
        scans.append(ascan(x, 0.0, z, root, photons, salt=i))
    return mesh, params, scans, 1


def ellipsoid_budgets(n=512, seed=7, lo=1e5, hi=1e7):
    rng = np.random.default_rng(seed)
    return np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi), n))).astype(np.int64)


def ellipsoid(cells=(64, 64, 32), n_scans=512, pitch=0.0075 * MM, budgets=None, bins=512):
    mesh = Mesh.generate(GEOM_ELLIPSOID2S, cells, (4 * MM, 4 * MM, 2 * MM),
                         p=(0.0, 0.0, 0.8 * MM, 1.2 * MM, 0.8 * MM, 0.5 * MM, 60.0, 0.2 * MM, 0.5 * MM))
    params = make_params([_reg(1.33, 0.02, 6, 0.9), _reg(1.38, 0.1, 12, 0.9), _reg(1.45, 0.5, 20, 0.95)],
                         num_bins=bins, depth=1.5 * MM, **BIAS)
    if budgets is None:
        budgets = ellipsoid_budgets(n_scans)
    scans = []
    for i in range(n_scans):
        x = (i - (n_scans - 1) / 2.0) * pitch
        root, z = mesh.find_root(x, 0.0, 0.0)
        scans.append(ascan(x, 0.0, z, root, int(budgets[i]), salt=i))
    return mesh, params, scans, 1


def slab128(cells=(128, 128, 128), bins=512):
    mesh = Mesh.generate(GEOM_LAYERS, cells, (4 * MM, 4 * MM, 4 * MM), p=(0.1 * MM, 0.3 * MM, 0.6 * MM),
                         n_layers=4)
    params = make_params(SLAB4_LAYERS, num_bins=bins, depth=1.5 * MM, **BIAS)
    return mesh, params


def bump(cells=(128, 128, 64), n_scans=1024, photons=1_000_000, pitch=0.0055 * MM, bins=1024):
    mesh = Mesh.generate(GEOM_BUMP, cells, (6 * MM, 6 * MM, 3 * MM), p=(0.3 * MM, 0.5 * MM, 0.2 * MM, 0.7 * MM))
    params = make_params([_reg(1.40, 0.1, 10, 0.9), _reg(1.37, 0.05, 7, 0.85), _reg(1.33, 0.01, 5, 0.9)],
                         num_bins=bins, depth=2.0 * MM, **BIAS)
    scans = []
    for i in range(n_scans):
        x = (i - (n_scans - 1) / 2.0) * pitch
        root, z = mesh.find_root(x, 0.0, 0.0, project=True)
        scans.append(ascan(x, 0.0, z, root, photons, salt=i))
    return mesh, params, scans, 1


def served_stream(rate, duration_s, seed=11, n_meshes=4, lo=1e4, hi=1e6):
    """Poisson request trace: (arrival_s, mesh index, n_scans, per-A-scan budgets), budgets
    log-uniform in [lo, hi] (the served stream's [1e4, 1e6] by default)."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration_s:
            break
        mesh = int(rng.integers(0, n_meshes))
        n = int(rng.integers(1, 257))
        budgets = np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi), n))).astype(np.int64)
        out.append((t, mesh, n, budgets))
    return out
