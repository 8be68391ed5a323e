"""Synthetic workloads of the benchmark configurations (host side, seeded):

* batch_trees: independent 3D adaptive trees, each refined around a random
  sphere (centre uniform in [0.2, 0.8]^3, radius uniform in [0.1, 0.3]) to the
  refinement level whose element count is closest, in log scale, to a target
  drawn log-uniformly from [lo, hi]; order p alternating 1 and 2;
* serve_pool: request meshes, dim 3 (a random-sphere tree) or 4 (a
  Gaussian-mixture tree, one point per leaf) in a 70/30 split, p = 1, with
  element counts log-uniform (stratified) in [lo, hi] (the sphere radius and
  the 4D point count are fitted to the target through the measured
  elements-per-size ratios of the generators).

There is no counterpart in the reference (its benchmark builds one tree per
run, bench/src/matvec_bench_adaptive.cpp:199-216); these generators only fix
the inputs the measurements and tests share.
"""
import math

import numpy as np


def sized_sphere_tree(dkt, target, center, radius, dim=3, maxdepth=30, base_level=2, max_level=14):
    """(tree, level): the 2:1 balanced sphere tree whose element count is
    closest to `target` in log scale among refinement levels base+1..max."""
    best = None
    for lev in range(base_level + 1, max_level + 1):
        tree = dkt.balance_tree(dkt.sphere_tree(dim, maxdepth, base_level, lev, center, radius), 1)
        n = len(tree)
        if best is None or abs(math.log(n / target)) < abs(math.log(len(best[0]) / target)):
            best = (tree, lev)
        if n >= target:
            break
    return best


def batch_trees(dkt, ntrees=64, lo=1e3, hi=2e5, seed=3):
    """[(tree, order, target elements, level)] for the batched configuration."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(ntrees):
        c = tuple(float(v) for v in rng.uniform(0.2, 0.8, 3))
        r = float(rng.uniform(0.1, 0.3))
        target = float(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        tree, lev = sized_sphere_tree(dkt, target, c, r)
        out.append((tree, 1 if k % 2 == 0 else 2, target, lev))
    return out


def gauss4d_points(target):
    """Point count of the 4D generator (4-component mixture, sigma 0.05, max
    level 9, one point per leaf, 2:1 balanced) whose tree has about `target`
    elements: measured 31,261 elements at 2,000 points growing as
    points^0.89 (1,285,366 at 130,000)."""
    return max(50, int(round(2000 * (target / 31261.0) ** (1 / 0.89))))


def serve_pool(serve_mod, n, lo=1e3, hi=1e6, seed=0, frac3d=0.7):
    """n MeshSpecs of the serving configuration.  The element targets are
    log-uniform and stratified: target i is drawn from the i-th of n equal
    slices of [log lo, log hi], so every pool spans the whole range."""
    rng = np.random.default_rng(seed)
    pool = []
    span = np.log(hi) - np.log(lo)
    for i in range(n):
        target = float(np.exp(np.log(lo) + span * (i + rng.uniform()) / n))
        if rng.uniform() < frac3d:
            c = tuple(float(v) for v in rng.uniform(0.2, 0.8, 3))
            lev, r = _sphere_fit(target, float(rng.uniform(0.1, 0.3)))
            pool.append((serve_mod.MeshSpec(3, 30, shape="sphere", center=c, radius=r, max_level=lev), target))
        else:
            pool.append((serve_mod.MeshSpec(4, 9, gauss4d_points(target), 1000 + i), target))
    return pool


def _sphere_fit(target, radius):
    """(level, radius) of a sphere tree (base level 2) with about `target`
    elements: measured, the 2:1 balanced tree holds about 58 r^2 4^level
    elements (the surface band of 4 pi r^2 / h^2 cells of size h = 2^-level
    with its transition layers) on top of the 64 base cells.  The level is
    the one whose fitting radius is closest to `radius` in log scale, and the
    radius is then set to fit the target (kept within [0.1, 0.3])."""
    best = None
    for lev in range(3, 14):
        r = math.sqrt(max(target - 64.0, 1.0) / (58.0 * 4.0 ** lev))
        err = abs(math.log(r / radius))
        if best is None or err < best[0]:
            best = (err, lev, min(max(r, 0.1), 0.3))
    return best[1], best[2]
