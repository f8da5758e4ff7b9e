"""Shared test helpers: the reference's default test grids and densities.

Grid3D_init_spheres (src/bubbles/grid.F90:856-960, Grid1D_init_spheres
:603-640) and the GaussGenerator extent rule
(src/bubbles/gauss_generator.F90:262-281).
"""
from __future__ import annotations

import json
import math
import os

import numpy as np

import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_golden(name: str):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def spheres_axis(centers, radii, step, nlip, grid_type=1) -> O.Axis:
    b0 = min(c - r for c, r in zip(centers, radii))
    b1 = max(c + r for c, r in zip(centers, radii))
    cell_len = step * (nlip - 1)
    ncell = math.ceil((b1 - b0) / cell_len)
    dx = ncell * cell_len - (b1 - b0)
    b0 -= 0.5 * dx
    b1 += 0.5 * dx
    return O.Axis.equidistant(b0, b1, nlip, step, grid_type)


def gauss_generator_axes(q, alpha, centers, step, nlip, thres, grid_type=1):
    q = np.asarray(q, float)
    alpha = np.asarray(alpha, float)
    centers = np.asarray(centers, float).reshape(-1, 3)
    extent = np.sqrt(np.log(np.abs(q) * (alpha / math.pi) ** 1.5 / thres) / alpha)
    return [spheres_axis(centers[:, i], extent, step, nlip, grid_type) for i in range(3)]


def random_gaussians(rng, n, alpha_range=(0.5, 5.0), box=2.0, logalpha=False):
    """Sum of n normalised Gaussians (the configs' test densities)."""
    if logalpha:
        alpha = np.exp(rng.uniform(np.log(alpha_range[0]), np.log(alpha_range[1]), n))
    else:
        alpha = rng.uniform(*alpha_range, n)
    centers = rng.uniform(-box, box, (n, 3))
    return np.ones(n), alpha, centers


def rel_maxdiff(a, b) -> float:
    """max |a-b| / max |b|: the potential error measure used throughout."""
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def slater_bubbles(rng, natom, lmax=4, zset=tuple(range(1, 9)), box=4.0, ncell=33, nlip=7,
                   r_max=20.0, k=0):
    """Config-2 style bubbles: Z uniform in zset, centres U[-box,box]^3, radial
    grid of ncell cells to r_max, f_lm(r) = c_lm r^l exp(-Z r), c_lm ~ U[-1,1]."""
    Z = rng.choice(np.asarray(zset, float), natom)
    centers = rng.uniform(-box, box, (natom, 3))
    steps, bfs = [], []
    for a in range(natom):
        st = O.radial_steps(Z[a], ncell, nlip, r_max)
        r = O.Axis(0.0, st, nlip).points
        c = rng.uniform(-1.0, 1.0, (lmax + 1) ** 2)
        bf = np.zeros((len(r), (lmax + 1) ** 2), order="F")
        for l in range(lmax + 1):
            for m in range(-l, l + 1):
                lm = l * l + l + m
                bf[:, lm] = c[lm] * r ** l * np.exp(-Z[a] * r)
        steps.append(st)
        bfs.append(bf)
    return O.BubbleSet(centers, steps, bfs, lmax, k, nlip), Z
