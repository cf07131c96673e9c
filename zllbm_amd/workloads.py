"""Synthetic workloads of the time step: grids, solver settings and initial fields.

* ``tgv(n)``: D3Q19 MRT + Smagorinsky (Cs 0.1), n^3-cell periodic Taylor-Green vortex
  (u0 0.05, Re 1600) with seeded Gaussian velocity noise 1e-3.
* ``decaying_turbulence(base, levels)``: D3Q19 MRT-LES on a ``levels``-level grid over a periodic
  base^3-cell box, initialised with 64 random Fourier modes (|k| in 1..8, E(k) ~ k^-5/3, rms
  speed 0.05/sqrt(3) per component); each level is refined where |vorticity| at its cell centres
  exceeds the 80th percentile of |vorticity| over the box (percentile from a 64^3 sample).
  base 32, levels 4 is the 256^3-equivalent configuration.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Tuple

import numpy as np

from zllbm_amd import grid as g
from zllbm_amd.fields import FourierModes, fourier_modes

PER = g.BC_PERIODIC


def periodic_bc(dims):
    return {d: (PER, (0., 0., 0.)) for d in range(1, 2 * dims + 1)}


@dataclass
class Workload:
    name: str
    grid: g.Grid
    solver: Dict                       # keyword arguments of zllbm_amd.hip.Solver (minus precision)
    init: List[Tuple[np.ndarray, np.ndarray]] = field(default_factory=list)   # (rho, u) per level
    modes: FourierModes = None

    @property
    def nodes(self):
        return [lev.n for lev in self.grid.levels]

    @property
    def updates_per_step(self):
        """Node updates per background step (level l takes 2^l substeps)."""
        return sum(n << l for l, n in enumerate(self.nodes))


def tgv(n=256, u0=0.05, re=1600.0, noise=1e-3, seed=20261202):
    grid = g.uniform_grid(3, (n, n, n))
    nu = u0 * (n / (2 * np.pi)) / re
    xyz = grid.coords(0).astype(np.float64)
    x = (xyz - 1.0) * (2 * np.pi / n)
    ux = u0 * np.sin(x[:, 0]) * np.cos(x[:, 1]) * np.cos(x[:, 2])
    uy = -u0 * np.cos(x[:, 0]) * np.sin(x[:, 1]) * np.cos(x[:, 2])
    u = np.stack([ux, uy, np.zeros_like(ux)], 1)
    rho = 1.0 + 3.0 * (u0 ** 2 / 16.0) * (np.cos(2 * x[:, 0]) + np.cos(2 * x[:, 1])) * (np.cos(2 * x[:, 2]) + 2.0)
    if noise:
        u = u + np.random.default_rng(seed).normal(0.0, noise, u.shape)
    # duplicated periodic planes carry the values of their twins
    wrap = (xyz - 1) % n
    key_twin = g.morton_encode(3, *[(wrap[:, a] + 1).astype(np.uint64) for a in range(3)])
    idx = np.searchsorted(grid.levels[0].keys, key_twin)
    rho, u = rho[idx], u[idx]
    solver = dict(nu0=nu, bc=periodic_bc(3), collision=(1,), les=1, les_tau_mode=1, cs=0.1)
    return Workload("TGV %d^3 periodic, u0 %g, Re %g, noise %g" % (n, u0, re, noise), grid, solver, [(rho, u)])


def vorticity_threshold(modes: FourierModes, percentile=80.0, sample=64):
    axes = [(np.arange(sample) + 0.5) * (b / sample) for b in modes.box]
    pts = np.stack([a.ravel() for a in np.meshgrid(*axes, indexing="ij")], 1)
    return float(np.percentile(np.linalg.norm(modes.vorticity(pts), axis=1), percentile))


def decaying_turbulence(base=32, levels=4, nu=None, seed=20261202, percentile=80.0, les=1, cs=0.1, buffer=6):
    """The multi-level decaying-turbulence workload (module docstring).  nu is the background-level
    viscosity; the default gives Re_lambda-scale turbulence on the finest level (nu_finest 1e-4)."""
    cells = (base, base, base)
    modes = fourier_modes(cells, seed=seed)
    thr = vorticity_threshold(modes, percentile)
    grid = g.refined_mask_grid(3, cells, levels,
                               lambda k, c: np.linalg.norm(modes.vorticity(c), axis=1) > thr, buffer=buffer)
    if nu is None:
        nu = 1e-4 * (1 << (levels - 1))
    init = []
    for l in range(levels):
        xyz = grid.coords(l).astype(np.float64) / (1 << l) - 1.0
        init.append((np.ones(len(xyz)), modes.velocity(xyz)))
    solver = dict(nu0=nu, bc=periodic_bc(3), collision=(1,), les=les, les_tau_mode=1, cs=cs)
    name = "decaying turbulence, %d levels over a %d^3 periodic box (%d^3-equivalent), 64 Fourier modes" % (
        levels, base, base << (levels - 1))
    return Workload(name, grid, solver, init, modes)
