"""Synthetic workloads of the time step: grids, solver settings and initial fields.

* ``tgv(n)``: D3Q19 MRT + Smagorinsky (Cs 0.1), n^3-cell periodic Taylor-Green vortex
  (u0 0.05, Re 1600) with seeded Gaussian velocity noise 1e-3.
* ``decaying_turbulence(base, levels)``: D3Q19 MRT-LES on a ``levels``-level grid over a periodic
  base^3-cell box, initialised with 64 random Fourier modes (|k| in 1..8, E(k) ~ k^-5/3, rms
  speed 0.05/sqrt(3) per component); each level is refined where |vorticity| at its cell centres
  exceeds the 80th percentile of |vorticity| over the box (percentile from a 64^3 sample).
  base 32, levels 4 is the 256^3-equivalent configuration.
* ``cavity(n)``: D2Q9 SRT lid-driven cavity, n x n cells, Re 100 on the lid speed 0.1, FP64 at rest.
* ``cylinder_amr_generated()``: the same channel and cylinder on a grid from the geometry-driven
  generator (zllbm_amd.gridgen) with the reference example's extension counts.
* ``cylinder_amr()``: D2Q9 MRT channel of 512 x 256 background cells with two nested refinement
  boxes around a cylinder of 32 finest cells diameter (101 direct-forcing markers on the finest
  level), Re 100 on the diameter and the mean inflow speed 0.05; parabolic inflow and outflow
  (peak 0.075, non-equilibrium extrapolation), bounce-back walls, started from the parabolic
  profile everywhere.
* ``sphere_amr()``: D3Q19 MRT-LES channel of 32 x 16 x 16 background cells (512 x 256 x 256 on
  the finest of 5 levels), each level refined within a band around a sphere of 48 finest cells
  diameter (7,239 Fibonacci markers on the finest level), Re 300, started from uniform flow 0.05;
  the inflow face carries 0.05 (1 + 0.01 xi) per node, xi seeded standard normal.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Tuple

import numpy as np

from zllbm_amd import grid as g
from zllbm_amd.fields import FourierModes, fourier_modes

PER, BB, NEE = g.BC_PERIODIC, g.BC_BOUNCE_BACK, g.BC_NEE


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


@dataclass
class IbSet:
    level: int
    xyz: np.ndarray
    area: np.ndarray


def _channel_bc(dims, u_in):
    bc = {g.XMIN: (NEE, (u_in, 0., 0.)), g.XMAX: (NEE, (u_in, 0., 0.))}
    for d in range(3, 2 * dims + 1):
        bc[d] = (BB, (0., 0., 0.))
    return bc


def cavity(n=128, lid=0.1, re=100.0):
    grid = g.uniform_grid(2, (n, n))
    bc = {d: (BB, (0., 0., 0.)) for d in range(1, 5)}
    bc[g.YMAX] = (BB, (lid, 0., 0.))
    nu = lid * (n + 1) / re
    init = [(np.ones(grid.levels[0].n), np.zeros((grid.levels[0].n, 2)))]
    return Workload("D2Q9 SRT lid-driven cavity %dx%d, Re %g" % (n, n, re), grid,
                    dict(nu0=nu, bc=bc, collision=(0,)), init)


def cylinder_amr(cells=(512, 256), diameter_fine=32, u_in=0.05, re=100.0):
    """Background units: node index i of level l sits at i / 2^l, the domain spans [1, cells + 1].
    u_in is the mean inflow speed; the parabolic profile peaks at 1.5 u_in."""
    levels = 3
    u_peak = 1.5 * u_in
    d_bg = diameter_fine / (1 << (levels - 1))
    cx, cy = cells[0] / 4.0 + 1.0, cells[1] / 2.0 + 1.0
    c0 = np.array([int(cx), int(cy)])
    box1 = (tuple(c0 - 24), tuple(c0 + 24))                 # level-0 node indices
    box2 = (tuple(2 * c0 - 20), tuple(2 * c0 + 20))         # level-1 node indices
    grid = g.refined_box_grid(2, cells, [box1, box2])
    nu = u_in * d_bg / re
    xy, area = _geometry().circle_markers((cx + 0.13, cy + 0.07), d_bg / 2.0, 1.0 / (1 << (levels - 1)))
    init = []
    for l in range(levels):
        y = grid.coords(l)[:, 1].astype(np.float64) / (1 << l)
        t = (y - 1.0) / cells[1]
        u = np.zeros((len(y), 2))
        u[:, 0] = u_peak * 4.0 * t * (1.0 - t)
        init.append((np.ones(len(y)), u))
    prof = {g.XMIN: 1, g.XMAX: 1}
    wl = Workload("D2Q9 MRT channel %dx%d, 3 levels, cylinder D=%d finest cells, Re %g, parabolic inflow" % (
        cells + (diameter_fine, re)), grid,
        dict(nu0=nu, bc=_channel_bc(2, u_peak), bc_profile=prof, collision=(1,), has_force=True), init)
    wl.ib = IbSet(levels - 1, xy, area)
    return wl


def cylinder_amr_generated(cells=(512, 256), diameter_fine=32, u_in=0.05, re=100.0, levels=3,
                           extend=((16, 8, 8, 8), 8)):
    """The cylinder channel of cylinder_amr on a grid from the geometry-driven generator
    (zllbm_amd.gridgen): the markers at the finest level are the tracking geometry, kInAndOut
    extension with the reference example's counts (x +16 / -8, y +-8 outside, 8 inside;
    examples/lbm/immersed_boundary/cylinder2d/cylinder2d.cpp:52-58), flood fill from the centre."""
    from zllbm_amd import gridgen
    u_peak = 1.5 * u_in
    d_bg = diameter_fine / (1 << (levels - 1))
    cx, cy = cells[0] / 4.0 + 1.0, cells[1] / 2.0 + 1.0
    xy, area = _geometry().circle_markers((cx + 0.13, cy + 0.07), d_bg / 2.0, 1.0 / (1 << (levels - 1)))
    (xp, xn, yp, yn), inner = extend
    geo = gridgen.GeometrySpec(level=levels - 1, vertices=xy, extend_type="in_and_out", x_pos=[xp] * levels,
                               x_neg=[xn] * levels, y_pos=[yp] * levels, y_neg=[yn] * levels,
                               inner=[inner] * levels, flood_fill_origin=(cx + 0.13, cy + 0.07))
    grid = gridgen.generate(2, cells, levels - 1, [geo]).grid
    nu = u_in * d_bg / re
    init = []
    for l in range(levels):
        y = grid.coords(l)[:, 1].astype(np.float64) / (1 << l)
        t = (y - 1.0) / cells[1]
        u = np.zeros((len(y), 2))
        u[:, 0] = u_peak * 4.0 * t * (1.0 - t)
        init.append((np.ones(len(y)), u))
    prof = {g.XMIN: 1, g.XMAX: 1}
    wl = Workload("D2Q9 MRT channel %dx%d, %d levels from the geometry-driven generator, cylinder D=%d finest cells, "
                  "Re %g, parabolic inflow" % (cells + (levels, diameter_fine, re)), grid,
                  dict(nu0=nu, bc=_channel_bc(2, u_peak), bc_profile=prof, collision=(1,), has_force=True), init)
    wl.ib = IbSet(levels - 1, xy, area)
    return wl


def perturbed_inflow(grid, u_in, perturb, seed):
    """Inflow velocity u_in (1 + perturb xi), xi ~ N(0, 1) seeded, per node of the x-minimum face,
    drawn level by level in key order; transverse components zero (zlh_set_boundary_velocities)."""
    rng = np.random.default_rng(seed)
    out = {}
    for l, lev in enumerate(grid.levels):
        on_face = grid.coords(l)[:, 0].astype(np.int64) == (int(grid.dom_min[0]) << l)
        n = int(np.count_nonzero(on_face))
        if n:
            ub = np.zeros((n, 3))
            ub[:, 0] = u_in * (1.0 + perturb * rng.standard_normal(n))
            out[(l, g.XMIN)] = ub
    return out


def sphere_amr(cells=(32, 16, 16), levels=5, diameter_fine=48, u_in=0.05, re=300.0, perturb=0.01,
               band=8.0, margin=3, les=2, seed=20261202, buffer=2):
    """Level k is refined where a level-k cell centre lies within r + band level-k cells of the
    sphere centre, at least ``margin`` level-k cells inside level k's region (refined_mask_grid;
    the 16-cell cross-section leaves room for 5 nested levels around the sphere with margin 3,
    the reference's minimum extension)."""
    d_bg = diameter_fine / (1 << (levels - 1))
    r = d_bg / 2.0
    ctr = np.array([cells[0] / 4.0, cells[1] / 2.0, cells[2] / 2.0])   # from the domain minimum

    def refine(k, c):
        return np.linalg.norm(c - ctr, axis=1) < r + band / (1 << k)

    grid = g.refined_mask_grid(3, cells, levels, refine, margin=margin, buffer=buffer)
    nu = u_in * d_bg / re
    xyz, area = _geometry().fibonacci_sphere_markers(tuple(ctr + 1.0 + 0.013), r, 1.0 / (1 << (levels - 1)))
    init = [(np.ones(lev.n), np.tile([u_in, 0.0, 0.0], (lev.n, 1))) for lev in grid.levels]
    wl = Workload("D3Q19 MRT-LES channel %dx%dx%d (%d levels), sphere D=%d finest cells, Re %g" % (
        cells + (levels, diameter_fine, re)), grid,
        dict(nu0=nu, bc=_channel_bc(3, u_in), collision=(1,), les=les, les_tau_mode=1, cs=0.1, has_force=True,
             bc_velocities=perturbed_inflow(grid, u_in, perturb, seed)), init)
    wl.ib = IbSet(levels - 1, xyz, area)
    return wl


def _geometry():
    from zllbm_amd import geometry
    return geometry
