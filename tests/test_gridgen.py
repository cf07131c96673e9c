"""The geometry-driven grid generator (zllbm_amd.gridgen), pinned by what its steps fix: the
tracking level's region as a Minkowski sum, the interface layer picture and flags, the flood
fill's split, and grids the oracle runs (complete interpolation stencils, uniform flow kept)."""
import itertools

import numpy as np
import pytest

from oracle.zlbm_oracle import Oracle
from zllbm_amd import geometry, gridgen, grid as g
from tests import scenarios as sc


def _circle(level=2, kind="in_and_out", **ext):
    xy, _ = geometry.circle_markers((20.3, 16.2), 2.0, 1.0 / (1 << level))
    return gridgen.GeometrySpec(level=level, vertices=xy, extend_type=kind, **ext)


def _sphere(level=2, kind="in_and_out", **ext):
    xyz, _ = geometry.fibonacci_sphere_markers((12.3, 12.2, 12.1), 2.0, 1.0 / (1 << level))
    return gridgen.GeometrySpec(level=level, vertices=xyz, extend_type=kind, **ext)


@pytest.mark.parametrize("dims", [2, 3])
def test_tracking_level_region_is_a_minkowski_sum(dims):
    """kOut geometry away from the domain edges: the tracking level's region (in the next coarser
    indices) is every tracking node's parent plus the box [-n_neg, n_pos] per axis, with
    n = max(extension, 3) and one more on the positive side (SetNumberOfExtendLayerForGrid): the
    neighbourhood layer gives one, the extension the rest."""
    if dims == 2:
        geo = _circle(kind="out", x_neg=[0, 0, 5], x_pos=[0, 0, 7], y_neg=[0, 0, 4], y_pos=[0, 0, 2])
        cells = (48, 32)
        neg, pos = (5, 4), (8, 4)
    else:
        geo = _sphere(kind="out", x_neg=[0, 0, 4], x_pos=[0, 0, 5], y_neg=[0, 0, 3], y_pos=[0, 0, 2],
                      z_neg=[0, 0, 2], z_pos=[0, 0, 4])
        cells = (32, 24, 24)
        neg, pos = (4, 3, 3), (6, 4, 5)
    r = gridgen.generate(dims, cells, 2, [geo])
    parents = {tuple(c >> 1 for c in t) for t in r.tracking[0]}
    box = list(itertools.product(*[range(-neg[a], pos[a] + 1) for a in range(dims)]))
    want = {tuple(p[a] + o[a] for a in range(dims)) for p in parents for o in box}
    assert r.one_lower[2] == want


@pytest.mark.parametrize("dims", [2, 3])
def test_interface_layers_and_flags(dims):
    """The layer picture (grid_manager.h:97-104) and flags (grid_manager.cpp:413-568): three fine
    layers and two coarse ones per refined level; the coarse outer layer sits on the fine inner
    layer and the coarse innermost layer on the fine outermost one; flags F2C0 on the outermost
    fine layer, F2C ghost on the two outer fine layers, C2F0 and C2F ghost on the coarse outer
    layer; every node inside the domain."""
    geo = _circle() if dims == 2 else _sphere()
    cells = (48, 32) if dims == 2 else (32, 24, 24)
    ext = dict(x_pos=[8, 8, 8], x_neg=[4, 4, 4], y_pos=[4, 4, 4], y_neg=[4, 4, 4], inner=[4, 4])
    if dims == 3:
        ext = dict(x_pos=[4, 4, 4], inner=[4, 4])
    for k, v in ext.items():
        setattr(geo, k, v)
    r = gridgen.generate(dims, cells, 2, [geo])
    grid = r.grid
    for l in range(1, 3):
        fine, coarse = grid.levels[l], grid.levels[l - 1]
        f2c = [fine.layers[(g.OUTER_F2C, i)] for i in range(3)]
        c2f = [coarse.layers[(g.OUTER_C2F, i)] for i in range(2)]
        assert all(len(x) > 0 for x in f2c + c2f)
        for a, b in itertools.combinations(f2c, 2):
            assert len(np.intersect1d(a, b)) == 0
        up = lambda keys: g.morton_encode(dims, *[(g.morton_decode(dims, keys)[:, a].astype(np.uint64) * 2)
                                                  for a in range(dims)])
        assert np.isin(up(c2f[1]), f2c[0]).all()
        assert np.isin(up(c2f[0]), f2c[2]).all()
        flag = dict(zip(fine.keys.tolist(), fine.flags.tolist()))
        cflag = dict(zip(coarse.keys.tolist(), coarse.flags.tolist()))
        assert all(flag[k] & g.FLAG_F2C0 and flag[k] & g.FLAG_F2C_GHOST for k in f2c[2].tolist())
        assert all(flag[k] & g.FLAG_F2C_GHOST and not flag[k] & g.FLAG_F2C0 for k in f2c[1].tolist())
        assert all(not flag[k] & (g.FLAG_F2C0 | g.FLAG_F2C_GHOST) for k in f2c[0].tolist())
        assert all(cflag[k] & g.FLAG_C2F0 and cflag[k] & g.FLAG_C2F_GHOST for k in c2f[1].tolist())
        assert all(not cflag[k] & (g.FLAG_C2F0 | g.FLAG_C2F_GHOST) for k in c2f[0].tolist())
    for l, lev in enumerate(grid.levels):
        c = grid.coords(l).astype(np.int64)
        assert (c >= 1 << l).all() and (c <= (np.asarray(cells) + 1) << l).all()


def test_flood_fill_splits_a_ring():
    """FloodFillForInAndOut on a two-node-thick ring around a disc: the ring nodes next to the
    inside come out inside, those next to the outside outside, and none both."""
    gen = gridgen._Gen(2, (40, 40), 0)
    ring = {(x, y) for x in range(1, 41) for y in range(1, 41) if 8.0 <= np.hypot(x - 20, y - 20) < 10.0}
    outside, inside = gen.flood_fill(0, (20.0, 20.0), ring)
    assert inside and outside and not (inside & outside)
    near = lambda n, test: any(test(np.hypot(n[0] + a - 20, n[1] + b - 20)) for a in (-1, 0, 1) for b in (-1, 0, 1))
    # every inside node touches the disc, every outside node the exterior, and every ring node
    # that touches the disc is inside
    assert all(near(n, lambda r: r < 8.0) for n in inside)
    assert all(near(n, lambda r: r >= 10.0) for n in outside)
    assert {n for n in ring if near(n, lambda r: r < 8.0)} == inside


@pytest.mark.parametrize("dims", [2, 3])
def test_generated_grid_runs_in_the_oracle(dims):
    """Every ghost node's interpolation stencil is complete (the oracle's transfer refuses
    otherwise) and a uniform flow stays uniform on every level; a shear flow runs 10 steps."""
    geo = _circle() if dims == 2 else _sphere()
    cells = (48, 32) if dims == 2 else (32, 24, 24)
    ext = dict(x_pos=[8, 8, 8], x_neg=[4, 4, 4], y_pos=[4, 4, 4], y_neg=[4, 4, 4], inner=[4, 4])
    if dims == 3:
        ext = dict(x_pos=[4, 4, 4], inner=[4, 4])
    for k, v in ext.items():
        setattr(geo, k, v)
    grid = gridgen.generate(dims, cells, 2, [geo]).grid
    u0 = (0.03, -0.01, 0.02)[:dims]
    o = Oracle(grid, nu0=0.02, u0=u0, bc=sc.all_bc(dims, g.BC_PERIODIC), collision=(1,))
    f0 = [o.f(l).copy() for l in range(3)]
    o.time_marching(3)
    for l in range(3):
        np.testing.assert_allclose(o.f(l), f0[l], atol=1e-15)
    o2 = Oracle(grid, nu0=0.02, bc=sc.all_bc(dims, g.BC_PERIODIC), collision=(1,))
    for l in range(3):
        o2.init_macro(l, *sc.level_fields(grid, l, 0.05))
    o2.time_marching(10)
    assert all(np.isfinite(o2.f(l)).all() for l in range(3))
