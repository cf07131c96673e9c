"""Shared test cases: one description drives both the HIP solver and the FP64 oracle."""
import numpy as np

from zllbm_amd import grid as g

BB, PER, NEE = g.BC_BOUNCE_BACK, g.BC_PERIODIC, g.BC_NEE


def all_bc(dims, scheme, val=(0., 0., 0.)):
    return {d: (scheme, val) for d in range(1, 2 * dims + 1)}


def cavity(n=32, lid=0.1, re=100.):
    """D2Q9 lid-driven cavity: n cells per side, bounce-back walls, moving lid at y max."""
    grid = g.uniform_grid(2, (n, n))
    bc = all_bc(2, BB)
    bc[g.YMAX] = (BB, (lid, 0.))
    nu = lid * (n + 1) / re
    return grid, dict(nu0=nu, bc=bc)


def tgv_fields(grid, u0=0.05, noise=0.0, seed=20261202):
    """Taylor-Green vortex on the unique nodes of a periodic box, copied to the duplicated
    planes, with optional seeded Gaussian velocity noise."""
    dims = grid.dims
    xyz = grid.coords(0).astype(np.float64)
    n = np.asarray(grid.cells, dtype=np.float64)
    k = 2 * np.pi / n
    x = (xyz - 1.0) * k
    if dims == 2:
        ux = u0 * np.sin(x[:, 0]) * np.cos(x[:, 1])
        uy = -u0 * np.cos(x[:, 0]) * np.sin(x[:, 1])
        rho = 1.0 + 3.0 * (u0 ** 2 / 4.0) * (np.cos(2 * x[:, 0]) + np.cos(2 * x[:, 1]))
        u = np.stack([ux, uy], axis=1)
    else:
        ux = u0 * np.sin(x[:, 0]) * np.cos(x[:, 1]) * np.cos(x[:, 2])
        uy = -u0 * np.cos(x[:, 0]) * np.sin(x[:, 1]) * np.cos(x[:, 2])
        uz = np.zeros_like(ux)
        rho = 1.0 + 3.0 * (u0 ** 2 / 16.0) * (np.cos(2 * x[:, 0]) + np.cos(2 * x[:, 1])) * (np.cos(2 * x[:, 2]) + 2.0)
        u = np.stack([ux, uy, uz], axis=1)
    if noise > 0:
        # noise drawn on unique nodes (index mod cells), so duplicated planes stay identical
        idx = (grid.coords(0).astype(np.int64) - 1) % np.asarray(grid.cells)
        lin = idx[:, 0]
        for a in range(1, dims):
            lin = lin * grid.cells[a] + idx[:, a]
        rng = np.random.default_rng(seed)
        tot = int(np.prod(grid.cells))
        field = rng.normal(0.0, noise, size=(tot, dims))
        u = u + field[lin]
    return rho, u


def periodic_tgv(dims=3, n=16, nu=0.01, **kw):
    grid = g.uniform_grid(dims, (n,) * dims)
    return grid, dict(nu0=nu, bc=all_bc(dims, PER), **kw)


def channel_nee(nx=48, ny=24, u_in=0.05, nu=0.02):
    """D2Q9 channel: NEE inflow/outflow (constant velocity), bounce-back walls."""
    grid = g.uniform_grid(2, (nx, ny))
    bc = {g.XMIN: (NEE, (u_in, 0.)), g.XMAX: (NEE, (u_in, 0.)), g.YMIN: (BB, (0., 0.)), g.YMAX: (BB, (0., 0.))}
    return grid, dict(nu0=nu, bc=bc)


def level_fields(grid, level, u0=0.05, noise=0.0, seed=20261202):
    """TGV velocity/density evaluated at the nodes of one level (background units)."""
    dims = grid.dims
    xyz = grid.coords(level).astype(np.float64) / (1 << level)
    k = 2 * np.pi / np.asarray(grid.cells, dtype=np.float64)
    x = (xyz - 1.0) * k
    if dims == 2:
        u = np.stack([u0 * np.sin(x[:, 0]) * np.cos(x[:, 1]), -u0 * np.cos(x[:, 0]) * np.sin(x[:, 1])], 1)
        rho = 1.0 + 3.0 * (u0 ** 2 / 4.0) * (np.cos(2 * x[:, 0]) + np.cos(2 * x[:, 1]))
    else:
        u = np.stack([u0 * np.sin(x[:, 0]) * np.cos(x[:, 1]) * np.cos(x[:, 2]),
                      -u0 * np.cos(x[:, 0]) * np.sin(x[:, 1]) * np.cos(x[:, 2]), np.zeros(len(x))], 1)
        rho = 1.0 + 3.0 * (u0 ** 2 / 16.0) * (np.cos(2 * x[:, 0]) + np.cos(2 * x[:, 1])) * (np.cos(2 * x[:, 2]) + 2.0)
    if noise > 0:
        rng = np.random.default_rng(seed + level)
        u = u + rng.normal(0.0, noise, size=u.shape)
    return rho, u


def amr2d(levels=3):
    """D2Q9 periodic box, 40 x 32 background cells, nested refinement boxes."""
    boxes = [((10, 8), (30, 24)), ((26, 22), (54, 40))][: levels - 1]
    grid = g.refined_box_grid(2, (40, 32), boxes)
    return grid, dict(nu0=0.02, bc=all_bc(2, PER))


def amr3d():
    """D3Q19 periodic box, 16^3 background cells, one refinement box."""
    grid = g.refined_box_grid(3, (16, 16, 16), [((5, 5, 5), (12, 12, 12))])
    return grid, dict(nu0=0.01, bc=all_bc(3, PER))


def cylinder_channel(nx=64, ny=32, radius=5.0, u_in=0.04, nu=0.04):
    """D2Q9 channel with an immersed cylinder: NEE inflow/outflow, bounce-back walls."""
    from zllbm_amd import geometry
    grid = g.uniform_grid(2, (nx, ny))
    bc = {g.XMIN: (NEE, (u_in, 0.)), g.XMAX: (NEE, (u_in, 0.)), g.YMIN: (BB, (0., 0.)), g.YMAX: (BB, (0., 0.))}
    xy, area = geometry.circle_markers((nx / 3.0 + 1.3, ny / 2.0 + 1.1), radius, 1.0)
    return grid, dict(nu0=nu, bc=bc, has_force=True), (xy, area)
