"""Pins of oracle functions by what the mathematics fixes (CPU): each check is chosen so that a
plausible mistake in the oracle (a wrong ratio, sign, factor or index) fails it.

* Coarse-to-fine rescale (PostStreamCoarse2Fine, lbm_collision_models.cpp:142-151, 244-251):
  from a uniform non-equilibrium state every ghost node is feq + (tau_f / (2 tau_c)) (f - feq).
* Moving-wall bounce-back (BoundaryBounceBack, lbm_bounce_back.cpp:29-91): plane Couette flow
  reaches the exact linear profile between walls half a spacing outside the boundary nodes.
* Non-equilibrium extrapolation (lbm_noneq_extrapolation.cpp:89-144) away from rest: the boundary
  node's moments are exactly (rho0, rho0 u_b) and its non-equilibrium part is the inner node's.
* Direct-forcing delta kernels (StencilDisOne / StencilDisTwo, immersed_boundary.cpp:365-384): for
  one marker the spread node forces carry the kernels' moments: sum 1, first moment 0, and for
  Peskin's 4-point kernel sum of squares 3/8 per axis at every offset.
* IB lazy reset (immersed_boundary.cpp:76-85, 284-296): a node force left from before the step is
  excluded from the reference velocity of the supported nodes.
"""
import numpy as np
import pytest

from oracle import zlbm_oracle as zo
from oracle.zlbm_oracle import Oracle
from zllbm_amd import grid as g
from tests import scenarios as sc


def _feq(rho, u, cx, cy, cz, w):
    cu = u[:, 0, None] * cx + u[:, 1, None] * cy + (u[:, 2, None] * cz if u.shape[1] > 2 else 0.)
    usq = (u ** 2).sum(1)[:, None]
    return w * rho[:, None] * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * usq)


def _moments(f, cx, cy, cz=None):
    rho = f.sum(1)
    j = [f @ cx, f @ cy] + ([f @ cz] if cz is not None else [])
    return rho, np.stack(j, 1)


@pytest.mark.parametrize("dims", [2, 3])
def test_c2f_rescale_of_uniform_nonequilibrium_state(dims):
    """Every coarse node holds the same populations f = feq(rho, u) + g with a fixed
    non-equilibrium part g; fine fallback nodes hold the rescaled value.  Interpolation weights
    sum to one, so every fine ghost node must be feq + r (f - feq) with r = tau_f / (2 tau_c),
    tau_l = 3 nu / dt_l + 1/2.  The fine-to-coarse ratio 2 tau_c / tau_f, or r without its factor
    1/2, fails by percents."""
    if dims == 2:
        grid, kw = sc.amr2d(levels=2)
    else:
        grid, kw = sc.amr3d()
    cx, cy, cz, w, _ = zo.lattice(dims)
    rho0, u0 = 1.02, np.array([0.03, -0.02, 0.01][:dims])
    rng = np.random.default_rng(5)
    g_neq = rng.normal(0, 2e-3, len(w))
    # keep g free of mass and momentum, so the macros of f are (rho0, u0)
    C = np.stack([np.ones_like(w), cx, cy] + ([cz] if dims == 3 else []))
    g_neq -= C.T @ np.linalg.solve(C @ C.T, C @ g_neq)
    feq = _feq(np.array([rho0]), np.pad(u0, (0, 3 - dims))[None, :], cx, cy, cz, w)[0]
    f = feq + g_neq
    nu = kw["nu0"]
    tau_c, tau_f = 3 * nu + 0.5, 3 * nu * 2 + 0.5
    r = tau_f / (2 * tau_c)
    assert abs(r - 0.5 * tau_f / tau_c) < 1e-15 and abs(r - 1) > 0.05 and abs(r - 2 * tau_c / tau_f) > 0.05
    expect = feq + r * (f - feq)
    o = Oracle(grid, **kw)
    o.set_f(0, np.tile(f, (o.n(0), 1)))
    o.set_f(1, np.tile(expect, (o.n(1), 1)))
    o.debug_transfer(1)
    lev = grid.levels[1]
    f1 = o.f(1)
    for layer in (1, 2):
        idx = np.searchsorted(lev.keys, lev.layers[(g.OUTER_F2C, layer)])
        assert idx.size > 0
        np.testing.assert_allclose(f1[idx], np.tile(expect, (idx.size, 1)), rtol=0, atol=4e-16)
    # and it differs from the f2c-ratio result by far more than rounding
    wrong = feq + (2 * tau_c / tau_f) * (f - feq)
    assert np.max(np.abs(expect - wrong)) > 1e-4


@pytest.mark.parametrize("tau", [0.6, 1.0, 1.7])
def test_couette_moving_wall_linear_profile(tau):
    """Plane Couette flow: periodic in x, bounce-back walls at y min (at rest) and y max (moving at
    U).  The boundary nodes' incoming populations come from their own outgoing ones less
    6 rho0 w (c . u_w), i.e. the walls sit half a spacing outside the boundary nodes, and the
    steady state is u_x(y) = U (y - y_lo + 1/2) / (y_hi - y_lo + 1) at every tau.  The wrong sign
    or factor of the wall term, or a wall on the nodes, fails by far more than 1e-10."""
    ny = 8
    grid = g.uniform_grid(2, (4, ny))
    U = 0.02
    bc = {g.XMIN: (sc.PER, (0., 0.)), g.XMAX: (sc.PER, (0., 0.)), g.YMIN: (sc.BB, (0., 0.)), g.YMAX: (sc.BB, (U, 0.))}
    nu = (tau - 0.5) / 3
    o = Oracle(grid, nu0=nu, bc=bc)
    steps = int(40 * (ny + 1) ** 2 / nu / np.pi ** 2) + 200
    o.time_marching(steps)
    cx, cy, cz, w, _ = zo.lattice(2)
    rho, j = _moments(o.f(0), cx, cy)
    y = grid.coords(0)[:, 1].astype(float)
    ylo, yhi = 1.0, 1.0 + ny
    expect = U * (y - ylo + 0.5) / (yhi - ylo + 1.0)
    np.testing.assert_allclose(j[:, 0] / rho, expect, rtol=0, atol=1e-10 * U)
    np.testing.assert_allclose(j[:, 1], 0.0, atol=1e-12)
    np.testing.assert_allclose(rho, 1.0, atol=1e-12)


@pytest.mark.parametrize("dims", [2, 3])
def test_nee_boundary_moments_away_from_rest(dims):
    """Random non-equilibrium populations everywhere; after a step each non-equilibrium-
    extrapolation node on the inflow face holds exactly mass rho0 and momentum rho0 u_b, and its
    non-equilibrium part equals the inner node's.  Dropping the inner node's feq, or taking the
    boundary feq at the inner density, breaks the moments."""
    cells = (10, 6) if dims == 2 else (6, 5, 4)
    grid = g.uniform_grid(dims, cells)
    ub = (0.04, 0.01, -0.02)[:dims]
    rho0 = 1.0
    bc = {d: (sc.BB, (0., 0., 0.)) for d in range(1, 2 * dims + 1)}
    bc[g.XMIN] = (sc.NEE, ub)
    o = Oracle(grid, nu0=0.03, rho0=rho0, bc=bc)
    cx, cy, cz, w, _ = zo.lattice(dims)
    rng = np.random.default_rng(11)
    n = o.n(0)
    rho = 1 + rng.normal(0, 0.01, n)
    u = rng.normal(0, 0.03, (n, 3))
    if dims == 2:
        u[:, 2] = 0
    f = _feq(rho, u, cx, cy, cz, w) + rng.normal(0, 1e-3, (n, len(w)))
    o.set_f(0, f)
    o.advance_level(0, 1)
    f1 = o.f(0)
    keys = grid.levels[0].keys
    xyz = grid.coords(0).astype(np.int64)
    b = np.nonzero(xyz[:, 0] == grid.dom_min[0])[0]
    # corner nodes take the later faces' conditions (y/z faces run after x min); keep the face
    inner_face = np.all([(xyz[b, a] > grid.dom_min[a]) & (xyz[b, a] < grid.dom_max[a]) for a in range(1, dims)], axis=0)
    b = b[inner_face]
    assert b.size > 0
    rb, jb = _moments(f1[b], cx, cy, cz if dims == 3 else None)
    np.testing.assert_allclose(rb, rho0, rtol=0, atol=2e-15)
    np.testing.assert_allclose(jb, np.tile(np.asarray(ub) * rho0, (b.size, 1)), rtol=0, atol=2e-15)
    # inner neighbours (x + 1)
    nb = np.searchsorted(keys, g.morton_encode(dims, *[(xyz[b, a] + (a == 0)).astype(np.uint64) for a in range(dims)]))
    ri, ji = _moments(f1[nb], cx, cy, cz if dims == 3 else None)
    ui = ji / ri[:, None]
    ui3 = np.pad(ui, ((0, 0), (0, 3 - dims)))
    ub3 = np.tile(np.pad(np.asarray(ub), (0, 3 - dims)), (b.size, 1))
    neq_b = f1[b] - _feq(np.full(b.size, rho0), ub3, cx, cy, cz, w)
    neq_i = f1[nb] - _feq(ri, ui3, cx, cy, cz, w)
    np.testing.assert_allclose(neq_b, neq_i, rtol=0, atol=3e-15)
    assert np.max(np.abs(neq_i)) > 1e-4       # the state is far from equilibrium


def _single_marker_spread(stencil, pos, dims=2):
    cells = (16, 16) if dims == 2 else (12, 12, 12)
    grid = g.uniform_grid(dims, cells)
    U = np.array([0.03, -0.01, 0.02][:dims])
    o = Oracle(grid, nu0=0.05, u0=tuple(U), bc=sc.all_bc(dims, sc.BB), has_force=True)
    xyz = np.asarray([pos[:dims]], dtype=float)
    area = np.array([0.7])
    o.set_ib(0, xyz, area, np.zeros_like(xyz), stencil_dis=stencil, stiffness=2.0)
    o.advance_level(0, 1)
    Fm = o.ib_force(0)[0]
    node = grid.coords(0).astype(float)
    return Fm, area[0], node, o.node_force(0), xyz[0], U


def _peskin(r):
    a = np.abs(r)
    out = np.zeros_like(a)
    m1 = a < 1
    m2 = (a >= 1) & (a < 2)
    out[m1] = (3 - 2 * a[m1] + np.sqrt(1 + 4 * a[m1] - 4 * a[m1] ** 2)) / 8
    out[m2] = (5 - 2 * a[m2] - np.sqrt(-7 + 12 * a[m2] - 4 * a[m2] ** 2)) / 8
    return out


@pytest.mark.parametrize("dims", [2, 3])
@pytest.mark.parametrize("pos", [(7.3, 8.05, 6.6), (7.0, 8.5, 6.999), (7.91, 8.38, 6.27)])
def test_peskin_kernel_moments(dims, pos):
    """One marker in a uniform flow U with stiffness 2 gets F = -2 U; its node forces are
    F area / dt * prod_a phi(x_a - X_a).  Peskin's kernel has, at every offset, sum 1, first
    moment 0 and sum of squares 3/8 per axis, so the node forces must carry
    sum = F area, sum (x - X) = 0 and sum f^2 = (F area)^2 (3/8)^dims, and match the closed form
    node by node."""
    Fm, area, node, nf, X, U = _single_marker_spread(2, pos, dims)
    np.testing.assert_allclose(Fm, -2 * U, rtol=1e-13)
    tot = Fm * area
    np.testing.assert_allclose(nf.sum(0), tot, rtol=1e-12)
    for a in range(dims):
        np.testing.assert_allclose((nf * (node[:, a] - X[a])[:, None]).sum(0), 0.0, atol=1e-15)
    np.testing.assert_allclose((nf ** 2).sum(0), tot ** 2 * (3 / 8) ** dims, rtol=1e-12)
    phi = np.prod([_peskin(node[:, a] - X[a]) for a in range(dims)], axis=0)
    np.testing.assert_allclose(nf, phi[:, None] * tot[None, :], rtol=0, atol=1e-15)


@pytest.mark.parametrize("dims", [2, 3])
def test_hat_kernel_moments(dims):
    """Stencil 1 is the hat kernel phi(r) = 1 - |r|: sum 1, first moment 0, and sum of squares
    (1 - t)^2 + t^2 per axis for the fractional offset t."""
    pos = (7.3, 8.05, 6.6)
    Fm, area, node, nf, X, U = _single_marker_spread(1, pos, dims)
    tot = Fm * area
    np.testing.assert_allclose(nf.sum(0), tot, rtol=1e-12)
    for a in range(dims):
        np.testing.assert_allclose((nf * (node[:, a] - X[a])[:, None]).sum(0), 0.0, atol=1e-15)
    t = X - np.floor(X)
    np.testing.assert_allclose((nf ** 2).sum(0), tot ** 2 * np.prod((1 - t) ** 2 + t ** 2), rtol=1e-12)


def test_ib_lazy_reset_excludes_prior_node_force():
    """Node forces left from an earlier step (here set by hand, random) do not enter the reference
    velocity: in a uniform equilibrium flow U every marker still gets exactly F = -2 U.  Without
    the reset, the supported nodes' velocity would carry dt F_prior / (2 rho) and the forces would
    move by about 1e-3.  Nodes outside every stencil keep their prior force."""
    grid = g.uniform_grid(2, (40, 24))
    U = np.array([0.03, -0.01])
    from zllbm_amd import geometry
    o = Oracle(grid, nu0=0.05, u0=tuple(U), bc=sc.all_bc(2, sc.BB), has_force=True)
    xy, area = geometry.circle_markers((17.3, 12.6), 4.0, 1.0)
    o.set_ib(0, xy, area, np.zeros_like(xy), stencil_dis=2, stiffness=2.0)
    prior = np.random.default_rng(2).normal(0, 1e-3, (o.n(0), 2))
    o.set_node_force(0, prior)
    o.advance_level(0, 1)
    np.testing.assert_allclose(o.ib_force(0), np.tile(-2.0 * U, (len(area), 1)), rtol=1e-13)
    node = grid.coords(0).astype(float)
    far = np.min(np.hypot(node[:, 0, None] - xy[None, :, 0], node[:, 1, None] - xy[None, :, 1]), axis=1) > 4.0
    np.testing.assert_array_equal(o.node_force(0)[far], prior[far])
