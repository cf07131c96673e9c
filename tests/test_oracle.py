"""Pins the FP64 oracle (oracle/zlbm_oracle.c) on the CPU.

The reference ships no tests or golden vectors and its build needs MPI and a C++23 header the
image lacks, so the oracle is checked against: constants restated from the reference's headers
(typed below as golden values with their file:line), closed forms and invariants of the lattice
Boltzmann method, an independent numpy restatement for periodic grids, and a behaviour of the
reference's shipped Smagorinsky mode that follows from its relaxation-time formula.
"""
import numpy as np
import pytest

from oracle import zlbm_oracle as zo
from oracle.zlbm_oracle import Oracle
from zllbm_amd import grid as g
from tests import scenarios as sc


# ------------------------------------------------------------------------------------------
# Morton codes (grid/morton_aux.cpp:238-351, grid/morton_aux.hpp:184-297)
# ------------------------------------------------------------------------------------------
@pytest.mark.parametrize("dims", [2, 3])
def test_morton_roundtrip_and_bit_layout(dims):
    L = zo.lib()
    rng = np.random.default_rng(11)
    bits = 32 if dims == 2 else 21
    xyz = rng.integers(0, 2 ** bits, size=(500, dims)).astype(np.uint64)
    for p in xyz:
        key = L.zo_encode(dims, int(p[0]), int(p[1]), int(p[2]) if dims == 3 else 0)
        # bit b of axis a sits at position dims*b + a
        expect = 0
        for a in range(dims):
            for b in range(bits):
                if (int(p[a]) >> b) & 1:
                    expect |= 1 << (dims * b + a)
        assert key == expect
        out = np.zeros(3, dtype=np.uint32)
        L.zo_decode(dims, key, out)
        assert list(out[:dims]) == [int(v) for v in p]


@pytest.mark.parametrize("dims", [2, 3])
def test_morton_neighbours_brute_force(dims):
    """FindXNeg/Pos etc.: +-1 in one axis lane; Neg returns the key unchanged at index 0."""
    L = zo.lib()
    rng = np.random.default_rng(5)
    bits = 32 if dims == 2 else 21
    pts = rng.integers(0, 2 ** bits - 1, size=(300, dims))
    pts[:20, 0] = 0
    for p in pts:
        c = [int(v) for v in p] + [0] * (3 - dims)
        key = L.zo_encode(dims, *c)
        for axis in range(dims):
            for d in (-1, 1):
                nb = L.zo_find_neighbor(dims, key, axis, d)
                q = list(c)
                if d < 0 and q[axis] == 0:
                    assert nb == key
                    continue
                q[axis] += d
                assert nb == L.zo_encode(dims, *q)


def test_grid_morton_matches_oracle():
    L = zo.lib()
    rng = np.random.default_rng(2)
    for dims in (2, 3):
        xyz = rng.integers(0, 2 ** (32 if dims == 2 else 21), size=(300, dims)).astype(np.uint64)
        keys = g.morton_encode(dims, *[xyz[:, a] for a in range(dims)])
        for p, k in zip(xyz, keys):
            c = [int(v) for v in p] + [0] * (3 - dims)
            assert int(k) == L.zo_encode(dims, *c)
        np.testing.assert_array_equal(g.morton_decode(dims, keys), xyz)


# ------------------------------------------------------------------------------------------
# lattice constants
# ------------------------------------------------------------------------------------------
@pytest.mark.parametrize("dims", [2, 3])
def test_lattice_isotropy(dims):
    cx, cy, cz, w, opp = zo.lattice(dims)
    c = np.stack([cx, cy, cz], 1)[:, :dims]
    assert abs(w.sum() - 1) < 1e-15
    np.testing.assert_allclose(w @ c, 0, atol=1e-16)
    np.testing.assert_allclose(np.einsum("q,qa,qb->ab", w, c, c), np.eye(dims) / 3, atol=1e-15)
    np.testing.assert_allclose(np.einsum("q,qa,qb,qc,qd->abcd", w, c, c, c, c)[0, 0, 1, 1], 1 / 9, atol=1e-15)
    np.testing.assert_array_equal(c[opp], -c)


def test_direction_order_golden():
    # lbm_d3q19.h:73-85 and lbm_d2q9.h:53-61
    cx, cy, cz, _, _ = zo.lattice(3)
    assert list(cx) == [0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0]
    assert list(cy) == [0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1]
    assert list(cz) == [0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1]
    cx, cy, _, w, _ = zo.lattice(2)
    assert list(cx) == [0, 1, 0, -1, 0, 1, -1, -1, 1]
    assert list(cy) == [0, 0, 1, 0, -1, 1, 1, -1, -1]


# ------------------------------------------------------------------------------------------
# MRT matrices (golden rows from lbm_d3q19.h:98-158 and lbm_d2q9.h:70-90)
# ------------------------------------------------------------------------------------------
D3_M_GOLDEN = {
    1: [-30] + [-11] * 6 + [8] * 12,
    2: [12] + [-4] * 6 + [1] * 12,
    4: [0, -4, 4, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0],
    9: [0, 2, 2, -1, -1, -1, -1, 1, 1, 1, 1, 1, 1, 1, 1, -2, -2, -2, -2],
    10: [0, -4, -4, 2, 2, 2, 2, 1, 1, 1, 1, 1, 1, 1, 1, -2, -2, -2, -2],
    12: [0, 0, 0, -2, -2, 2, 2, 1, 1, 1, 1, -1, -1, -1, -1, 0, 0, 0, 0],
    16: [0, 0, 0, 0, 0, 0, 0, 1, -1, 1, -1, -1, 1, -1, 1, 0, 0, 0, 0],
    17: [0, 0, 0, 0, 0, 0, 0, -1, -1, 1, 1, 0, 0, 0, 0, 1, -1, 1, -1],
    18: [0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, -1, -1, -1, -1, 1, 1],
}
D2_M_GOLDEN = [
    [1, 1, 1, 1, 1, 1, 1, 1, 1], [-4, -1, -1, -1, -1, 2, 2, 2, 2], [4, -2, -2, -2, -2, 1, 1, 1, 1],
    [0, 1, 0, -1, 0, 1, -1, -1, 1], [0, -2, 0, 2, 0, 1, -1, -1, 1], [0, 0, 1, 0, -1, 1, 1, -1, -1],
    [0, 0, -2, 0, 2, 1, 1, -1, -1], [0, 1, -1, 1, -1, 0, 0, 0, 0], [0, 0, 0, 0, 0, 1, -1, 1, -1]]


def test_mrt_matrices_golden():
    m3, minv3, s3 = zo.mrt_matrices(3)
    for k, row in D3_M_GOLDEN.items():
        np.testing.assert_array_equal(m3[k], row)
    # inverse rows as rationals in lbm_d3q19.h:120-157 (exact doubles of the same rationals)
    assert list(minv3[0, :3]) == [1. / 19, -5. / 399, 1. / 21]
    assert list(minv3[1, :11]) == [1. / 19, -11. / 2394, -1. / 63, 0.1, -0.1, 0, 0, 0, 0, 1. / 18, -1. / 18]
    assert list(minv3[7, :]) == [1. / 19, 4. / 1197, 1. / 252, 0.1, 0.025, 0.1, 0.025, 0., 0., 1. / 36, 1. / 72,
                                 1. / 12, 1. / 24, 0.25, 0., 0., 0.125, -0.125, 0.]
    np.testing.assert_allclose(m3 @ minv3, np.eye(19), atol=1e-15)
    assert list(s3) == [0., 1.5, 1.4, 0., 1.2, 0., 1.2, 0., 1.2, 0., 1.4, 0., 1.4, 0., 0., 0., 1.98, 1.98, 1.98]
    m2, minv2, s2 = zo.mrt_matrices(2)
    np.testing.assert_array_equal(m2, D2_M_GOLDEN)
    assert list(minv2[5, :]) == [1. / 9., 1. / 18., 1. / 36., 1. / 6., 1. / 12., 1. / 6., 1. / 12., 0., 1. / 4.]
    np.testing.assert_allclose(m2 @ minv2, np.eye(9), atol=1e-15)


# ------------------------------------------------------------------------------------------
# equilibrium, collision invariants
# ------------------------------------------------------------------------------------------
@pytest.mark.parametrize("dims", [2, 3])
def test_feq_moments(dims):
    grid = g.uniform_grid(dims, (4,) * dims)
    o = Oracle(grid, nu0=0.1, bc=sc.all_bc(dims, sc.PER))
    rng = np.random.default_rng(1)
    n = grid.levels[0].n
    rho = 1 + 0.1 * rng.random(n)
    u = 0.1 * (rng.random((n, dims)) - 0.5)
    o.init_macro(0, rho, u)
    f = o.f(0)
    cx, cy, cz, _, _ = zo.lattice(dims)
    c = np.stack([cx, cy, cz], 1)[:, :dims]
    np.testing.assert_allclose(f.sum(1), rho, rtol=1e-14)
    np.testing.assert_allclose(f @ c, rho[:, None] * u, atol=1e-15)
    pi = np.einsum("nq,qa,qb->nab", f, c, c)
    np.testing.assert_allclose(pi, rho[:, None, None] * (np.eye(dims) / 3 + np.einsum("na,nb->nab", u, u)),
                               atol=1e-14)


@pytest.mark.parametrize("coll,les", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_uniform_equilibrium_is_fixed_point(coll, les):
    grid = g.uniform_grid(3, (6, 6, 6))
    o = Oracle(grid, nu0=0.01, u0=(0.03, -0.02, 0.01), bc=sc.all_bc(3, sc.PER), collision=(coll,), les=les)
    f0 = o.f(0).copy()
    o.time_marching(3)
    np.testing.assert_allclose(o.f(0), f0, atol=1e-16)


def test_mrt_with_uniform_rates_is_srt():
    """MRT with every relaxation rate equal to 1/tau reduces to SRT (textbook identity)."""
    grid, kw = sc.periodic_tgv(3, 8, nu=0.02)
    rho, u = sc.tgv_fields(grid, 0.05, noise=1e-3)
    tau = 0.02 * 3 + 0.5
    srt = Oracle(grid, collision=(0,), **kw)
    mrt = Oracle(grid, collision=(1,), **kw)
    mrt.debug_set_s_base(np.full(19, 1 / tau))
    for o in (srt, mrt):
        o.init_macro(0, rho, u)
        o.time_marching(10)
    np.testing.assert_allclose(mrt.f(0), srt.f(0), rtol=0, atol=1e-15)


def test_mrt_conserved_moments():
    """The MRT collision leaves density and momentum unchanged (rows 0,3,5,7 relax at 0)."""
    grid, kw = sc.periodic_tgv(3, 6, nu=0.02)
    o = Oracle(grid, collision=(1,), **kw)
    f = sc_noneq(grid)
    o.set_f(0, f)
    o.time_marching(1)
    fc = o.fc(0)
    cx, cy, cz, _, _ = zo.lattice(3)
    c = np.stack([cx, cy, cz], 1)
    np.testing.assert_allclose(fc.sum(1), f.sum(1), rtol=1e-14)
    np.testing.assert_allclose(fc @ c, f @ c, atol=1e-16)


def sc_noneq(grid, seed=3):
    o = Oracle(grid, nu0=0.02, bc=sc.all_bc(grid.dims, sc.PER))
    rho, u = sc.tgv_fields(grid, 0.05, noise=1e-3)
    o.init_macro(0, rho, u)
    f = o.f(0)
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 1e-4, f.shape)
    return f + p - p.mean(1, keepdims=True)


def test_smagorinsky_zero_cs_is_plain():
    grid, kw = sc.periodic_tgv(3, 8, nu=0.01)
    rho, u = sc.tgv_fields(grid, 0.05, noise=1e-3)
    a = Oracle(grid, collision=(1,), les=1, les_tau_mode=1, cs=0.0, **kw)
    b = Oracle(grid, collision=(1,), **kw)
    for o in (a, b):
        o.init_macro(0, rho, u)
        o.time_marching(5)
    np.testing.assert_array_equal(a.f(0), b.f(0))


def test_smagorinsky_tau_closed_form():
    """Single node: tau_t = (sqrt(tau0^2 + 18 Cs^2 sqrt(2 Pi:Pi)) - tau0) / 2."""
    grid, kw = sc.periodic_tgv(3, 4, nu=0.01)
    f = sc_noneq(grid)
    cs = 0.1
    o = Oracle(grid, collision=(0,), les=1, les_tau_mode=1, cs=cs, **kw)
    o.set_f(0, f)
    o.time_marching(1)
    fc = o.fc(0)
    cx, cy, cz, w, _ = zo.lattice(3)
    c = np.stack([cx, cy, cz], 1)
    rho = f.sum(1)
    u = (f @ c) / rho[:, None]
    cu = u @ c.T
    feq = w * rho[:, None] * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * (u ** 2).sum(1)[:, None])
    pi = np.einsum("nq,qa,qb->nab", f - feq, c, c)
    q2 = np.einsum("nab,nab->n", pi, pi)
    tau0 = 0.01 * 3 + 0.5
    tau = tau0 + 0.5 * (np.sqrt(tau0 ** 2 + 18 * cs ** 2 * np.sqrt(2 * q2)) - tau0)
    np.testing.assert_allclose(fc, f - (f - feq) / tau[:, None], atol=2e-16)


def test_reference_smagorinsky_blows_up():
    """The shipped Smagorinsky mode sets tau_eff = tau_t (lbm_solver_interface.cpp:276-295), which
    is zero at equilibrium: D3Q19 8^3 periodic TGV, nu 0.01, u0 0.05, Cs 0.16, 20 steps -> every
    population non-finite, for SRT and MRT; with LES off every population stays finite."""
    grid, kw = sc.periodic_tgv(3, 8, nu=0.01)
    rho, u = sc.tgv_fields(grid, 0.05)
    for coll in (0, 1):
        o = Oracle(grid, collision=(coll,), les=1, les_tau_mode=0, cs=0.16, **kw)
        o.init_macro(0, rho, u)
        o.time_marching(20)
        assert not np.any(np.isfinite(o.f(0)))
        o = Oracle(grid, collision=(coll,), **kw)
        o.init_macro(0, rho, u)
        o.time_marching(20)
        assert np.all(np.isfinite(o.f(0)))


# ------------------------------------------------------------------------------------------
# streaming and boundaries
# ------------------------------------------------------------------------------------------
def _numpy_periodic_bgk(f, nu, steps, cells, dims):
    """Independent periodic BGK on the unique nodes: roll streaming."""
    cx, cy, cz, w, _ = zo.lattice(dims)
    c = np.stack([cx, cy, cz], 1)[:, :dims].astype(int)
    tau = nu * 3 + 0.5
    for _ in range(steps):
        rho = f.sum(-1)
        u = np.einsum("...q,qa->...a", f, c) / rho[..., None]
        cu = np.einsum("...a,qa->...q", u, c)
        feq = w * rho[..., None] * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * (u ** 2).sum(-1)[..., None])
        fc = f - (f - feq) / tau
        for q in range(len(w)):
            fc[..., q] = np.roll(fc[..., q], shift=tuple(c[q]), axis=tuple(range(dims)))
        f = fc
    return f


@pytest.mark.parametrize("dims", [2, 3])
def test_periodic_stream_matches_roll(dims):
    cells = (10, 6) if dims == 2 else (6, 5, 4)
    grid = g.uniform_grid(dims, cells)
    nu = 0.03
    o = Oracle(grid, nu0=nu, bc=sc.all_bc(dims, sc.PER))
    rho, u = sc.tgv_fields(grid, 0.04, noise=2e-3)
    o.init_macro(0, rho, u)
    xyz = grid.coords(0).astype(int) - 1
    uniq = np.all(xyz < np.asarray(cells), axis=1)
    # field on unique nodes as an array
    f0 = o.f(0)
    arr = np.zeros(tuple(cells) + (f0.shape[1],))
    arr[tuple(xyz[uniq].T)] = f0[uniq]
    o.time_marching(7)
    ref = _numpy_periodic_bgk(arr, nu, 7, cells, dims)
    f = o.f(0)
    np.testing.assert_allclose(f[uniq], ref[tuple(xyz[uniq].T)], rtol=1e-13, atol=1e-15)
    # duplicated planes equal their twins
    wrapped = xyz % np.asarray(cells)
    np.testing.assert_allclose(f, ref[tuple(wrapped.T)], rtol=1e-13, atol=1e-15)


def test_absent_source_keeps_previous_value():
    """Pull streaming writes f[q](x) only if x - c_q exists (lbm_d3q19_solver.cpp:31-133)."""
    grid = g.uniform_grid(2, (8, 8))
    lev = grid.levels[0]
    xyz = grid.coords(0).astype(int)
    hole = (xyz[:, 0] == 4) & (xyz[:, 1] == 5)
    grid.levels[0] = g.Level(lev.keys[~hole], lev.flags[~hole])
    o = Oracle(grid, nu0=0.05, bc=sc.all_bc(2, sc.BB))
    f = sc_noneq2(grid)
    o.set_f(0, f)
    src = o.source_table(0)
    o.time_marching(1)
    f1 = o.f(0)
    miss = src < 0
    np.testing.assert_array_equal(f1[miss & ~_boundary_mask(o)], f[miss & ~_boundary_mask(o)])
    assert miss.sum() > 0


def sc_noneq2(grid):
    rng = np.random.default_rng(4)
    o = Oracle(grid, nu0=0.05, bc=sc.all_bc(grid.dims, sc.BB))
    f = o.f(0)
    return f + rng.normal(0, 1e-4, f.shape)


def _boundary_mask(o):
    n, q = o.n(0), o.q
    m = np.zeros((n, q), dtype=bool)
    for d in range(1, 2 * o.dims + 1):
        m[o.boundary(0, d)] = True
    return m


def test_bounce_back_closed_box_conserves_mass():
    grid = g.uniform_grid(3, (8, 8, 8))
    bc = sc.all_bc(3, sc.BB)
    o = Oracle(grid, nu0=0.02, bc=bc, collision=(1,))
    rho, u = sc.tgv_fields(grid, 0.05, noise=1e-3)
    o.init_macro(0, rho, u)
    m0 = o.conservation(0)[0]
    o.time_marching(40)
    assert abs(o.conservation(0)[0] - m0) < 1e-12 * m0


@pytest.mark.slow
def test_poiseuille_profile():
    """Body-force driven channel between bounce-back walls: the steady SRT profile is the
    parabola u(y) = F/(2 nu) (y - y0)(y1 - y) with walls half a spacing outside the wall nodes,
    exactly so for (tau - 1/2)^2 = 3/16 (the known exactness of half-way bounce-back)."""
    ny = 16
    grid = g.uniform_grid(2, (2, ny))
    bc = {g.XMIN: (sc.PER, ()), g.XMAX: (sc.PER, ()), g.YMIN: (sc.BB, (0, 0)), g.YMAX: (sc.BB, (0, 0))}
    nu, F = np.sqrt(3. / 16.) / 3., 1e-6
    o = Oracle(grid, nu0=nu, bc=bc, has_force=True, const_force=(F, 0.))
    o.time_marching(4000)
    f = o.f(0)
    cx, cy, _, _, _ = zo.lattice(2)
    y = grid.coords(0)[:, 1].astype(float)
    ux = (f @ cx + 0.5 * F) / f.sum(1)
    y0, y1 = 0.5, ny + 1.5
    ref = F / (2 * nu) * (y - y0) * (y1 - y)
    np.testing.assert_allclose(ux, ref, rtol=1e-6, atol=1e-12)


# ------------------------------------------------------------------------------------------
# time stepping order
# ------------------------------------------------------------------------------------------
def _reference_order_loop(max_level):
    """The level-order construction of MultiTimeSteppingC2F (grid/grid_manager.cpp:997-1024)
    restated step by step, independent of the oracle's recursion."""
    if max_level == 0:
        return [0]
    if max_level == 1:
        return [0, 1, 1]
    acc = [0] * (max_level + 1)
    lvl = 1
    acc[0] = 2 ** max_level
    order = [0]
    while lvl > 0:
        if acc[lvl - 1] == acc[lvl]:
            lvl -= 1
        else:
            order.append(lvl)
            acc[lvl] += 2 ** (max_level - lvl)
            if lvl + 1 < max_level:
                lvl += 1
            else:
                lvl += 1
                order += [lvl, lvl]
                acc[lvl] += 2 ** (max_level - lvl + 1)
    return order


def test_time_stepping_order():
    assert list(zo.time_stepping_order(0)) == [0]
    assert list(zo.time_stepping_order(1)) == [0, 1, 1]
    # MultiTimeSteppingC2F (grid_manager.cpp:997-1024): each coarse step, then two of the next level
    assert list(zo.time_stepping_order(2)) == [0, 1, 2, 2, 1, 2, 2]
    for ml in range(2, 7):
        assert list(zo.time_stepping_order(ml)) == _reference_order_loop(ml)


def test_inflow_node_velocities():
    """Per-node face velocities (zo_set_bc_velocities): a channel at rest, after one step every
    interior inlet node carries exactly the velocity given for it (feq(rho0, u_b) + f_in -
    feq(f_in) with the inner node at rest), in key order of the face's nodes."""
    grid, kw = sc.channel_nee(nx=24, ny=16, u_in=0.06)
    xy = grid.coords(0).astype(np.int64)
    lo, hi = grid.dom_min, grid.dom_max
    face = xy[:, 0] == lo[0]
    ub = np.zeros((int(face.sum()), 3))
    rng = np.random.default_rng(7)
    ub[:, 0] = 0.05 * (1 + 0.01 * rng.standard_normal(len(ub)))
    ub[:, 1] = 1e-3 * rng.standard_normal(len(ub))
    o = Oracle(grid, collision=(0,), bc_velocities={(0, g.XMIN): ub}, **kw)
    o.advance_level(0, 1)
    f = o.f(0)
    cx, cy, _, _, _ = zo.lattice(2)
    rho = f.sum(1)
    u = np.stack([f @ np.asarray(cx, np.float64), f @ np.asarray(cy, np.float64)], 1) / rho[:, None]
    inner = (xy[face, 1] > lo[1]) & (xy[face, 1] < hi[1])    # wall corners take bounce-back after
    np.testing.assert_allclose(u[face][inner], ub[inner, :2], rtol=1e-12, atol=1e-15)
    with pytest.raises(RuntimeError, match="boundary velocities"):
        Oracle(grid, collision=(0,), bc_velocities={(0, g.XMIN): ub[:-1]}, **kw)


def test_parabolic_inflow_profile():
    """A channel at rest: after one step the inlet nodes carry feq(rho0, u_b) + f_in - feq(f_in) with
    the inner node still at rest, so their velocity is the profile itself, u_in 4 t (1 - t) across
    the channel (t = 0, 1 on the walls), peaking at u_in on the centre line."""
    u_in = 0.06
    grid, kw = sc.channel_nee(nx=24, ny=16, u_in=u_in)
    o = Oracle(grid, collision=(0,), bc_profile={g.XMIN: 1}, **kw)
    o.advance_level(0, 1)
    f = o.f(0)
    cx, cy, _, _, _ = zo.lattice(2)
    rho = f.sum(1)
    ux = (f @ np.asarray(cx, np.float64)) / rho
    xy = grid.coords(0).astype(np.int64)
    lo, hi = grid.dom_min, grid.dom_max
    inlet = (xy[:, 0] == lo[0]) & (xy[:, 1] > lo[1]) & (xy[:, 1] < hi[1])
    t = (xy[inlet, 1] - lo[1]) / float(hi[1] - lo[1])
    np.testing.assert_allclose(ux[inlet], u_in * 4 * t * (1 - t), rtol=1e-12, atol=1e-15)
    assert abs(ux[inlet].max() - u_in) < 1e-12        # even node count across: a node on the centre line
