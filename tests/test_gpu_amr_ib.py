"""Parity of the HIP coarse/fine transfer and immersed boundary with the FP64 oracle (GPU)."""
import numpy as np
import pytest

from oracle.zlbm_oracle import Oracle
from zllbm_amd import grid as g
from zllbm_amd import hip
from tests import scenarios as sc
from tests.test_gpu_parity import assert_close

pytestmark = pytest.mark.gpu


def _pair(grid, kw, precision=hip.FP64, noise=1e-3):
    o = Oracle(grid, **kw)
    s = hip.Solver(grid, precision=precision, **kw).build()
    for l in range(grid.max_level + 1):
        rho, u = sc.level_fields(grid, l, 0.05, noise=noise)
        o.init_macro(l, rho, u)
        s.init_macro(l, rho, u)
    return o, s


@pytest.mark.parametrize("coll", [hip.SRT, hip.MRT])
def test_amr2d_three_levels_fp64(coll):
    grid, kw = sc.amr2d(levels=3)
    kw["collision"] = (coll,)
    o, s = _pair(grid, kw)
    o.time_marching(12)
    s.time_marching(12)
    for l in range(3):
        assert_close(s.f(l), o.f(l))


def test_amr3d_mrt_les_fp64():
    grid, kw = sc.amr3d()
    kw.update(collision=(hip.MRT,), les=1, cs=0.1)
    o, s = _pair(grid, kw)
    o.time_marching(8)
    s.time_marching(8)
    for l in range(2):
        assert_close(s.f(l), o.f(l))


@pytest.mark.parametrize("dims", [2, 3])
def test_amr_linear_interpolation_fp64(dims):
    """Linear coarse-to-fine interpolation (InterpolationLinear{2D,3D}) on every refined level:
    populations of every level against the oracle after 10 background steps."""
    grid, kw = sc.amr2d(levels=3) if dims == 2 else sc.amr3d()
    kw.update(collision=(hip.MRT,), interp_method=hip.INTERP_LINEAR)
    o, s = _pair(grid, kw)
    assert s.transfer_info(1)["contributors"] < _pair(grid, dict(kw, interp_method=hip.INTERP_LAGRANGIAN))[1] \
        .transfer_info(1)["contributors"]
    o.time_marching(10)
    s.time_marching(10)
    for l in range(grid.max_level + 1):
        assert_close(s.f(l), o.f(l))


def test_amr_substep_by_substep_fp64():
    """Level substeps driven one at a time through the C ABI, including odd substeps."""
    grid, kw = sc.amr2d(levels=2)
    kw["collision"] = (hip.MRT,)
    o, s = _pair(grid, kw)
    for lvl, tsl in [(0, 1), (1, 1), (1, 2), (0, 1), (1, 1)]:
        o.advance_level(lvl, tsl)
        s.step_level(lvl, tsl)
        if lvl > 0 and tsl % 2 == 0:
            s.transfer_levels(lvl)
        for l in range(2):
            assert_close(s.f(l), o.f(l))


def test_split_transfer_equals_whole_and_odd_substep_refused():
    """The coarse-to-fine and fine-to-coarse halves, called in the reference's order, give the
    whole transfer bit for bit; a transfer after an odd substep of the fine level is refused (its
    ghost targets were not deferred)."""
    grid, kw = sc.amr2d(levels=2)
    kw["collision"] = (hip.MRT,)
    _, a = _pair(grid, kw)
    _, b = _pair(grid, kw)
    for s in (a, b):
        s.step_level(0, 1)
        s.step_level(1, 1)
        with pytest.raises(RuntimeError, match="odd substep"):
            s.transfer_levels(1)
        s.step_level(1, 2)
    a.transfer_levels(1)
    b.transfer_from_coarse(1)
    b.transfer_to_coarse(1)
    for s in (a, b):
        s.step_level(0, 1)
    for l in range(2):
        np.testing.assert_array_equal(a.f(l), b.f(l))


def test_amr_fp32_tracks_fp64():
    grid, kw = sc.amr3d()
    kw.update(collision=(hip.MRT,), les=1, cs=0.1)
    o, _ = _pair(grid, kw)
    _, s = _pair(grid, kw, precision=hip.FP32_SHIFTED)
    o.time_marching(10)
    s.time_marching(10)
    for l in range(2):
        d = s.f(l) - o.f(l)
        assert np.sqrt(np.mean(d ** 2)) < 2e-6


@pytest.mark.parametrize("stencil", [1, 2])
def test_ib_cylinder_fp64(stencil):
    grid, kw, (xy, area) = sc.cylinder_channel()
    kw["collision"] = (hip.SRT,)
    o = Oracle(grid, **kw)
    s = hip.Solver(grid, precision=hip.FP64, **kw)
    o.set_ib(0, xy, area, np.zeros_like(xy), stencil_dis=stencil)
    s.set_ib(0, xy, area, np.zeros_like(xy), stencil_dis=stencil)
    s.build().init_feq()
    for _ in range(40):
        o.time_marching(1)
    s.time_marching(40)
    # FP64 atomics sum the spread in a different order than the reference's marker loop
    assert_close(s.f(0), o.f(0), rtol=1e-10)
    np.testing.assert_allclose(s.ib_force(0), o.ib_force(0), rtol=1e-8, atol=1e-12)
    # node forces (GridNodeLbm::force_, as checkpoints write them): the spread on the supports,
    # zero elsewhere
    fs, fo = s.node_forces(0)[:, :2], o.node_force(0)[:, :2]
    np.testing.assert_allclose(fs, fo, rtol=1e-8, atol=1e-13)
    assert np.count_nonzero(np.abs(fo).sum(1)) > 0 and np.count_nonzero(np.abs(fs).sum(1)) == np.count_nonzero(np.abs(fo).sum(1))


@pytest.mark.parametrize("dims,stencil", [(2, 2), (2, 1), (3, 2)])
def test_ib_periodic_domain_fp64(dims, stencil):
    """Immersed boundary in a periodic domain (the overlap walk, immersed_boundary.cpp:250-256,
    341-363): markers whose stencils wrap across periodic planes, at corners and past a maximum
    plane. Populations, marker forces and node forces (duplicated plane nodes included) against the
    oracle after 10 steps."""
    from tests.test_oracle_amr_ib import _periodic_field
    cells = (16, 12) if dims == 2 else (12, 10, 8)
    grid = g.uniform_grid(dims, cells)
    kw = dict(nu0=0.02, bc=sc.all_bc(dims, g.BC_PERIODIC), collision=(hip.MRT,), has_force=True)
    if dims == 2:
        mk = np.array([[1.3, 6.2], [cells[0] + 0.7, 3.4], [8.0, cells[1] + 0.9], [1.2, 1.1],
                       [cells[0] + 1.2, 5.5], [1.0, 7.3]])
    else:
        mk = np.array([[1.3, 5.2, 4.4], [cells[0] + 0.7, 1.4, 3.3], [6.0, cells[1] + 0.9, cells[2] + 0.6],
                       [1.2, 1.1, 1.3], [cells[0] + 1.2, 5.5, 2.5]])
    area = np.linspace(0.3, 0.7, len(mk))
    phys = (grid.coords(0).astype(float) - 1.0) % np.asarray(cells, float)
    init = _periodic_field(phys, cells)
    o = Oracle(grid, **kw)
    s = hip.Solver(grid, precision=hip.FP64, **kw)
    o.set_ib(0, mk, area, np.zeros_like(mk), stencil_dis=stencil)
    s.set_ib(0, mk, area, np.zeros_like(mk), stencil_dis=stencil)
    s.build()
    o.init_macro(0, *init)
    s.init_macro(0, *init)
    o.time_marching(10)
    s.time_marching(10)
    assert_close(s.f(0), o.f(0), rtol=1e-11)
    np.testing.assert_allclose(s.ib_force(0), o.ib_force(0), rtol=1e-9, atol=1e-14)
    fs, fo = s.node_forces(0)[:, :dims], o.node_force(0)
    np.testing.assert_allclose(fs, fo, rtol=1e-9, atol=1e-15)
    assert np.abs(fo).max() > 1e-4


def test_ib_uniform_flow_force():
    grid = g.uniform_grid(2, (40, 24))
    U = (0.03, -0.01)
    from zllbm_amd import geometry
    xy, area = geometry.circle_markers((17.3, 12.6), 4.0, 1.0)
    s = hip.Solver(grid, nu0=0.05, u0=U, bc=sc.all_bc(2, sc.BB), has_force=True)
    s.set_ib(0, xy, area, np.zeros_like(xy))
    s.build().init_feq()
    s.step_level(0, 1)
    np.testing.assert_allclose(s.ib_force(0), np.tile(-2.0 * np.asarray(U), (len(area), 1)), rtol=1e-12)


def _turbulence_pair(levels, precision=hip.FP64, **extra):
    from zllbm_amd import workloads
    W = workloads.decaying_turbulence(24, levels, buffer=2)
    kw = dict(W.solver, **extra)
    o = Oracle(W.grid, **kw)
    s = hip.Solver(W.grid, precision=precision, **kw).build()
    for l, (rho, u) in enumerate(W.init):
        o.init_macro(l, rho, u)
        s.init_macro(l, rho, u)
    return W, o, s


@pytest.mark.parametrize("levels,steps", [(2, 6), (3, 3)])
def test_vorticity_refined_turbulence_fp64(levels, steps):
    """Irregular (vorticity-criterion) interfaces: MRT-LES on 2 and 3 levels against the oracle."""
    W, o, s = _turbulence_pair(levels)
    print("nodes per level", W.nodes, flush=True)
    o.time_marching(steps)
    s.time_marching(steps)
    for l in range(levels):
        assert_close(s.f(l), o.f(l))


def test_vorticity_refined_turbulence_fp32_velocity_l2():
    """FP32 shifted populations on the 3-level irregular grid: velocity L2 against the FP64 oracle."""
    W, o, s = _turbulence_pair(3, precision=hip.FP32_SHIFTED)
    o.time_marching(3)
    s.time_marching(3)
    num = den = 0.0
    for l in range(3):
        f = o.f(l)
        q = f.shape[1]
        from tests.test_gpu_parity import _macros
        _, u_o = _macros(f, 3)
        _, u_s = s.macros(l)
        num += np.sum((u_s - u_o) ** 2)
        den += np.sum(u_o ** 2)
    l2 = np.sqrt(num / den)
    print("FP32 vs FP64 velocity L2 on 3 levels: %.2e" % l2)
    assert l2 < 1e-5


@pytest.mark.parametrize("base,levels", [
    (24, 3),
    pytest.param(24, 5, marks=pytest.mark.skip(
        reason="with the two coarsest levels mapped onto one shared stream this 5-level case crashed the "
               "host inside zlh_time_marching; that mapping was taken out and its cause is not shown, so "
               "the case stays off until it is")),
])
def test_graph_replay_matches_direct_launches(base, levels):
    """Replayed graphs (of two background steps, and of one for an odd count) and direct
    launches, each with the levels on streams of their own,
    against direct launches on one stream, over calls of odd and even length with manual substeps
    and downloads (synchronisations) in between: every replay bit-identical."""
    from zllbm_amd import workloads
    W = workloads.decaying_turbulence(base, levels, buffer=2)
    runs = []
    for graphs, streams in ((True, True), (False, True), (False, False)):
        s = hip.Solver(W.grid, precision=hip.FP32_SHIFTED, **W.solver).build()
        s.graphs_enable(graphs)
        s.level_streams_enable(streams)
        for l, (rho, u) in enumerate(W.init):
            s.init_macro(l, rho, u)
        runs.append(s)
    ref = runs[-1]
    for n in (2, 3, 4, 1, 2, 2):
        for s in runs:
            s.time_marching(n)
        for l in range(levels):
            f = ref.f(l)
            for s in runs[:-1]:
                np.testing.assert_array_equal(s.f(l), f)
        if n == 1:          # a finest-level substep outside the graph flips that level's buffers
            for s in runs:
                s.step_level(levels - 1, 1)


def test_graph_replay_with_immersed_boundary():
    grid, kw, (xy, area) = sc.cylinder_channel()
    kw["collision"] = (hip.MRT,)
    runs = []
    for graphs in (True, False):
        s = hip.Solver(grid, precision=hip.FP64, **kw)
        s.set_ib(0, xy, area, np.zeros_like(xy), stencil_dis=2)
        s.build().init_feq()
        s.graphs_enable(graphs)
        runs.append(s)
    for n in (4, 3, 6):
        for s in runs:
            s.time_marching(n)
        # FP64 atomics in the force spread may sum in another order
        assert_close(runs[0].f(0), runs[1].f(0), rtol=1e-12)


@pytest.mark.parametrize("dims", [2, 3])
def test_time_interpolation_fp64(dims):
    """Time-interpolated ghost refills after odd substeps (zlh_set_time_interpolation): every
    level against the oracle with the same option after 8 background steps, level streams and
    graph replay on; the run must differ from the reference schedule's."""
    grid, kw = sc.amr2d(levels=3) if dims == 2 else sc.amr3d()
    kw.update(collision=(hip.MRT,), time_interp=True)
    o, s = _pair(grid, kw)
    o.time_marching(8)
    s.time_marching(8)
    for l in range(grid.max_level + 1):
        assert_close(s.f(l), o.f(l))
    o2, _ = _pair(grid, dict(kw, time_interp=False))
    o2.time_marching(8)
    assert np.abs(o2.f(1) - o.f(1)).max() > 1e-9


def test_time_interpolation_substeps_and_restart():
    """The time-interpolated refill through the C ABI (transfer_levels after an odd substep)
    equals time marching bit for bit, with level streams and graphs (a) or on one stream (c); the
    fine-to-coarse half stays refused after an odd substep; a solver restarted from the downloaded
    populations (the upload takes the earlier coarse rows from the state it receives) continues
    bit-exactly."""
    grid, kw = sc.amr2d(levels=2)
    kw.update(collision=(hip.MRT,), time_interp=True)
    _, a = _pair(grid, kw)
    _, b = _pair(grid, kw)
    _, c = _pair(grid, kw)
    c.level_streams_enable(False)
    c.graphs_enable(False)
    a.time_marching(3)
    c.time_marching(3)
    for _ in range(3):
        for lvl, tsl in [(0, 1), (1, 1), (1, 2)]:
            b.step_level(lvl, tsl)
            if lvl > 0:
                b.transfer_levels(lvl)
    for l in range(2):
        np.testing.assert_array_equal(c.f(l), a.f(l))
        np.testing.assert_array_equal(b.f(l), a.f(l))
    b.step_level(0, 1)
    b.step_level(1, 1)
    with pytest.raises(RuntimeError, match="odd substep"):
        b.transfer_to_coarse(1)
    # restart: a fresh solver from a's populations, two more steps each
    r = hip.Solver(grid, precision=hip.FP64, **kw).build()
    for l in range(2):
        r.set_f(l, a.f(l))
    a.time_marching(2)
    r.time_marching(2)
    for l in range(2):
        np.testing.assert_array_equal(r.f(l), a.f(l))
