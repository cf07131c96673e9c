"""Every benchmark configuration on the GPU against the FP64 oracle (reduced where the oracle
would take minutes), plus FP32 against FP64 at full size.

* Configuration 1 (D2Q9 SRT lid-driven cavity 128^2): FP64 against the oracle.
* Configuration 2 (D2Q9 MRT channel, 3 levels, cylinder of 101 direct-forcing markers on the finest
  level, parabolic inflow): a 128 x 64 channel with the same nesting, cylinder and markers, FP64
  against the oracle for 100 background steps (populations of every level and marker forces);
  the full 512 x 256 channel FP32 against FP64.
* Configuration 3 (D3Q19 MRT + Smagorinsky TGV): the same TGV at 32^3 against the oracle.
* Configuration 4 (4-level turbulence with the learned closure): the closure on a 2-level 24^3 box
  against the oracle.
* Configuration 5 (D3Q19 MRT-LES sphere channel, IB, learned closure): a 3-level channel whose
  finest region hugs the sphere, so that marker stencils reach the fine interface layers, FP64
  against the oracle; the full 5-level channel FP32 against FP64.
* D3Q19 direct forcing (DirectForcingScheme3D, immersed_boundary.cpp:232-364) on a uniform
  channel with Fibonacci-sphere markers, stencil 1 and 2.
* Level streams with transfers that share nodes of the level between them.
"""
import os

import numpy as np
import pytest

from oracle.zlbm_oracle import Oracle, lib as oracle_lib
from zllbm_amd import closure, geometry, grid as g, hip, workloads
from tests import scenarios as sc
from tests.test_gpu_parity import _macros, assert_close

pytestmark = pytest.mark.gpu

MODEL = os.path.join(os.path.dirname(closure.__file__), "data", "closure_turbulence.npz")


def _threads():
    oracle_lib().zo_set_threads(min(16, os.cpu_count() or 1))


def gpu_solver(W, precision, model=None, operands=hip.MLP_OPERANDS_FP16):
    kw = dict(W.solver)
    if kw.get("les") == hip.LES_MLP:
        kw["mlp_operands"] = operands
    s = hip.Solver(W.grid, precision=precision, **kw)
    ib = getattr(W, "ib", None)
    if ib is not None:
        s.set_ib(ib.level, ib.xyz, ib.area, np.zeros_like(ib.xyz), stencil_dis=2)
    s.build()
    if kw.get("les") == hip.LES_MLP:
        s.mlp_set(model)
    for lvl, (rho, u) in enumerate(W.init):
        s.init_macro(lvl, rho, u)
    return s


def oracle_solver(W, model=None):
    o = Oracle(W.grid, **W.solver)
    ib = getattr(W, "ib", None)
    if ib is not None:
        o.set_ib(ib.level, ib.xyz, ib.area, np.zeros_like(ib.xyz), stencil_dis=2)
    if W.solver.get("les") == hip.LES_MLP:
        o.set_mlp(model.params.astype(np.float64), model.mean, model.std, model.fmin, model.fmax, model.nu_scale)
    for lvl, (rho, u) in enumerate(W.init):
        o.init_macro(lvl, rho, u)
    return o


def state(sim, W):
    """Density and velocity of the post-stream populations of every level (same moments on both
    sides)."""
    out = [_macros(sim.f(l), W.grid.dims) for l in range(W.grid.max_level + 1)]
    return np.concatenate([r for r, _ in out]), np.concatenate([u for _, u in out])


def errors(a, b):
    (ra, ua), (rb, ub) = a, b
    return float(np.sqrt(np.sum((ua - ub) ** 2) / np.sum(ub ** 2))), float(np.sqrt(np.mean((ra - rb) ** 2)))


def fp32_vs_fp64(W, steps, model=None):
    a = gpu_solver(W, hip.FP32_SHIFTED, model)
    b = gpu_solver(W, hip.FP64, model)
    a.time_marching(steps)
    b.time_marching(steps)
    return errors(state(a, W), state(b, W))


def test_config1_cavity_fp64_vs_oracle():
    """D2Q9 SRT lid-driven cavity 128^2 (Re 100), 2,000 steps: every population to 1e-11."""
    _threads()
    W = workloads.cavity(128)
    s = gpu_solver(W, hip.FP64)
    o = oracle_solver(W)
    s.time_marching(2000)
    o.time_marching(2000)
    assert_close(s.f(0), o.f(0))


def test_config2_reduced_fp64_vs_oracle():
    """3-level D2Q9 MRT channel (128 x 64 background cells, the nesting, cylinder and 101 markers of
    the full configuration on the finest level, parabolic NEE inflow and outflow, bounce-back walls):
    100 background steps, populations of every level and the Lagrangian forces."""
    _threads()
    W = workloads.cylinder_amr(cells=(128, 64))
    assert len(W.ib.area) == 101 and W.ib.level == 2
    s = gpu_solver(W, hip.FP64)
    o = oracle_solver(W)
    s.time_marching(100)
    o.time_marching(100)
    for l in range(3):
        assert_close(s.f(l), o.f(l), rtol=1e-10)
    # the spread sums in fixed per-node order, the oracle in marker order
    np.testing.assert_allclose(s.ib_force(2), o.ib_force(2), rtol=1e-8, atol=1e-12)
    assert np.max(np.abs(o.ib_force(2))) > 1e-4        # the cylinder is loaded


def test_config2_generated_grid_fp64_vs_oracle():
    """The cylinder channel (128 x 64 background cells, cylinder of 16 finest cells) on the grid of
    the geometry-driven generator (tracking markers on level 2, kInAndOut extension with the
    reference example's counts): 100 background steps, populations of every level and the
    Lagrangian forces against the oracle."""
    _threads()
    W = workloads.cylinder_amr_generated(cells=(128, 64), diameter_fine=16)
    assert all(n > 1000 for n in W.nodes)
    s = gpu_solver(W, hip.FP64)
    o = oracle_solver(W)
    s.time_marching(100)
    o.time_marching(100)
    for l in range(3):
        assert_close(s.f(l), o.f(l), rtol=1e-10)
    np.testing.assert_allclose(s.ib_force(2), o.ib_force(2), rtol=1e-8, atol=1e-12)
    assert np.max(np.abs(o.ib_force(2))) > 1e-4


def test_config2_full_fp32_tracks_fp64():
    """The full 512 x 256 3-level cylinder channel: FP32 shifted populations against FP64 in
    lockstep over 500 background steps."""
    W = workloads.cylinder_amr()
    l2_u, rms_rho = fp32_vs_fp64(W, 500)
    print("configuration 2, FP32 vs FP64 after 500 steps: L2(u) %.2e, RMS(rho) %.2e" % (l2_u, rms_rho))
    assert l2_u < 1e-4 and rms_rho < 1e-5


def test_config3_tgv_fp64_vs_oracle():
    """D3Q19 MRT + Smagorinsky (Cs 0.1, physical tau) TGV with noise at 32^3 (u0 0.05, Re 1600),
    100 steps."""
    _threads()
    W = workloads.tgv(32)
    s = gpu_solver(W, hip.FP64)
    o = oracle_solver(W)
    s.time_marching(100)
    o.time_marching(100)
    assert_close(s.f(0), o.f(0))


def test_config4_closure_two_levels_fp64_vs_oracle():
    """The turbulence workload's trained closure on 2 levels over a 24^3 box, 20 steps (the device
    network runs in FP32 inside the FP64 collision, the oracle's in FP64)."""
    _threads()
    model = closure.ClosureModel.load(MODEL)
    W = workloads.decaying_turbulence(24, 2, buffer=2, les=hip.LES_MLP)
    s = gpu_solver(W, hip.FP64, model)
    o = oracle_solver(W, model)
    s.time_marching(20)
    o.time_marching(20)
    l2_u, rms_rho = errors(state(s, W), state(o, W))
    print("configuration 4 (2 levels, 24^3), FP64 GPU vs oracle: L2(u) %.2e, RMS(rho) %.2e" % (l2_u, rms_rho))
    assert l2_u < 1e-8 and rms_rho < 1e-10
    for l in range(2):
        assert_close(s.f(l), o.f(l), rtol=2e-7)


def _sphere_reduced():
    """3-level sphere channel: 32 x 16 x 16 background cells, sphere of 12 finest cells (453
    Fibonacci markers), finest region within 1 background cell of the sphere and no buffer, so
    that marker stencils take fine interface (transfer) nodes."""
    return workloads.sphere_amr(cells=(32, 16, 16), levels=3, diameter_fine=12, band=4.0, buffer=0)


def _supports_on_layers(W):
    """Stencil nodes (4^3 per marker, stencil 2) that lie on interface layers of the marker level."""
    lev = W.grid.levels[W.ib.level]
    lay = np.concatenate([v for v in lev.layers.values()])
    idx = np.floor(W.ib.xyz * (1 << W.ib.level) + 1e-6).astype(np.int64)
    sup = []
    for dz in range(-1, 3):
        for dy in range(-1, 3):
            for dx in range(-1, 3):
                sup.append(g.morton_encode(3, *[(idx[:, a] + d).astype(np.uint64) for a, d in enumerate((dx, dy, dz))]))
    sup = np.unique(np.concatenate(sup))
    return int(np.isin(sup, lay).sum()), sup.size


def test_config5_reduced_fp64_vs_oracle():
    """3-level D3Q19 MRT-LES channel with the immersed sphere (stencil 2) and the trained closure:
    FP64 on the GPU against the oracle for 4 background steps, populations of every level and
    marker forces.  Marker stencils take nodes of the fine interface layers (transfer targets and
    sources that are also forcing supports)."""
    _threads()
    model = closure.ClosureModel.load(MODEL)
    W = _sphere_reduced()
    on_layers, total = _supports_on_layers(W)
    assert on_layers > 500, (on_layers, total)
    s = gpu_solver(W, hip.FP64, model)
    o = oracle_solver(W, model)
    s.time_marching(4)
    o.time_marching(4)
    l2_u, rms_rho = errors(state(s, W), state(o, W))
    print("configuration 5 (3 levels), FP64 GPU vs oracle after 4 steps: L2(u) %.2e, RMS(rho) %.2e" % (l2_u, rms_rho))
    assert l2_u < 1e-8 and rms_rho < 1e-10
    for l in range(3):
        assert_close(s.f(l), o.f(l), rtol=2e-7)
    np.testing.assert_allclose(s.ib_force(2), o.ib_force(2), rtol=1e-6, atol=1e-10)
    assert np.max(np.abs(o.ib_force(2))) > 1e-3


def test_config5_full_fp32_tracks_fp64():
    """The full 5-level sphere channel (7,239 markers, trained closure, FP16-operand network):
    FP32 against FP64 in lockstep over 100 background steps."""
    model = closure.ClosureModel.load(MODEL)
    W = workloads.sphere_amr()
    assert len(W.ib.area) == 7239
    l2_u, rms_rho = fp32_vs_fp64(W, 100, model)
    print("configuration 5, FP32 vs FP64 after 100 steps: L2(u) %.2e, RMS(rho) %.2e" % (l2_u, rms_rho))
    assert l2_u < 1e-4 and rms_rho < 1e-5


def _sphere_channel_uniform():
    grid = g.uniform_grid(3, (32, 20, 20))
    bc = {g.XMIN: (sc.NEE, (0.04, 0., 0.)), g.XMAX: (sc.NEE, (0.04, 0., 0.))}
    for d in range(3, 7):
        bc[d] = (sc.BB, (0., 0., 0.))
    xyz, area = geometry.fibonacci_sphere_markers((11.3, 10.6, 10.4), 4.0, 1.0)
    return grid, dict(nu0=0.02, bc=bc, has_force=True, collision=(hip.MRT,)), xyz, area


@pytest.mark.parametrize("stencil", [1, 2])
def test_ib_sphere_uniform_fp64(stencil):
    """D3Q19 MRT channel with a Fibonacci sphere (r = 4 cells, 202 markers) on one level: FP64
    against the oracle after 30 steps, populations and marker forces."""
    grid, kw, xyz, area = _sphere_channel_uniform()
    assert len(area) == 202
    o = Oracle(grid, **kw)
    s = hip.Solver(grid, precision=hip.FP64, **kw)
    o.set_ib(0, xyz, area, np.zeros_like(xyz), stencil_dis=stencil)
    s.set_ib(0, xyz, area, np.zeros_like(xyz), stencil_dis=stencil)
    s.build().init_feq()
    for _ in range(30):
        o.time_marching(1)
    s.time_marching(30)
    assert_close(s.f(0), o.f(0), rtol=1e-10)
    np.testing.assert_allclose(s.ib_force(0), o.ib_force(0), rtol=1e-8, atol=1e-12)
    assert np.max(np.abs(o.ib_force(0))) > 1e-3


def test_level_streams_with_transfer_overlap():
    """A 3-level grid whose finest region keeps only the minimum 3 cells to level 1's edge, so that
    the transfer into level 2 writes level-1 nodes that the transfer into level 1 reads (the host
    reports the overlap): graph replays and direct launches on level streams must equal one stream
    bit for bit."""
    grid = g.refined_mask_grid(2, (24, 24), 3, lambda k, c: np.ones(len(c), bool), margin=3, buffer=0)
    kw = dict(nu0=0.02, bc=sc.all_bc(2, sc.BB), collision=(hip.MRT,))
    runs = []
    for graphs, streams in ((True, True), (False, True), (False, False)):
        s = hip.Solver(grid, precision=hip.FP32_SHIFTED, **kw).build()
        s.graphs_enable(graphs)
        s.level_streams_enable(streams)
        for l in range(3):
            rho, u = sc.level_fields(grid, l, 0.05, noise=1e-3)
            s.init_macro(l, rho, u)
        runs.append(s)
    assert runs[0].transfer_info(1)["overlaps_finer_transfer"] == 1
    ref = runs[-1]
    for n in (2, 3, 4, 2):
        for s in runs:
            s.time_marching(n)
        for l in range(3):
            f = ref.f(l)
            for s in runs[:-1]:
                np.testing.assert_array_equal(s.f(l), f)


@pytest.mark.parametrize("precision", [hip.FP64, hip.FP32_SHIFTED])
def test_inflow_node_velocities(precision):
    """Per-node inflow velocities (zlh_set_boundary_velocities) on a D2Q9 MRT channel, against the
    oracle after 200 steps: FP64 populations to 1e-11, FP32 shifted velocity L2 below 1e-5."""
    grid, kw = sc.channel_nee(nx=48, ny=24, u_in=0.05)
    kw["collision"] = (hip.MRT,)
    face = grid.coords(0)[:, 0].astype(np.int64) == grid.dom_min[0]
    rng = np.random.default_rng(11)
    ub = np.zeros((int(face.sum()), 3))
    ub[:, 0] = 0.05 * (1 + 0.01 * rng.standard_normal(len(ub)))
    kw["bc_velocities"] = {(0, g.XMIN): ub}
    o = Oracle(grid, **kw)
    s = hip.Solver(grid, precision=precision, **kw).build().init_feq()
    o.time_marching(200)
    s.time_marching(200)
    if precision == hip.FP64:
        assert_close(s.f(0), o.f(0), rtol=1e-11)
    else:
        (ra, ua), (rb, ub_) = _macros(s.f(0), 2), _macros(o.f(0), 2)
        assert np.sqrt(np.sum((ua - ub_) ** 2) / np.sum(ub_ ** 2)) < 1e-5
