"""Parity of the HIP time step with the FP64 oracle on single-level grids (runs on the GPU)."""
import numpy as np
import pytest

from oracle.zlbm_oracle import Oracle
from zllbm_amd import grid as g
from zllbm_amd import hip
from tests import scenarios as sc

pytestmark = pytest.mark.gpu

# FP64 device arithmetic differs from the oracle only in rounding order (the MRT transform is
# applied through the integer moment basis, FMA contraction): relative 1e-12 after 100 steps.
RTOL64 = 1e-11


def run_pair(grid, steps, precision=hip.FP64, init=None, **kw):
    okw = {k: v for k, v in kw.items()}
    o = Oracle(grid, **okw)
    s = hip.Solver(grid, precision=precision, **kw).build()
    if init is None:
        s.init_feq()
    else:
        rho, u = init
        o.init_macro(0, rho, u)
        s.init_macro(0, rho, u)
    for _ in range(steps):
        o.time_marching(1)
    s.time_marching(steps)
    return o, s


def assert_close(a, b, rtol=RTOL64, atol=1e-14):
    err = np.max(np.abs(a - b))
    scale = np.max(np.abs(b))
    assert np.all(np.isfinite(a)), "non-finite device populations"
    assert err <= atol + rtol * scale, "max abs diff %.3e (scale %.3e)" % (err, scale)


@pytest.mark.parametrize("dims,cells", [(2, (13, 9)), (2, (40, 24)), (3, (7, 5, 9)), (3, (12, 12, 12))])
def test_source_table_bit_exact(dims, cells):
    grid = g.uniform_grid(dims, cells)
    # remove a scattered subset of nodes so that absent neighbours occur inside the box
    rng = np.random.default_rng(7)
    keep = rng.random(grid.levels[0].n) > 0.15
    lev = grid.levels[0]
    grid.levels[0] = g.Level(lev.keys[keep], lev.flags[keep])
    bc = sc.all_bc(dims, sc.BB)
    o = Oracle(grid, nu0=0.1, bc=bc)
    s = hip.Solver(grid, nu0=0.1, bc=bc).build()
    np.testing.assert_array_equal(s.source_table(0), o.source_table(0))


def test_cavity_srt_fp64():
    grid, kw = sc.cavity(n=32)
    o, s = run_pair(grid, 200, collision=(hip.SRT,), **kw)
    assert_close(s.f(0), o.f(0))


def test_cavity_mrt_fp64():
    grid, kw = sc.cavity(n=24)
    o, s = run_pair(grid, 150, collision=(hip.MRT,), **kw)
    assert_close(s.f(0), o.f(0))


@pytest.mark.parametrize("coll", [hip.SRT, hip.MRT])
def test_tgv3d_fp64(coll):
    grid, kw = sc.periodic_tgv(3, 12, nu=0.02)
    init = sc.tgv_fields(grid, 0.05, noise=1e-3)
    o, s = run_pair(grid, 60, init=init, collision=(coll,), **kw)
    assert_close(s.f(0), o.f(0))


@pytest.mark.parametrize("coll", [hip.SRT, hip.MRT])
def test_tgv3d_les_physical_fp64(coll):
    grid, kw = sc.periodic_tgv(3, 12, nu=0.002)
    init = sc.tgv_fields(grid, 0.05, noise=1e-3)
    o, s = run_pair(grid, 60, init=init, collision=(coll,), les=1, les_tau_mode=1, cs=0.1, **kw)
    assert_close(s.f(0), o.f(0))


def _noneq_start(grid, seed=3):
    """feq of a TGV plus a seeded non-equilibrium perturbation, so that every node has a
    non-zero stress at the first collision."""
    from oracle.zlbm_oracle import Oracle as _O
    rho, u = sc.tgv_fields(grid, 0.05, noise=1e-3)
    o = _O(grid, nu0=0.01, bc=sc.all_bc(grid.dims, sc.PER))
    o.init_macro(0, rho, u)
    f = o.f(0)
    rng = np.random.default_rng(seed)
    pert = rng.normal(0.0, 1e-4, size=f.shape)
    pert -= pert.mean(axis=1, keepdims=True)      # keep the density
    return f + pert


@pytest.mark.parametrize("coll", [hip.SRT, hip.MRT])
def test_les_reference_mode_fp64(coll):
    """tau_eff = tau_t (the reference's shipped Smagorinsky) from a non-equilibrium start."""
    grid, kw = sc.periodic_tgv(3, 8, nu=0.01)
    f0 = _noneq_start(grid)
    kw.update(collision=(coll,), les=1, les_tau_mode=0, cs=0.16)
    o = Oracle(grid, **kw)
    s = hip.Solver(grid, precision=hip.FP64, **kw).build()
    o.set_f(0, f0)
    s.set_f(0, f0)
    o.time_marching(1)
    s.time_marching(1)
    assert_close(s.f(0), o.f(0), rtol=1e-10)


def test_les_reference_mode_blows_up_from_equilibrium():
    """From an equilibrium start the shipped Smagorinsky form divides by tau_t = 0: every
    population turns non-finite, as observed on the reference (8^3 TGV, nu 0.01, 20 steps)."""
    grid, kw = sc.periodic_tgv(3, 8, nu=0.01)
    rho, u = sc.tgv_fields(grid, 0.05)
    for coll in (hip.SRT, hip.MRT):
        s = hip.Solver(grid, precision=hip.FP64, collision=(coll,), les=1, les_tau_mode=0, cs=0.16, **kw).build()
        s.init_macro(0, rho, u)
        s.time_marching(20)
        assert not np.any(np.isfinite(s.f(0)))


def test_channel_nee_fp64():
    grid, kw = sc.channel_nee()
    o, s = run_pair(grid, 150, collision=(hip.MRT,), **kw)
    assert_close(s.f(0), o.f(0))


def test_channel_parabolic_inflow_fp64():
    """Parabolic inflow on the non-equilibrium-extrapolation inlet (per-node boundary velocity),
    FP64 against the oracle, and the FP32 form close to it."""
    grid, kw = sc.channel_nee(u_in=0.06)
    prof = {g.XMIN: hip.PROFILE_PARABOLIC}
    o, s = run_pair(grid, 150, collision=(hip.MRT,), bc_profile=prof, **kw)
    assert_close(s.f(0), o.f(0))
    _, s32 = run_pair(grid, 150, precision=hip.FP32_SHIFTED, collision=(hip.MRT,), bc_profile=prof, **kw)
    d = s32.f(0) - o.f(0)
    assert np.sqrt(np.mean(d ** 2)) < 1e-6


def test_const_force_fp64():
    grid = g.uniform_grid(3, (6, 14, 6))
    bc = sc.all_bc(3, sc.PER)
    bc[g.YMIN] = (sc.BB, (0, 0, 0))
    bc[g.YMAX] = (sc.BB, (0, 0, 0))
    for coll in (hip.SRT, hip.MRT):
        o, s = run_pair(grid, 80, collision=(coll,), nu0=0.05, bc=bc, has_force=True,
                        const_force=(1e-5, 0., 0.))
        assert_close(s.f(0), o.f(0))


@pytest.mark.parametrize("dims", [2, 3])
@pytest.mark.parametrize("compressible", [True, False])
@pytest.mark.parametrize("coll", [hip.SRT, hip.MRT])
def test_forced_les_fp64(dims, compressible, coll):
    """Forcing (Guo terms, force-corrected stress) with the Smagorinsky closure, both equilibrium
    forms, on a periodic channel with bounce-back walls."""
    grid = g.uniform_grid(dims, (8, 12, 6)[:dims])
    bc = sc.all_bc(dims, sc.PER)
    bc[g.YMIN] = (sc.BB, (0, 0, 0))
    bc[g.YMAX] = (sc.BB, (0.02, 0, 0))
    o, s = run_pair(grid, 60, init=sc.tgv_fields(grid, 0.03, noise=1e-3), collision=(coll,), nu0=0.01, bc=bc,
                    has_force=True, const_force=(2e-5, -1e-5, 5e-6)[:dims], les=1, cs=0.15,
                    compressible=compressible)
    assert_close(s.f(0), o.f(0))


def _macros(f, dims):
    cx, cy, cz = (np.array(v, dtype=np.float64) for v in _lat(dims))
    rho = f.sum(1)
    u = np.stack([f @ cx, f @ cy] + ([f @ cz] if dims == 3 else []), axis=1) / rho[:, None]
    return rho, u


def _lat(dims):
    from oracle.zlbm_oracle import lattice
    cx, cy, cz, _, _ = lattice(dims)
    return cx, cy, cz


def test_fp32_shifted_tracks_fp64():
    grid, kw = sc.periodic_tgv(3, 16, nu=0.005)
    init = sc.tgv_fields(grid, 0.05, noise=1e-3)
    o, s = run_pair(grid, 100, init=init, precision=hip.FP32_SHIFTED, collision=(hip.MRT,), les=1, cs=0.1, **kw)
    rho_o, u_o = _macros(o.f(0), 3)
    rho_s, u_s = s.macros(0)
    l2_u = np.sqrt(np.sum((u_s - u_o) ** 2) / np.sum(u_o ** 2))
    l2_rho = np.sqrt(np.mean((rho_s - rho_o) ** 2))
    assert l2_u < 1e-4, l2_u
    assert l2_rho < 1e-6, l2_rho


def test_closed_box_mass_conservation_fp64():
    grid = g.uniform_grid(3, (10, 10, 10))
    bc = sc.all_bc(3, sc.BB)
    init = sc.tgv_fields(grid, 0.05, noise=1e-3)
    s = hip.Solver(grid, precision=hip.FP64, collision=(hip.MRT,), nu0=0.01, bc=bc).build()
    s.init_macro(0, *init)
    m0 = s.conservation(0)
    s.time_marching(50)
    m1 = s.conservation(0)
    assert abs(m1[0] - m0[0]) < 1e-11 * m0[0]


@pytest.mark.parametrize("precision", [hip.FP64, hip.FP32_SHIFTED])
def test_full_size_closed_box_conserves_mass(precision):
    """At the 256^3 size of the uniform MRT + Smagorinsky workload (257^3 nodes, ~260k bricks),
    in a bounce-back box: the total mass after 20 steps equals the initial one (FP64 to 1e-12
    relative, measured 1.6e-14; FP32 shifted storage to 1e-10, measured 1.1e-12), a
    size-independent check of the full-size launch shapes and 64-bit indexing."""
    grid = g.uniform_grid(3, (256, 256, 256))
    bc = sc.all_bc(3, sc.BB)
    init = sc.tgv_fields(grid, 0.05, noise=1e-3)
    s = hip.Solver(grid, precision=precision, collision=(hip.MRT,), nu0=0.0005, bc=bc, les=1,
                   les_tau_mode=1, cs=0.1).build()
    s.init_macro(0, *init)
    m0 = s.conservation(0)
    s.time_marching(20)
    m1 = s.conservation(0)
    rel = abs(m1[0] - m0[0]) / m0[0]
    print("256^3 closed box, precision %d: mass drift %.3e relative after 20 steps" % (precision, rel))
    assert rel < (1e-12 if precision == hip.FP64 else 1e-10)
