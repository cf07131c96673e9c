"""Learned closure on the CPU: the FP64 reference network and its gradient, the oracle's fused
closure, and the a-priori data pipeline."""
import numpy as np
import pytest

from oracle import mlp_oracle as mo
from oracle.zlbm_oracle import Oracle
from zllbm_amd import closure
from tests import scenarios as sc


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = closure.init_params(1).astype(np.float64) + rng.normal(0, 0.05, mo.NPARAM)
    X = rng.normal(size=(40, mo.NF))
    T = rng.uniform(0.1, 2.0, 40)
    _, g = mo.loss_and_grad(p, X, T)
    idx = rng.choice(mo.NPARAM, 60, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros(mo.NPARAM)
        e[i] = h
        fd = (mo.loss_and_grad(p + e, X, T)[0] - mo.loss_and_grad(p - e, X, T)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(fd), (i, fd, g[i])


def test_feature_formulas_agree():
    """The product's input formula (from a stress) and the oracle's (from populations)."""
    grid, kw = sc.periodic_tgv(3, 6, nu=0.02)
    o = Oracle(grid, **kw)
    rho, u = sc.tgv_fields(grid, 0.05, noise=1e-3)
    o.init_macro(0, rho, u)
    rng = np.random.default_rng(2)
    f = o.f(0) + rng.normal(0, 1e-4, (o.n(0), 19))
    x_o = mo.features_from_populations(f, 0.56)
    from oracle.zlbm_oracle import lattice
    cx, cy, cz, w, _ = lattice(3)
    c = np.stack([cx, cy, cz], 1)
    r = f.sum(1)
    uu = (f @ c) / r[:, None]
    cu = uu @ c.T
    feq = w * r[:, None] * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * (uu ** 2).sum(1)[:, None])
    pi = np.einsum("nq,qa,qb->nab", f - feq, c, c)
    pi6 = np.stack([pi[:, 0, 0], pi[:, 1, 1], pi[:, 2, 2], pi[:, 0, 1], pi[:, 0, 2], pi[:, 1, 2]], 1)
    np.testing.assert_allclose(closure.features_from_stress(pi6, r, uu, 0.56), x_o, rtol=1e-10, atol=1e-18)


def _constant_model(nu_t, tau0):
    p = np.zeros(mo.NPARAM)
    z = np.log(np.expm1(1.0))           # softplus(z) = 1
    p[-1] = z
    return dict(params=p, mean=np.zeros(9), std=np.ones(9), fmin=np.full(9, -1e30), fmax=np.full(9, 1e30),
                nu_scale=nu_t)


def test_oracle_closure_constant_network_is_srt():
    """A network with constant output nu_t gives tau_eff = tau0 + 3 nu_t."""
    grid, kw = sc.periodic_tgv(3, 6, nu=0.02)
    tau0 = 3 * 0.02 + 0.5
    nu_t = 0.013
    f0 = None
    a = Oracle(grid, collision=(0,), les=2, **kw)
    a.set_mlp(**_constant_model(nu_t, tau0))
    b = Oracle(grid, nu0=0.02 + nu_t, bc=kw["bc"], collision=(0,))
    rho, u = sc.tgv_fields(grid, 0.05, noise=1e-3)
    for o in (a, b):
        o.init_macro(0, rho, u)
        o.time_marching(5)
    np.testing.assert_allclose(a.f(0), b.f(0), rtol=1e-14, atol=1e-17)


def test_oracle_closure_fallback_is_smagorinsky():
    grid, kw = sc.periodic_tgv(3, 6, nu=0.01)
    m = _constant_model(0.5, 0.53)
    m["fmin"] = np.full(9, 1e29)        # every input outside the range
    a = Oracle(grid, collision=(1,), les=2, cs=0.1, **kw)
    a.set_mlp(**m)
    b = Oracle(grid, collision=(1,), les=1, les_tau_mode=1, cs=0.1, **kw)
    rho, u = sc.tgv_fields(grid, 0.05, noise=1e-3)
    for o in (a, b):
        o.init_macro(0, rho, u)
        o.time_marching(5)
    np.testing.assert_array_equal(a.f(0), b.f(0))


def test_filtered_samples_on_synthetic_field():
    """Box filtering of a random solenoidal field: non-negative targets, finite inputs, and the
    filtered strain term dominates where the sub-grid stress aligns with it."""
    from zllbm_amd import grid as g
    grid = g.uniform_grid(3, (32, 32, 32))
    u = closure.random_fourier_velocity(grid, nmodes=48, kmax=12, seed=3)
    uf = closure.unique_field(grid, u)
    x, nu, pi = closure.filtered_samples(uf, 4, 0.52)
    assert x.shape == (8 ** 3, 9) and np.all(np.isfinite(x))
    assert np.all(nu >= 0) and np.mean(nu) > 0
    # spectral divergence vanishes (solenoidal modes) and the rms speed is the requested one
    k = np.fft.fftfreq(32) * 32
    kx, ky, kz = np.meshgrid(k, k, k, indexing="ij")
    uh = [np.fft.fftn(uf[..., a]) for a in range(3)]
    div = kx * uh[0] + ky * uh[1] + kz * uh[2]
    assert np.max(np.abs(div)) < 1e-9 * np.max(np.abs(uh[0]))
    np.testing.assert_allclose(np.sqrt(np.mean(uf ** 2)), 0.05 / np.sqrt(3.0), rtol=1e-12)


def test_bf16_round_known_values():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.0e-5, 65504.0], np.float32)
    r = mo.bf16_round(v)
    # 1 + 2^-8 is a tie and rounds to even (1.0); 1 + 3 * 2^-8 rounds up to 1 + 2^-6
    np.testing.assert_array_equal(r[:4], [1.0, 1.0, 1.0 + 2 ** -6, -2.5])
    assert abs(r[4] - 3.0e-5) <= 2 ** -8 * 3.0e-5
    assert r[5] == 65536.0
    # a BF16 value has at most 8 significant bits
    m, _ = np.frexp(r)
    assert np.all((m * 2 ** 8) == np.round(m * 2 ** 8))


def test_forward_bf16_close_to_exact():
    rng = np.random.default_rng(3)
    p = rng.normal(0, 0.3, mo.NPARAM)
    X = rng.normal(size=(200, 9))
    y, _ = mo.forward(p, X)
    yb = mo.forward_bf16(p, X)
    rel = np.abs(yb - y) / np.abs(y)
    assert 0 < rel.max() < 3e-2 and np.median(rel) < 5e-3


def test_lowp_gradient_with_exact_rounding_is_the_gradient():
    """The emulation of the matrix-core training kernel reduces to the finite-difference-pinned
    gradient when its rounding is the identity, and stays within BF16 accuracy of it."""
    rng = np.random.default_rng(4)
    p = closure.init_params(2).astype(np.float64) + rng.normal(0, 0.05, mo.NPARAM)
    X = rng.normal(size=(300, mo.NF))
    T = rng.uniform(0.1, 2.0, 300)
    loss, g = mo.loss_and_grad(p, X, T)
    l_id, g_id = mo.loss_and_grad_lowp(p, X, T, lambda a: np.asarray(a, np.float64))
    assert abs(l_id - loss) <= 1e-13 * loss
    np.testing.assert_allclose(g_id, g, rtol=1e-11, atol=1e-14 * np.abs(g).max())
    l_b, g_b = mo.loss_and_grad_lowp(p, X, T, mo.bf16_round)
    rel = np.linalg.norm(g_b - g) / np.linalg.norm(g)
    assert 1e-4 < rel < 2e-2, rel
    assert abs(l_b - loss) < 1e-2 * loss
