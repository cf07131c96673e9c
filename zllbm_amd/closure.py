"""Learned sub-grid closure: training data from explicitly filtered fine runs of this solver,
training on the device, and the model container.

No counterpart in the reference.  The closure replaces the Smagorinsky relaxation time
(les_models/lbm_les_smagorinsky.cpp:22-94) with tau_eff = tau0 + 3 nu_t, nu_t from a
9 -> 32 -> 32 -> 1 MLP of local non-equilibrium invariants, and falls back to Smagorinsky where
an input leaves the training range.

Training data (a-priori): a fine periodic run is box-filtered by `ratio` onto the coarse grid.
The sub-grid stress tau_ij = <u_i u_j> - <u_i><u_j> and the resolved strain S of the filtered
velocity give the target nu_t = max(0, -tau^d:S / (2 S:S)) in coarse lattice units.  The inputs
are those the collision computes from the coarse non-equilibrium stress, which for a closure with
eddy viscosity nu_t is Pi = -(2/3) rho (tau0 + 3 nu_t) S (first-order Chapman-Enskog).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

NF, NH = 9, 32
NPARAM = NH * NF + NH + NH * NH + NH + NH + 1


@dataclass
class ClosureModel:
    params: np.ndarray
    mean: np.ndarray
    std: np.ndarray
    fmin: np.ndarray
    fmax: np.ndarray
    nu_scale: float

    def save(self, path):
        np.savez(path, params=self.params, mean=self.mean, std=self.std, fmin=self.fmin, fmax=self.fmax,
                 nu_scale=np.float64(self.nu_scale))

    @staticmethod
    def load(path):
        z = np.load(path)
        return ClosureModel(z["params"], z["mean"], z["std"], z["fmin"], z["fmax"], float(z["nu_scale"]))

    def normalise(self, x):
        return ((x - self.mean) / self.std).astype(np.float32)


def init_params(seed=0):
    """He-style initial weights, zero biases except a negative output bias (small nu_t)."""
    rng = np.random.default_rng(seed)
    W1 = rng.normal(0, np.sqrt(2.0 / NF), (NH, NF))
    W2 = rng.normal(0, np.sqrt(2.0 / NH), (NH, NH))
    w3 = rng.normal(0, np.sqrt(1.0 / NH), NH)
    p = np.concatenate([W1.ravel(), np.zeros(NH), W2.ravel(), np.zeros(NH), w3, [-1.0]])
    return p.astype(np.float32)


def features_from_stress(pi6, rho, u, tau0):
    """The 9 raw inputs the collision computes (mlp.hpp closure_features) from the
    non-equilibrium stress pi6 = (xx, yy, zz, xy, xz, yz) [n, 6], density and velocity."""
    P = pi6 / rho[:, None]
    pxx, pyy, pzz, pxy, pxz, pyz = (P[:, k] for k in range(6))
    pp = pxx ** 2 + pyy ** 2 + pzz ** 2 + 2 * (pxy ** 2 + pxz ** 2 + pyz ** 2)
    a = np.sqrt(2 * pp)
    i1 = pxx + pyy + pzz
    i2 = 0.5 * (i1 * i1 - pp)
    i3 = pxx * (pyy * pzz - pyz * pyz) - pxy * (pxy * pzz - pyz * pxz) + pxz * (pxy * pyz - pyy * pxz)
    u2 = (u ** 2).sum(1)
    t0 = np.full_like(a, tau0)
    return np.stack([a, i1, i2, i3, u2, rho - 1.0, t0, a * tau0, a * a], axis=1)


def random_fourier_velocity(grid, level=0, nmodes=64, kmin=1, kmax=8, urms=0.05 / np.sqrt(3.0), seed=20261202):
    """Random-Fourier-mode velocity (zllbm_amd.fields) at the nodes of one level of a periodic
    grid, rms speed per component `urms` over the box."""
    from zllbm_amd.fields import fourier_modes
    modes = fourier_modes(grid.cells, nmodes, kmin, kmax, urms, seed)
    xyz = grid.coords(level).astype(np.float64) / (1 << level) - 1.0
    return modes.velocity(xyz)


def unique_field(grid, values, level=0):
    """Values on the unique nodes of a periodic level as an (n, n, n, ...) array (duplicated
    max planes dropped)."""
    cells = np.asarray(grid.cells) << level
    xyz = grid.coords(level).astype(np.int64) - (1 << level)
    keep = np.all(xyz < cells, axis=1)
    arr = np.zeros(tuple(cells) + values.shape[1:], dtype=values.dtype)
    arr[tuple(xyz[keep].T)] = values[keep]
    return arr


def box_filter(arr, r):
    """Average over r^3 blocks of a periodic (n, n, n, ...) array."""
    n = arr.shape[:3]
    s = arr.reshape(n[0] // r, r, n[1] // r, r, n[2] // r, r, *arr.shape[3:])
    return s.mean(axis=(1, 3, 5))


def strain(u):
    """Symmetric velocity gradient of a periodic (n, n, n, 3) field, unit spacing, central
    differences; returns (xx, yy, zz, xy, xz, yz)."""
    g = np.empty(u.shape[:3] + (3, 3))
    for a in range(3):
        for b in range(3):
            g[..., a, b] = 0.5 * (np.roll(u[..., a], -1, axis=b) - np.roll(u[..., a], 1, axis=b))
    S = 0.5 * (g + np.swapaxes(g, -1, -2))
    return np.stack([S[..., 0, 0], S[..., 1, 1], S[..., 2, 2], S[..., 0, 1], S[..., 0, 2], S[..., 1, 2]], -1)


def filtered_samples(u_fine, ratio, tau0_coarse, rho_fine=None):
    """A-priori samples from one fine velocity snapshot (n, n, n, 3) and, if given, its density
    (n, n, n): raw inputs [m, 9] and target nu_t [m] in coarse lattice units, plus the coarse
    stress for the Smagorinsky comparison."""
    pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
    ub = box_filter(u_fine, ratio)
    S6 = strain(ub)
    tau6 = np.stack([box_filter(u_fine[..., a] * u_fine[..., b], ratio) - ub[..., a] * ub[..., b]
                     for a, b in pairs], -1)
    tr = tau6[..., 0] + tau6[..., 1] + tau6[..., 2]
    for a in range(3):
        tau6[..., a] -= tr / 3.0
    w = np.array([1.0, 1.0, 1.0, 2.0, 2.0, 2.0])          # off-diagonal entries count twice
    SS = np.sum(w * S6 * S6, -1)
    nu_t = np.maximum(0.0, -np.sum(w * tau6 * S6, -1) / np.maximum(2.0 * SS, 1e-30))
    rho = np.ones(SS.shape) if rho_fine is None else box_filter(rho_fine, ratio)
    pi = -(2.0 / 3.0) * rho[..., None] * (tau0_coarse + 3.0 * nu_t)[..., None] * S6
    # the collision evaluates the inputs in FP32; keep them so, so that the training range
    # bounds compare exactly
    x = features_from_stress(pi.reshape(-1, 6), rho.ravel(), ub.reshape(-1, 3), tau0_coarse).astype(np.float32)
    return x, nu_t.ravel(), pi.reshape(-1, 6)


def smagorinsky_nu_t(pi6, tau0, cs):
    """nu_t of the Smagorinsky closure for a given non-equilibrium stress (physical tau mode)."""
    q = np.sqrt(2 * (pi6[:, 0] ** 2 + pi6[:, 1] ** 2 + pi6[:, 2] ** 2 +
                     2 * (pi6[:, 3] ** 2 + pi6[:, 4] ** 2 + pi6[:, 5] ** 2)))
    return 0.5 * (np.sqrt(tau0 ** 2 + 18 * cs * cs * q) - tau0) / 3.0


def fit_model(solver, x_raw, nu_t, steps=3000, batch=4096, lr=2e-3, seed=20261202, margin=0.0, form=None):
    """Normalise, train on the device (solver: any zllbm_amd.hip.Solver context; form: gradient
    kernel, hip.MLP_FP32 by default or hip.MLP_BF16_MFMA) and return the model with the training
    range of the raw inputs."""
    mean = x_raw.mean(0)
    std = x_raw.std(0)
    const = std == 0
    std[const] = 1.0
    nu_scale = float(np.mean(nu_t)) if np.mean(nu_t) > 0 else 1.0
    xn = ((x_raw - mean) / std).astype(np.float32)
    t = (nu_t / nu_scale).astype(np.float32)
    kw = {} if form is None else {"form": form}
    params, hist = solver.mlp_train(init_params(seed), xn, t, steps, min(batch, len(t)), lr, seed, **kw)
    span = x_raw.max(0) - x_raw.min(0)
    # a feature without spread in the data does not bound the range
    fmin = np.where(const, -np.inf, x_raw.min(0) - margin * span).astype(np.float32)
    fmax = np.where(const, np.inf, x_raw.max(0) + margin * span).astype(np.float32)
    model = ClosureModel(params, mean.astype(np.float32), std.astype(np.float32), fmin, fmax, nu_scale)
    return model, hist
