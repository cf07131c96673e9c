"""Analytic initial fields: the random-Fourier-mode turbulence of the decaying-turbulence
workload (64 solenoidal modes, |k| in [kmin, kmax] in units of 2 pi / L, amplitudes from
E(k) ~ k^(-5/3), seeded uniform phases) and its vorticity for the refinement criterion.

No counterpart in the reference, which starts every node at feq(rho0, u0)
(lbm_solver_interface.cpp:344-367); the field is imposed through the initial-macro upload.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class FourierModes:
    k: np.ndarray          # [m, 3] integer wave vectors
    amp: np.ndarray        # [m]
    dirs: np.ndarray       # [m, 3] unit vectors orthogonal to k
    phase: np.ndarray      # [m]
    box: np.ndarray        # [3] box length per axis (background cells)

    def _args(self, xyz):
        return 2 * np.pi * (np.asarray(xyz, np.float64) @ (self.k / self.box).T) + self.phase

    def velocity(self, xyz):
        """u at positions xyz [n, 3] (background units, origin at the domain minimum)."""
        return np.cos(self._args(xyz)) * self.amp @ self.dirs

    def vorticity(self, xyz):
        """curl u at positions xyz [n, 3]."""
        kd = np.cross(self.k / self.box, self.dirs) * 2 * np.pi
        return -(np.sin(self._args(xyz)) * self.amp) @ kd


def fourier_modes(box, nmodes=64, kmin=1, kmax=8, urms=0.05 / np.sqrt(3.0), seed=20261202):
    """Draw the modes and scale them to rms speed `urms` per component.  The mean square is taken
    on a uniform periodic sample of 4 kmax points per axis, which is exact: products of two modes
    have wavenumbers of at most 2 kmax < 4 kmax, so none aliases to zero."""
    rng = np.random.default_rng(seed)
    ks = []
    while len(ks) < nmodes:
        k = rng.integers(-kmax, kmax + 1, size=3)
        r = np.linalg.norm(k)
        if kmin <= r <= kmax:
            ks.append(k.astype(np.float64))
    ks = np.asarray(ks)
    kn = np.linalg.norm(ks, axis=1)
    amp = kn ** (-11.0 / 6.0)
    dirs = rng.normal(size=(nmodes, 3))
    dirs -= (np.sum(dirs * ks, 1) / kn ** 2)[:, None] * ks        # divergence free
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    phase = rng.uniform(0, 2 * np.pi, nmodes)
    box = np.asarray(box, np.float64)
    modes = FourierModes(ks, amp, dirs, phase, box)
    ns = 4 * kmax
    axes = [np.arange(ns, dtype=np.float64) * (b / ns) for b in box]
    sample = np.stack([a.ravel() for a in np.meshgrid(*axes, indexing="ij")], 1)
    u = modes.velocity(sample)
    modes.amp = amp * urms / np.sqrt(np.mean(np.sum(u ** 2, 1)) / 3.0)
    return modes
