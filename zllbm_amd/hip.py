"""ctypes binding of the HIP library (C ABI in include/zllbm_hip.h).

The library is the only compute path: :func:`lib` raises when ``libzllbm_hip.so`` is missing
and nothing here computes populations on the CPU.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Dict, Optional, Sequence

import numpy as np

from . import grid as _grid

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libzllbm_hip.so")
_lib = None

FP64, FP32_SHIFTED, BF16_SHIFTED = 0, 1, 2
SRT, MRT = 0, 1
LES_NONE, LES_SMAGORINSKY, LES_MLP = 0, 1, 2
LES_TAU_REFERENCE, LES_TAU_PHYSICAL = 0, 1
PROFILE_UNIFORM, PROFILE_PARABOLIC = 0, 1
MLP_FP32, MLP_BF16_MFMA, MLP_FP16_MFMA = 0, 1, 2
MLP_OPERANDS_BF16, MLP_OPERANDS_FP16 = 0, 1
MLP_NOALLOC_MAX_SAMPLES = 65536   # ZLH_MLP_NOALLOC_MAX_SAMPLES
INTERP_LAGRANGIAN, INTERP_LINEAR = 0, 1
TRANSFORM_VALU, TRANSFORM_MFMA = 0, 1


class SolverDesc(C.Structure):
    _fields_ = [
        ("dims", C.c_int32), ("max_level", C.c_int32), ("precision", C.c_int32),
        ("compressible", C.c_int32), ("nu0", C.c_double), ("rho0", C.c_double),
        ("u0", C.c_double * 3), ("les", C.c_int32), ("les_tau_mode", C.c_int32),
        ("cs", C.c_double), ("has_force", C.c_int32), ("mlp_operands", C.c_int32),
        ("const_force", C.c_double * 3), ("collision", C.c_int32 * 16),
        ("bc_scheme", C.c_int32 * 7), ("pad1", C.c_int32), ("bc_value", (C.c_double * 3) * 7),
        ("dom_min", C.c_uint32 * 3), ("dom_max", C.c_uint32 * 3),
    ]


EXPORTED = (
    "zlh_create", "zlh_destroy", "zlh_last_error", "zlh_set_level_nodes", "zlh_set_interface_layer",
    "zlh_ib_set_markers", "zlh_build", "zlh_init_feq", "zlh_init_macros", "zlh_upload_populations",
    "zlh_step_level", "zlh_transfer_levels", "zlh_time_marching", "zlh_download_populations", "zlh_download_macros",
    "zlh_download_source_table", "zlh_ib_download_forces", "zlh_conservation", "zlh_level_info", "zlh_transfer_info", "zlh_level_streams_enable", "zlh_set_boundary_profile",
    "zlh_synchronize", "zlh_timing_enable", "zlh_timing_read", "zlh_mlp_set", "zlh_mlp_gradient",
    "zlh_mlp_gradient_noalloc", "zlh_set_boundary_velocities", "zlh_set_interpolation_method",
    "zlh_download_node_forces", "zlh_transfer_from_coarse", "zlh_transfer_to_coarse",
    "zlh_mlp_train", "zlh_mlp_training_form", "zlh_mlp_eval", "zlh_closure_features", "zlh_graphs_enable",
    "zlh_max_level_bricks", "zlh_set_time_interpolation", "zlh_set_moment_transform",
)


def lib():
    """Load the HIP library; raises when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libzllbm_hip.so not built: run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        D = np.ctypeslib.ndpointer
        L.zlh_create.argtypes = [C.POINTER(SolverDesc), C.POINTER(P)]
        L.zlh_destroy.argtypes = [P]
        L.zlh_destroy.restype = None
        L.zlh_last_error.argtypes = [P]
        L.zlh_last_error.restype = C.c_char_p
        L.zlh_set_level_nodes.argtypes = [P, C.c_int, D(np.uint64), D(np.int32), C.c_int64]
        L.zlh_set_interface_layer.argtypes = [P, C.c_int, C.c_int, C.c_int, D(np.uint64), C.c_int64]
        L.zlh_ib_set_markers.argtypes = [P, C.c_int, D(np.float64), D(np.float64), D(np.float64), C.c_int64,
                                         C.c_int, C.c_double]
        L.zlh_build.argtypes = [P]
        L.zlh_init_feq.argtypes = [P]
        L.zlh_init_macros.argtypes = [P, C.c_int, D(np.float64), D(np.float64)]
        L.zlh_upload_populations.argtypes = [P, C.c_int, D(np.float64)]
        L.zlh_step_level.argtypes = [P, C.c_int, C.c_int]
        L.zlh_transfer_levels.argtypes = [P, C.c_int]
        L.zlh_transfer_from_coarse.argtypes = [P, C.c_int]
        L.zlh_transfer_to_coarse.argtypes = [P, C.c_int]
        L.zlh_time_marching.argtypes = [P, C.c_int]
        L.zlh_download_populations.argtypes = [P, C.c_int, D(np.float64)]
        L.zlh_download_macros.argtypes = [P, C.c_int, D(np.float64), D(np.float64)]
        L.zlh_download_source_table.argtypes = [P, C.c_int, D(np.int32)]
        L.zlh_ib_download_forces.argtypes = [P, C.c_int, D(np.float64)]
        L.zlh_download_node_forces.argtypes = [P, C.c_int, D(np.float64)]
        L.zlh_conservation.argtypes = [P, C.c_int, D(np.float64)]
        L.zlh_level_info.argtypes = [P, C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.zlh_transfer_info.argtypes = [P, C.c_int, C.POINTER(C.c_int64)]
        L.zlh_synchronize.argtypes = [P]
        L.zlh_timing_enable.argtypes = [P, C.c_int]
        L.zlh_timing_read.argtypes = [P, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int64)]
        F32 = D(np.float32)
        L.zlh_mlp_set.argtypes = [P, F32, F32, F32, F32, F32, C.c_double]
        L.zlh_mlp_gradient.argtypes = [P, F32, F32, F32, C.c_int64, F32, F32]
        L.zlh_mlp_gradient_noalloc.argtypes = [P, F32, F32, F32, C.c_int64, F32, F32]
        L.zlh_mlp_train.argtypes = [P, F32, F32, C.c_int64, C.c_int, C.c_int, C.c_double, C.c_uint64, F32, F32]
        L.zlh_mlp_training_form.argtypes = [P, C.c_int]
        L.zlh_mlp_eval.argtypes = [P, F32, C.c_int64, C.c_int, F32]
        L.zlh_closure_features.argtypes = [P, C.c_int, F32, F32, D(np.int32)]
        L.zlh_graphs_enable.argtypes = [P, C.c_int]
        L.zlh_level_streams_enable.argtypes = [P, C.c_int]
        L.zlh_set_boundary_profile.argtypes = [P, C.c_int, C.c_int]
        L.zlh_set_boundary_velocities.argtypes = [P, C.c_int, C.c_int, D(np.float64), C.c_int64]
        L.zlh_set_interpolation_method.argtypes = [P, C.c_int, C.c_int]
        L.zlh_set_time_interpolation.argtypes = [P, C.c_int]
        L.zlh_set_moment_transform.argtypes = [P, C.c_int]
        L.zlh_max_level_bricks.argtypes = [C.POINTER(SolverDesc), C.POINTER(C.c_int64)]
        _lib = L
    return _lib


def make_desc(grid: "_grid.Grid", *, nu0: float, precision: int = FP64, rho0: float = 1.0,
              u0=(0., 0., 0.), collision: Sequence[int] = (SRT,), les: int = LES_NONE,
              les_tau_mode: int = LES_TAU_PHYSICAL, cs: float = 0.1, has_force: bool = False,
              const_force=(0., 0., 0.), bc: Optional[Dict[int, tuple]] = None,
              compressible: bool = True, mlp_operands: int = MLP_OPERANDS_BF16) -> SolverDesc:
    d = SolverDesc()
    d.dims = grid.dims
    d.max_level = grid.max_level
    d.precision = precision
    d.compressible = 1 if compressible else 0
    d.nu0, d.rho0 = nu0, rho0
    for a in range(3):
        d.u0[a] = u0[a] if a < len(u0) else 0.
        d.const_force[a] = const_force[a] if a < len(const_force) else 0.
    d.les, d.les_tau_mode, d.cs = les, les_tau_mode, cs
    d.has_force = 1 if has_force else 0
    d.mlp_operands = mlp_operands
    for l in range(16):
        d.collision[l] = collision[min(l, len(collision) - 1)]
    for dr, (scheme, val) in (bc or {}).items():
        d.bc_scheme[dr] = scheme
        for a in range(3):
            d.bc_value[dr][a] = val[a] if a < len(val) else 0.
    for a in range(grid.dims):
        d.dom_min[a] = grid.dom_min[a]
        d.dom_max[a] = grid.dom_max[a]
    return d


def max_level_bricks(desc: SolverDesc) -> int:
    """Largest brick count (64 nodes each) one level may hold with these settings; no device call."""
    out = C.c_int64()
    rc = lib().zlh_max_level_bricks(C.byref(desc), C.byref(out))
    if rc != 0:
        raise ValueError("zlh_max_level_bricks: status %d" % rc)
    return out.value


def _checked(a, dtype, size, what):
    """Contiguous copy of `a` as `dtype`, raising ValueError unless it holds `size` elements."""
    a = np.ascontiguousarray(a, dtype).ravel()
    if a.size != size:
        raise ValueError("%s: expected %d values, got %d" % (what, size, a.size))
    return a


class Solver:
    """One device solver over a :class:`zllbm_amd.grid.Grid` (the SolverLbmInterface +
    GridInfoLbmInteface pair of the reference, lbm_interface.h:140-474)."""

    def __init__(self, grid: "_grid.Grid", bc_profile=None, bc_velocities=None, interp_method=INTERP_LAGRANGIAN,
                 time_interp=False, moment_transform=TRANSFORM_VALU, **kw):
        """bc_profile: {face: PROFILE_PARABOLIC} for non-equilibrium-extrapolation faces;
        bc_velocities: {(level, face): [n][3] velocities of the level's nodes on that face, in key
        order} (zlh_set_boundary_velocities); interp_method: INTERP_LAGRANGIAN or INTERP_LINEAR
        ghost-node interpolation of every refined level (zlh_set_interpolation_method);
        time_interp: refill the ghost layers after odd substeps too (zlh_set_time_interpolation);
        moment_transform: TRANSFORM_VALU or TRANSFORM_MFMA, where the FP32 learned-closure kernels
        run the MRT transform (zlh_set_moment_transform)."""
        self.grid = grid
        self.dims = grid.dims
        self.q = 9 if grid.dims == 2 else 19
        L = lib()
        self.desc = make_desc(grid, **kw)
        h = C.c_void_p()
        rc = L.zlh_create(C.byref(self.desc), C.byref(h))
        if rc != 0:
            raise RuntimeError("zlh_create failed with status %d" % rc)
        self.h = h
        for face, prof in (bc_profile or {}).items():
            self._chk(L.zlh_set_boundary_profile(self.h, face, prof))
        for level in range(1, grid.max_level + 1):
            self._chk(L.zlh_set_interpolation_method(self.h, level, interp_method))
        for (level, face), ub in (bc_velocities or {}).items():
            ub = np.ascontiguousarray(ub, np.float64).reshape(-1, 3)
            self._chk(L.zlh_set_boundary_velocities(self.h, level, face, ub.ravel(), len(ub)))
        self._chk(L.zlh_set_time_interpolation(self.h, 1 if time_interp else 0))
        self._chk(L.zlh_set_moment_transform(self.h, moment_transform))
        for l, lev in enumerate(grid.levels):
            self._chk(L.zlh_set_level_nodes(self.h, l, _checked(lev.keys, np.uint64, lev.n, "level %d keys" % l),
                                            _checked(lev.flags, np.int32, lev.n, "level %d flags" % l), lev.n))
        for l, lev in enumerate(grid.levels):
            for (kind, layer), keys in lev.layers.items():
                self._chk(L.zlh_set_interface_layer(self.h, l, kind, layer,
                                                    np.ascontiguousarray(keys, np.uint64), keys.size))
        self._ib_n = {}

    def _chk(self, rc):
        if rc != 0:
            raise RuntimeError("zllbm_hip: status %d: %s" % (rc, lib().zlh_last_error(self.h).decode()))

    def close(self):
        if getattr(self, "h", None):
            lib().zlh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_ib(self, level, xyz, area, vel, stencil_dis=2, stiffness=2.0):
        n = len(area)

        def pad3(a):
            a = np.asarray(a, np.float64).reshape(n, -1)
            return np.ascontiguousarray(np.pad(a, ((0, 0), (0, 3 - a.shape[1]))).ravel())
        self._ib_n[level] = n
        self._chk(lib().zlh_ib_set_markers(self.h, level, pad3(xyz), np.ascontiguousarray(area, np.float64),
                                           pad3(vel), n, stencil_dis, stiffness))

    def build(self):
        self._chk(lib().zlh_build(self.h))
        return self

    def init_feq(self):
        self._chk(lib().zlh_init_feq(self.h))
        return self

    def init_macro(self, level, rho, u):
        n = self.n(level)
        self._chk(lib().zlh_init_macros(self.h, level, _checked(rho, np.float64, n, "rho"),
                                        _checked(u, np.float64, n * self.dims, "u")))

    def set_f(self, level, f):
        self._chk(lib().zlh_upload_populations(self.h, level, _checked(f, np.float64, self.n(level) * self.q, "f")))

    def step_level(self, level, time_step_level):
        self._chk(lib().zlh_step_level(self.h, level, time_step_level))

    def transfer_from_coarse(self, fine_level):
        """Coarse-to-fine half of transfer_levels (TransferInfoFromCoarseGrid)."""
        self._chk(lib().zlh_transfer_from_coarse(self.h, fine_level))

    def transfer_to_coarse(self, fine_level):
        """Fine-to-coarse half of transfer_levels (TransferInfoToCoarseGrid)."""
        self._chk(lib().zlh_transfer_to_coarse(self.h, fine_level))

    def transfer_levels(self, fine_level):
        self._chk(lib().zlh_transfer_levels(self.h, fine_level))

    def time_marching(self, nsteps=1):
        self._chk(lib().zlh_time_marching(self.h, nsteps))

    def n(self, level=0) -> int:
        return self.grid.levels[level].n

    def info(self, level=0):
        n, a, b = C.c_int64(), C.c_int64(), C.c_int64()
        self._chk(lib().zlh_level_info(self.h, level, C.byref(n), C.byref(a), C.byref(b)))
        return n.value, a.value, b.value

    def transfer_info(self, level) -> dict:
        """Interpolation tables of a refined level: sizes of the gather's staged data."""
        out = (C.c_int64 * 9)()
        self._chk(lib().zlh_transfer_info(self.h, level, out))
        keys = ("ghost_nodes", "source_rows", "f2c_nodes", "groups", "staged_rows", "contributors",
                "fallback_contributors", "lds_bytes", "overlaps_finer_transfer")
        return dict(zip(keys, list(out)))

    def f(self, level=0) -> np.ndarray:
        out = np.zeros(self.n(level) * self.q)
        self._chk(lib().zlh_download_populations(self.h, level, out))
        return out.reshape(-1, self.q)

    def macros(self, level=0):
        rho = np.zeros(self.n(level))
        u = np.zeros(self.n(level) * self.dims)
        self._chk(lib().zlh_download_macros(self.h, level, rho, u))
        return rho, u.reshape(-1, self.dims)

    def source_table(self, level=0) -> np.ndarray:
        out = np.zeros(self.n(level) * self.q, dtype=np.int32)
        self._chk(lib().zlh_download_source_table(self.h, level, out))
        return out.reshape(-1, self.q)

    def ib_force(self, level) -> np.ndarray:
        out = np.zeros(self._ib_n[level] * 3)
        self._chk(lib().zlh_ib_download_forces(self.h, level, out))
        return out.reshape(-1, 3)[:, :self.dims]

    def node_forces(self, level=0) -> np.ndarray:
        """Node forces [n][3] in key order (the immersed-boundary force on its supports)."""
        out = np.zeros(self.n(level) * 3)
        self._chk(lib().zlh_download_node_forces(self.h, level, out))
        return out.reshape(-1, 3)

    def conservation(self, level=0) -> np.ndarray:
        out = np.zeros(4)
        self._chk(lib().zlh_conservation(self.h, level, out))
        return out

    # ---- learned closure ----
    def mlp_set(self, model):
        f = lambda v: np.ascontiguousarray(v, np.float32)
        self._chk(lib().zlh_mlp_set(self.h, f(model.params), f(model.mean), f(model.std), f(model.fmin),
                                    f(model.fmax), float(model.nu_scale)))

    def mlp_gradient(self, params, x, t, form=MLP_FP32):
        """Mean-squared-error gradient over all samples; form MLP_FP32 (vector units) or
        MLP_BF16_MFMA (matrix cores, BF16 operands)."""
        f = lambda v: np.ascontiguousarray(v, np.float32)
        g = np.zeros(len(params), np.float32)
        loss = np.zeros(1, np.float32)
        self._chk(lib().zlh_mlp_training_form(self.h, form))
        self._chk(lib().zlh_mlp_gradient(self.h, f(params), f(x), f(t), len(t), g, loss))
        return float(loss[0]), g

    def mlp_gradient_noalloc(self, params, x, t, form=MLP_FP32):
        """mlp_gradient through zlh_mlp_gradient_noalloc: no device allocation, at most
        MLP_NOALLOC_MAX_SAMPLES samples."""
        t = np.ascontiguousarray(t, np.float32).ravel()
        params = _checked(params, np.float32, 1409, "params")
        x = _checked(x, np.float32, 9 * len(t), "x")
        g = np.zeros(len(params), np.float32)
        loss = np.zeros(1, np.float32)
        self._chk(lib().zlh_mlp_training_form(self.h, form))
        self._chk(lib().zlh_mlp_gradient_noalloc(self.h, params, x, t, len(t), g, loss))
        return float(loss[0]), g

    def mlp_train(self, params, x, t, steps, batch, lr, seed=20261202, form=MLP_FP32):
        """Adam on the device with the gradient kernel of `form` (as mlp_gradient)."""
        f = lambda v: np.ascontiguousarray(v, np.float32)
        p = f(params).copy()
        hist = np.zeros(steps, np.float32)
        self._chk(lib().zlh_mlp_training_form(self.h, form))
        self._chk(lib().zlh_mlp_train(self.h, f(x), f(t), len(t), steps, batch, lr, seed, p, hist))
        return p, hist

    def mlp_eval(self, x, form=None):
        """Network output of normalised inputs; form MLP_FP32 (default) or MLP_BF16_MFMA."""
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros(len(x), np.float32)
        self._chk(lib().zlh_mlp_eval(self.h, x, len(x), MLP_FP32 if form is None else form, y))
        return y

    def closure_features(self, level=0):
        n = self.n(level)
        x = np.zeros(n * 9, np.float32)
        nu = np.zeros(n, np.float32)
        fb = np.zeros(n, np.int32)
        self._chk(lib().zlh_closure_features(self.h, level, x, nu, fb))
        return x.reshape(n, 9), nu, fb

    def graphs_enable(self, on=True):
        self._chk(lib().zlh_graphs_enable(self.h, 1 if on else 0))

    def set_time_interpolation(self, on=True):
        """Refill the ghost layers after odd substeps too (zlh_set_time_interpolation)."""
        self._chk(lib().zlh_set_time_interpolation(self.h, 1 if on else 0))

    def set_moment_transform(self, where):
        """TRANSFORM_VALU or TRANSFORM_MFMA (zlh_set_moment_transform)."""
        self._chk(lib().zlh_set_moment_transform(self.h, where))

    def level_streams_enable(self, on=True):
        """Run each level's substeps on a stream of its own (default on); results are identical."""
        self._chk(lib().zlh_level_streams_enable(self.h, 1 if on else 0))

    def synchronize(self):
        self._chk(lib().zlh_synchronize(self.h))

    def timing_enable(self, on=True):
        self._chk(lib().zlh_timing_enable(self.h, 1 if on else 0))

    def timing_read(self, level=-1):
        """(total ms, launches) of the fused kernel since timing_enable, one level or all (-1)."""
        ms, n = C.c_double(), C.c_int64()
        self._chk(lib().zlh_timing_read(self.h, level, C.byref(ms), C.byref(n)))
        return ms.value, n.value
