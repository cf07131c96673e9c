"""MI355X-native lattice Boltzmann time step on ZLLBM's Morton-keyed AMR grid.

The compute path is the HIP library ``libzllbm_hip.so`` (built from ``zllbm_amd/csrc``) behind
the C ABI declared in ``include/zllbm_hip.h``.  :mod:`zllbm_amd.hip` loads it and fails loudly
when it is missing; there is no CPU fallback.
"""
from . import grid  # noqa: F401

__all__ = ["grid"]
