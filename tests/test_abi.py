"""The C-ABI library loads, exports every declared entry point and validates arguments
before touching a device (CPU only)."""
import ctypes as C
import os
import re

import pytest

from zllbm_amd import hip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(zlh_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    L = hip.lib()
    declared = _declared()
    assert len(declared) >= 20
    missing = [n for n in sorted(declared) if not hasattr(L, n)]
    assert not missing, missing
    assert set(hip.EXPORTED) <= declared


def test_argument_validation_without_device():
    L = hip.lib()
    h = C.c_void_p()
    assert L.zlh_create(None, C.byref(h)) == -1
    d = hip.SolverDesc()
    d.dims = 4
    assert L.zlh_create(C.byref(d), C.byref(h)) == -1
    d.dims, d.nu0 = 3, 0.0
    assert L.zlh_create(C.byref(d), C.byref(h)) == -1          # viscosity must be positive
    d.nu0, d.precision = 0.01, 7
    assert L.zlh_create(C.byref(d), C.byref(h)) == -1          # unknown precision
    assert L.zlh_last_error(None) == b"null context"
    assert L.zlh_step_level(None, 0, 1) == -1
    assert L.zlh_build(None) == -1


def _declared_arity():
    """Parameter count of every function declared in include/*.h."""
    out = {}
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            for name, args in re.findall(r"\b(zlh_[a-z0-9_]+)\s*\(([^;{]*?)\)\s*;", text, flags=re.S):
                args = args.strip()
                out[name] = 0 if args in ("", "void") else args.count(",") + 1
    return out


def test_binding_arity_matches_header():
    L = hip.lib()
    arity = _declared_arity()
    for name in hip.EXPORTED:
        assert name in arity, name
        assert len(getattr(L, name).argtypes) == arity[name], name


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus 2 exits with status 2 and says so, before any device work."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    proc = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                          text=True, timeout=120)
    assert proc.returncode == 2
    assert "one GPU" in proc.stderr


def test_bench_refuses_bf16_with_learned_closure():
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    proc = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--populations", "bf16"],
                          capture_output=True, text=True, timeout=300)
    assert proc.returncode == 2
    assert "learned closure" in proc.stderr


def test_level_size_limit_without_device():
    """The FP32 learned closure's fused kernel forms 32-bit byte offsets of its pulls and of its
    deferred-node stores, so a level holds fewer than 2^32 / (Q 64 4) bricks with it (883,011 in D3Q19); other settings keep the
    32-bit slot-index bound.  zlh_build refuses larger levels with this limit."""
    from zllbm_amd import grid as g
    grid = g.uniform_grid(3, (4, 4, 4))
    general = (2 ** 31 // 19 - 1) // 64
    assert hip.max_level_bricks(hip.make_desc(grid, nu0=0.01)) == general
    assert hip.max_level_bricks(hip.make_desc(grid, nu0=0.01, precision=hip.FP32_SHIFTED)) == general
    assert hip.max_level_bricks(hip.make_desc(grid, nu0=0.01, precision=hip.FP64, les=hip.LES_MLP)) == general
    mlp = hip.max_level_bricks(hip.make_desc(grid, nu0=0.01, precision=hip.FP32_SHIFTED, les=hip.LES_MLP))
    assert mlp == (2 ** 32 - 1) // (19 * 64 * 4) == 883011
    # the largest admitted level's last pull offset stays below 2^32, and so does the spare slot
    # of its deferred-node stores (Q naux 4 with every node deferred)
    assert (mlp - 1) * 19 * 256 + 19 * 256 - 1 < 2 ** 32
    assert 19 * (mlp * 64) * 4 < 2 ** 32
    grid2 = g.uniform_grid(2, (4, 4))
    assert hip.max_level_bricks(hip.make_desc(grid2, nu0=0.01, precision=hip.FP32_SHIFTED,
                                              les=hip.LES_MLP)) == min(general, 2 ** 32 // (9 * 256))
    bad = hip.make_desc(grid, nu0=0.01)
    bad.dims = 5
    with pytest.raises(ValueError):
        hip.max_level_bricks(bad)


def test_wrapper_refuses_short_arrays():
    """init_macro / set_f check their array lengths before the library reads n * dims or n * Q
    values through the raw pointer."""
    import numpy as np
    from zllbm_amd import grid as g
    grid = g.uniform_grid(2, (4, 4))
    s = hip.Solver.__new__(hip.Solver)       # no device: only the wrapper's checks run
    s.grid, s.dims, s.q, s.h = grid, 2, 9, None
    n = grid.levels[0].n
    with pytest.raises(ValueError):
        s.init_macro(0, np.ones(n - 1), np.zeros((n, 2)))
    with pytest.raises(ValueError):
        s.init_macro(0, np.ones(n), np.zeros((n, 3)))
    with pytest.raises(ValueError):
        s.set_f(0, np.zeros((n, 8)))
    with pytest.raises(ValueError):
        hip._checked(np.zeros(3), np.uint64, 4, "keys")
