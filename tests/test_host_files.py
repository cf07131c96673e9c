"""The C++ example driver's input files as zllbm_amd.host_driver writes them: the initial-field
file (per level: int64 count, densities, velocity components) and the immersed-boundary file
(records of int32 level, int32 stencil, double stiffness, int64 n, coordinates and velocities
padded to 3 components, areas), read back field by field."""
import numpy as np

from zllbm_amd import host_driver


def test_init_file_layout(tmp_path):
    rng = np.random.default_rng(1)
    init = [(rng.random(5), rng.random((5, 3))), (rng.random(2), rng.random((2, 3)))]
    path = tmp_path / "init.bin"
    host_driver.write_init(str(path), init)
    raw = path.read_bytes()
    off = 0
    for rho, u in init:
        n = int(np.frombuffer(raw, "<i8", 1, off)[0])
        off += 8
        assert n == len(rho)
        np.testing.assert_array_equal(np.frombuffer(raw, "<f8", n, off), rho)
        off += 8 * n
        np.testing.assert_array_equal(np.frombuffer(raw, "<f8", 3 * n, off).reshape(n, 3), u)
        off += 24 * n
    assert off == len(raw)


def test_ib_file_layout(tmp_path):
    xy = np.array([[1.0, 2.0], [3.0, 4.0], [5.0, 6.5]])
    area = np.array([0.1, 0.2, 0.3])
    vel = np.array([[0.01, 0.0], [0.0, 0.02], [0.03, 0.04]])
    xyz = np.array([[7.0, 8.0, 9.0]])
    path = tmp_path / "ib.bin"
    host_driver.write_ib(str(path), [(2, xy, area, vel, 1, 2.5), (0, xyz, [1.5], np.zeros((1, 3)), 2, 2.0)])
    raw = path.read_bytes()
    off = 0
    for level, x, a, v, stencil, stiff in [(2, xy, area, vel, 1, 2.5), (0, xyz, [1.5], np.zeros((1, 3)), 2, 2.0)]:
        lv, st = np.frombuffer(raw, "<i4", 2, off)
        off += 8
        k = np.frombuffer(raw, "<f8", 1, off)[0]
        off += 8
        n = int(np.frombuffer(raw, "<i8", 1, off)[0])
        off += 8
        assert (lv, st, k, n) == (level, stencil, stiff, len(a))
        pad = lambda m: np.pad(np.asarray(m, float), ((0, 0), (0, 3 - np.asarray(m).shape[1])))
        np.testing.assert_array_equal(np.frombuffer(raw, "<f8", 3 * n, off).reshape(n, 3), pad(x))
        off += 24 * n
        np.testing.assert_array_equal(np.frombuffer(raw, "<f8", n, off), a)
        off += 8 * n
        np.testing.assert_array_equal(np.frombuffer(raw, "<f8", 3 * n, off).reshape(n, 3), pad(v))
        off += 24 * n
    assert off == len(raw)
