"""Grid construction: the criterion-refined builder against the box builder, its structural
invariants on an irregular (vorticity) refinement, and the analytic turbulence field."""
import numpy as np
import pytest

from zllbm_amd import grid as g, workloads
from zllbm_amd.fields import fourier_modes


def _box_criterion(boxes):
    """Boxes in level-k node indices -> cell criterion in background units."""
    def crit(k, c):
        lo, hi = (np.asarray(v, np.float64) / (1 << k) - 1.0 for v in boxes[k])
        return np.all((c > lo) & (c < hi), axis=1)
    return crit


@pytest.mark.parametrize("dims,cells,boxes", [
    (2, (40, 32), [((10, 8), (30, 24)), ((26, 22), (54, 40))]),
    (3, (16, 16, 16), [((5, 5, 5), (12, 12, 12))]),
])
def test_mask_builder_reproduces_box_builder(dims, cells, boxes):
    gb = g.refined_box_grid(dims, cells, boxes)
    gm = g.refined_mask_grid(dims, cells, len(boxes) + 1, _box_criterion(boxes), buffer=0)
    for a, b in zip(gb.levels, gm.levels):
        np.testing.assert_array_equal(a.keys, b.keys)
        np.testing.assert_array_equal(a.flags, b.flags)
        assert sorted(a.layers) == sorted(b.layers)
        for k in a.layers:
            np.testing.assert_array_equal(a.layers[k], b.layers[k])


def test_vorticity_refinement_invariants():
    W = workloads.decaying_turbulence(24, 3, buffer=2)
    G = W.grid
    for l, lev in enumerate(G.levels):
        assert np.all(np.diff(lev.keys.astype(np.int64)) != 0) and np.all(lev.keys[1:] > lev.keys[:-1])
        for keys in lev.layers.values():
            assert np.all(np.isin(keys, lev.keys))
    for l in range(1, len(G.levels)):
        fine, coarse = G.levels[l], G.levels[l - 1]
        # every fine node coinciding with a coarse position whose coarse node was kept lies on
        # the two coarse interface rows, which coincide with the fine outer layers
        xyz = g.morton_decode(3, fine.keys).astype(np.int64)
        even = np.all(xyz % 2 == 0, axis=1)
        parent = g.morton_encode(3, *[(xyz[even, a] // 2).astype(np.uint64) for a in range(3)])
        on_coarse = np.isin(parent, coarse.keys)
        c2f = np.concatenate([coarse.layers[(g.OUTER_C2F, 0)], coarse.layers[(g.OUTER_C2F, 1)]])
        assert np.all(np.isin(parent[on_coarse], c2f))
        # coarse ring r coincides with fine ring 2 r: ring 0 with fine layer 2, ring 1 with layer 0
        assert np.all(np.isin(coarse.layers[(g.OUTER_C2F, 0)] << np.uint64(3), fine.layers[(g.OUTER_F2C, 2)]))
        assert np.all(np.isin(coarse.layers[(g.OUTER_C2F, 1)] << np.uint64(3), fine.layers[(g.OUTER_F2C, 0)]))
        # three fine interface rings, outermost first flagged
        assert all(len(fine.layers[(g.OUTER_F2C, k)]) > 0 for k in range(3))
    # the finest level concentrates where the criterion asked: more of its nodes exceed the
    # threshold than the 20 % of the box (the buffer and the nesting margins dilute it)
    thr = workloads.vorticity_threshold(W.modes)
    xyz = G.coords(2).astype(np.float64) / 4 - 1.0
    assert np.mean(np.linalg.norm(W.modes.vorticity(xyz), axis=1) > thr) > 0.23


def test_fourier_field_rms_divergence_and_curl():
    m = fourier_modes((48, 48, 48), seed=5)
    ax = np.arange(48.0)
    pts = np.stack([a.ravel() for a in np.meshgrid(ax, ax, ax, indexing="ij")], 1)
    u = m.velocity(pts)
    assert abs(np.sqrt(np.mean(u ** 2)) - 0.05 / np.sqrt(3)) < 1e-12
    x0 = np.array([[10.3, 20.1, 5.7], [40.0, 1.5, 33.2]])
    h = 1e-4
    J = np.zeros((2, 3, 3))
    for j in range(3):
        e = np.zeros(3)
        e[j] = h
        J[:, :, j] = (m.velocity(x0 + e) - m.velocity(x0 - e)) / (2 * h)
    assert np.all(np.abs(np.trace(J, axis1=1, axis2=2)) < 1e-9)
    curl = np.stack([J[:, 2, 1] - J[:, 1, 2], J[:, 0, 2] - J[:, 2, 0], J[:, 1, 0] - J[:, 0, 1]], 1)
    np.testing.assert_allclose(m.vorticity(x0), curl, rtol=1e-6, atol=1e-10)
    # E(k) ~ k^-5/3: amplitude^2 of the drawn modes against |k|
    kn = np.linalg.norm(m.k, axis=1)
    slope = np.polyfit(np.log(kn), np.log(m.amp ** 2), 1)[0]
    assert abs(slope - (-11.0 / 3.0)) < 1e-9
