"""Morton-keyed grid description: node key sets, domain bounds and refinement-interface layers.

This is the host-side grid ingest for the LBM time step.  The reference keeps one
``unordered_map`` per level keyed by a 64-bit Morton code (defs_libs.h:35-45,
grid/grid_info_interface.h:258); here a level is the sorted array of those codes plus a flag per
node, which is what the device library builds its slot and neighbour tables from.

Conventions restated from the reference:

* Morton layout: x at bit 3i, y at 3i+1, z at 3i+2 (3D); x at 2i, y at 2i+1 (2D)
  (grid/morton_aux.cpp:238-286).
* Background node indices run over [1, N+1] per axis for a domain of N cells, i.e. both end
  planes exist and periodic domains carry a duplicated plane (grid/grid_manager3d.cpp:105-112).
* A level-l node has level-l indices; a background index i is i << l at level l
  (grid/grid_manager3d.cpp:165-205).
* Refinement interfaces: the fine level keeps three layers (0 innermost .. 2 outermost) along
  the coarse-fine boundary; the coarse level keeps two layers (0 at the fine level's outermost
  row, 1 one coarse spacing inside), layer counts from grid/grid_info_interface.h:207-210 and the
  layer picture at grid/grid_manager.h:97-104.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

# node flags (NodeBitStatus, grid/grid_manager.h:64-73)
FLAG_C2F0 = 1 << 1
FLAG_C2F_GHOST = 1 << 2
FLAG_F2C0 = 1 << 3
FLAG_F2C_GHOST = 1 << 4

# boundary directions (EDomainBoundaryDirection, grid/grid_enumerates.h:37-45)
XMIN, XMAX, YMIN, YMAX, ZMIN, ZMAX = 1, 2, 3, 4, 5, 6
# boundary schemes (ELbmBoundaryConditionScheme, boundary_conditions/lbm_boundary_conditions.h)
BC_BOUNCE_BACK, BC_PERIODIC, BC_NEE = 1, 2, 3
# interface layer kinds (InterfaceLayerInfo, grid/grid_info_interface.h:117-127)
OUTER_F2C, OUTER_C2F, INNER_F2C, INNER_C2F = 0, 1, 2, 3


_U = np.uint64


def _spread3(v):
    v = np.asarray(v, dtype=_U) & _U(0x1FFFFF)
    v = (v | (v << _U(32))) & _U(0x1F00000000FFFF)
    v = (v | (v << _U(16))) & _U(0x1F0000FF0000FF)
    v = (v | (v << _U(8))) & _U(0x100F00F00F00F00F)
    v = (v | (v << _U(4))) & _U(0x10C30C30C30C30C3)
    v = (v | (v << _U(2))) & _U(0x1249249249249249)
    return v


def _compact3(v):
    v = np.asarray(v, dtype=_U) & _U(0x1249249249249249)
    v = (v | (v >> _U(2))) & _U(0x10C30C30C30C30C3)
    v = (v | (v >> _U(4))) & _U(0x100F00F00F00F00F)
    v = (v | (v >> _U(8))) & _U(0x1F0000FF0000FF)
    v = (v | (v >> _U(16))) & _U(0x1F00000000FFFF)
    v = (v | (v >> _U(32))) & _U(0x1FFFFF)
    return v


def _spread2(v):
    v = np.asarray(v, dtype=_U) & _U(0xFFFFFFFF)
    v = (v | (v << _U(16))) & _U(0x0000FFFF0000FFFF)
    v = (v | (v << _U(8))) & _U(0x00FF00FF00FF00FF)
    v = (v | (v << _U(4))) & _U(0x0F0F0F0F0F0F0F0F)
    v = (v | (v << _U(2))) & _U(0x3333333333333333)
    v = (v | (v << _U(1))) & _U(0x5555555555555555)
    return v


def _compact2(v):
    v = np.asarray(v, dtype=_U) & _U(0x5555555555555555)
    v = (v | (v >> _U(1))) & _U(0x3333333333333333)
    v = (v | (v >> _U(2))) & _U(0x0F0F0F0F0F0F0F0F)
    v = (v | (v >> _U(4))) & _U(0x00FF00FF00FF00FF)
    v = (v | (v >> _U(8))) & _U(0x0000FFFF0000FFFF)
    v = (v | (v >> _U(16))) & _U(0xFFFFFFFF)
    return v


def morton_encode(dims: int, x, y, z=None) -> np.ndarray:
    """Interleave integer index arrays into 64-bit Morton codes (SFBitsetEncoding,
    grid/morton_aux.cpp:274-286): x at bit dims*i, y at dims*i+1, z at 3i+2."""
    if dims == 2:
        return _spread2(x) | (_spread2(y) << _U(1))
    return _spread3(x) | (_spread3(y) << _U(1)) | (_spread3(z) << _U(2))


def morton_decode(dims: int, keys) -> np.ndarray:
    """Inverse of :func:`morton_encode`; returns an (n, dims) uint32 index array
    (SFBitsetComputeIndices, grid/morton_aux.cpp:336-351)."""
    keys = np.asarray(keys, dtype=_U)
    if dims == 2:
        cols = [_compact2(keys), _compact2(keys >> _U(1))]
    else:
        cols = [_compact3(keys), _compact3(keys >> _U(1)), _compact3(keys >> _U(2))]
    return np.stack(cols, axis=-1).astype(np.uint32)


@dataclass
class Level:
    """One refinement level: sorted Morton keys, flags and interface layer key sets."""
    keys: np.ndarray
    flags: np.ndarray
    layers: Dict[Tuple[int, int], np.ndarray] = field(default_factory=dict)

    @property
    def n(self) -> int:
        return int(self.keys.size)


@dataclass
class Grid:
    """All levels of a Morton-keyed AMR grid over a box of background cells."""
    dims: int
    cells: Tuple[int, ...]                 # background cells per axis (domain_size / dx)
    levels: List[Level]
    dom_min: Tuple[int, ...] = ()
    dom_max: Tuple[int, ...] = ()

    def __post_init__(self):
        if not self.dom_min:
            self.dom_min = tuple([1] * self.dims)
        if not self.dom_max:
            self.dom_max = tuple(1 + c for c in self.cells)

    @property
    def max_level(self) -> int:
        return len(self.levels) - 1

    def coords(self, level: int) -> np.ndarray:
        return morton_decode(self.dims, self.levels[level].keys)


def _box_keys(dims: int, lo: Sequence[int], hi: Sequence[int]) -> np.ndarray:
    rng = [np.arange(lo[a], hi[a] + 1, dtype=np.uint64) for a in range(dims)]
    mesh = np.meshgrid(*rng, indexing="ij")
    flat = [m.ravel() for m in mesh]
    return morton_encode(dims, *flat)


def uniform_grid(dims: int, cells: Sequence[int]) -> Grid:
    """Single-level grid of ``cells`` background cells per axis, indices [1, cells+1]."""
    cells = tuple(int(c) for c in cells)
    keys = np.sort(_box_keys(dims, [1] * dims, [c + 1 for c in cells]))
    return Grid(dims, cells, [Level(keys, np.zeros(keys.size, dtype=np.int32))])


def refined_box_grid(dims: int, cells: Sequence[int],
                     boxes: Sequence[Tuple[Sequence[int], Sequence[int]]]) -> Grid:
    """Nested box refinement.

    ``boxes[k] = (lo, hi)`` is the region refined to level k+1, given in level-k node indices
    (inclusive); each box must lie strictly inside the previous level's interior and keep at
    least three coarse spacings to the previous box's edge (the reference's minimum extension,
    grid/grid_manager.h:97).  The fine level covers the box's nodes at level k+1 resolution.  The
    coarse level keeps its nodes outside the box plus its two outermost rows inside the box; the
    fine level's three outermost rows form the fine-to-coarse layers.
    """
    cells = tuple(int(c) for c in cells)
    sets: List[np.ndarray] = [_box_keys(dims, [1] * dims, [c + 1 for c in cells])]
    prev_lo = np.full(dims, 1, dtype=np.int64)
    prev_hi = np.asarray([c + 1 for c in cells], dtype=np.int64)
    for k, (lo, hi) in enumerate(boxes):
        lo = np.asarray(lo, dtype=np.int64)
        hi = np.asarray(hi, dtype=np.int64)
        if np.any(hi - lo < 6):
            raise ValueError("refinement box too small")
        if np.any(lo - prev_lo < 4) or np.any(prev_hi - hi < 4):
            raise ValueError("refinement box %d must keep 4 spacings to the enclosing region" % k)
        sets.append(_box_keys(dims, list(2 * lo), list(2 * hi)))
        prev_lo, prev_hi = 2 * lo, 2 * hi
    nlev = len(sets)
    levels: List[Level] = []
    # remove coarse nodes strictly inside each box beyond the two interface rows
    for k in range(nlev):
        keys = sets[k]
        lay: Dict[Tuple[int, int], np.ndarray] = {}
        fl = np.zeros(keys.size, dtype=np.int32)
        xyz = morton_decode(dims, keys).astype(np.int64)
        if k < nlev - 1:
            lo, hi = (np.asarray(boxes[k][0], dtype=np.int64), np.asarray(boxes[k][1], dtype=np.int64))
            # ring index: 0 on the box edge, 1 one row inside, ... (Chebyshev distance to edge)
            inside = np.all((xyz >= lo) & (xyz <= hi), axis=1)
            dist = np.min(np.concatenate([xyz - lo, hi - xyz], axis=1), axis=1)
            keep = ~inside | (dist <= 1)
            c2f0 = inside & (dist == 0)
            c2f1 = inside & (dist == 1)
            lay[(OUTER_C2F, 0)] = keys[c2f0]
            lay[(OUTER_C2F, 1)] = keys[c2f1]
            # layer 1 is the outmost coarse-to-fine layer and the only ghost one
            # (InstantiateOverlapLayerOfRefinementInterface, grid/grid_manager.cpp:413-568)
            fl[c2f1] |= FLAG_C2F_GHOST | FLAG_C2F0
            keys, fl, xyz = keys[keep], fl[keep], xyz[keep]
        if k > 0:
            lo, hi = (2 * np.asarray(boxes[k - 1][0], dtype=np.int64),
                      2 * np.asarray(boxes[k - 1][1], dtype=np.int64))
            dist = np.min(np.concatenate([xyz - lo, hi - xyz], axis=1), axis=1)
            for layer, d in ((2, 0), (1, 1), (0, 2)):
                sel = dist == d
                lay[(OUTER_F2C, layer)] = keys[sel]
                if layer == 2:
                    fl[sel] |= FLAG_F2C0 | FLAG_F2C_GHOST
                elif layer == 1:
                    fl[sel] |= FLAG_F2C_GHOST
        order = np.argsort(keys)
        keys, fl = keys[order], fl[order]
        lev = Level(keys, fl)
        for kk, v in lay.items():
            lev.layers[kk] = np.sort(v)
        levels.append(lev)
    return Grid(dims, cells, levels)


def _node_closure(cells: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """Nodes of a cell set (array shape + 1 per axis): the closure (some incident cell in the set)
    and the interior (every incident cell in the set)."""
    import itertools
    pad = np.pad(cells, 1)
    shape = tuple(n + 1 for n in cells.shape)
    anyc = np.zeros(shape, dtype=bool)
    allc = np.ones(shape, dtype=bool)
    for off in itertools.product((0, 1), repeat=cells.ndim):
        sl = pad[tuple(slice(o, o + n) for o, n in zip(off, shape))]
        anyc |= sl
        allc &= sl
    return anyc, allc


def _rings(closure: np.ndarray, interior: np.ndarray, count: int) -> np.ndarray:
    """Chebyshev ring index of the closure nodes from its boundary (0 on the boundary), for the
    first `count` rings; 255 beyond and outside."""
    from scipy import ndimage
    ring = np.full(closure.shape, 255, dtype=np.uint8)
    front = closure & ~interior
    reached = front.copy()
    ring[front] = 0
    box = np.ones((3,) * closure.ndim, dtype=bool)
    for r in range(1, count):
        front = ndimage.binary_dilation(front, structure=box) & closure & ~reached
        ring[front] = r
        reached |= front
    return ring


def refined_mask_grid(dims: int, cells: Sequence[int], levels: int, refine, margin: int = 4,
                      buffer: int = 2) -> Grid:
    """Refinement by a cell criterion (the general form of refined_box_grid).

    ``refine(k, centres)`` returns, for the cell centres of level k (background units, origin at
    the domain minimum, array [n, dims]), which of them ask for level k+1.  A flagged cell is
    refined if it lies at least ``margin`` level-k cells inside level k's region and inside the
    domain faces; the flags grow by ``buffer`` cells (the reference's extension layers,
    grid/grid_manager.h:97) within that allowance.  Interface layers and flags follow
    refined_box_grid with the ring index (Chebyshev distance to the region boundary) in place of
    the box distance.
    """
    from scipy import ndimage
    cells = tuple(int(c) for c in cells)
    box = np.ones((3,) * dims, dtype=bool)
    regions = [np.ones(cells, dtype=bool)]            # level-k cells of level k's region
    flagged: List[np.ndarray] = []
    for k in range(levels - 1):
        region = regions[k]
        allowed = ndimage.binary_erosion(region, structure=box, iterations=margin, border_value=0)
        axes = [(np.arange(n, dtype=np.float64) + 0.5) / (1 << k) for n in region.shape]
        centres = np.stack([a.ravel() for a in np.meshgrid(*axes, indexing="ij")], 1)
        flag = np.asarray(refine(k, centres), dtype=bool).reshape(region.shape) & allowed
        if buffer:
            flag = ndimage.binary_dilation(flag, structure=box, iterations=buffer) & allowed
        if not flag.any():
            raise ValueError("level %d: no cell is refined" % (k + 1))
        flagged.append(flag)
        fine = flag
        for a in range(dims):
            fine = np.repeat(fine, 2, axis=a)
        regions.append(fine)
    out: List[Level] = []
    for k in range(levels):
        lo = 1 << k
        closure, interior = _node_closure(regions[k])
        keep = closure.copy()
        fl = np.zeros(closure.shape, dtype=np.int32)
        lay_masks: Dict[Tuple[int, int], np.ndarray] = {}
        if k < levels - 1:
            c_clo, c_int = _node_closure(flagged[k])
            ring = _rings(c_clo, c_int, 2)
            keep &= ~c_clo | (ring <= 1)
            lay_masks[(OUTER_C2F, 0)] = ring == 0
            lay_masks[(OUTER_C2F, 1)] = ring == 1
            # layer 1 is the outmost coarse-to-fine layer and the only ghost one
            # (InstantiateOverlapLayerOfRefinementInterface, grid/grid_manager.cpp:413-568)
            fl[ring == 1] |= FLAG_C2F_GHOST | FLAG_C2F0
        if k > 0:
            ring = _rings(closure, interior, 3)
            for layer, r in ((2, 0), (1, 1), (0, 2)):
                lay_masks[(OUTER_F2C, layer)] = ring == r
            fl[ring == 0] |= FLAG_F2C0 | FLAG_F2C_GHOST
            fl[ring == 1] |= FLAG_F2C_GHOST
        idx = np.nonzero(keep)
        keys = morton_encode(dims, *[(i + lo).astype(np.uint64) for i in idx])
        flags = fl[idx]
        order = np.argsort(keys)
        lev = Level(keys[order], flags[order])
        for kk, m in lay_masks.items():
            li = np.nonzero(m & keep)
            lev.layers[kk] = np.sort(morton_encode(dims, *[(i + lo).astype(np.uint64) for i in li]))
        out.append(lev)
    return Grid(dims, cells, out)


def boundary_masks(grid: Grid, level: int) -> Dict[int, np.ndarray]:
    """Nodes on each domain face at ``level`` (CheckIfNodeOutsideCubicDomain,
    grid/grid_info.cpp:84-142)."""
    xyz = grid.coords(level).astype(np.int64)
    out = {}
    for a in range(grid.dims):
        lo = grid.dom_min[a] << level
        hi = grid.dom_max[a] << level
        out[1 + 2 * a] = xyz[:, a] == lo
        out[2 + 2 * a] = xyz[:, a] == hi
    return out


def write_grid_file(grid: Grid, path: str) -> None:
    """Serialise a grid for the C++ host driver (zllbm_lbm_driver): magic "ZLGRID1", dims,
    level count, background index bounds, then per level the keys, flags and interface layers."""
    with open(path, "wb") as fh:
        fh.write(b"ZLGRID1\0")
        lo = list(grid.dom_min) + [0] * (3 - grid.dims)
        hi = list(grid.dom_max) + [0] * (3 - grid.dims)
        np.asarray([grid.dims, len(grid.levels)], dtype=np.int32).tofile(fh)
        np.asarray(lo + hi, dtype=np.uint32).tofile(fh)
        for lev in grid.levels:
            np.asarray([lev.n], dtype=np.int64).tofile(fh)
            np.ascontiguousarray(lev.keys, dtype=np.uint64).tofile(fh)
            np.ascontiguousarray(lev.flags, dtype=np.int32).tofile(fh)
            np.asarray([len(lev.layers)], dtype=np.int32).tofile(fh)
            for (kind, layer), keys in sorted(lev.layers.items()):
                np.asarray([kind, layer], dtype=np.int32).tofile(fh)
                np.asarray([keys.size], dtype=np.int64).tofile(fh)
                np.ascontiguousarray(keys, dtype=np.uint64).tofile(fh)
