"""Geometry-driven grid generation: the node sets, interface layers and flags of every level.

A restatement of the reference's serial generator for the geometry criterion
(GridManagerInterface::GenerateGridFromHighToLowLevelSerial, grid/grid_generation_serial.cpp:20-225,
and InstantiateGridNodeAllLevel, grid/grid_manager.cpp:413-634, with the 2D/3D helpers of
grid/grid_generation_serial{2d,3d}.cpp and grid/grid_manager{2d,3d}.cpp).  Levels are sets of
integer coordinate tuples (level-l index of a background node i is i << l); the result is a
:class:`zllbm_amd.grid.Grid` that the solver, the oracle and the C++ host take as they take the box
and mask grids.

Pipeline, from the finest level down (``one_lower[l]`` is level l's region written with level l-1
indices, the reference's ``sfbitset_one_lower_level``):

1. tracking nodes: every geometry vertex's node at the geometry's level, floor(x 2^l + 1e-6)
   (GeometryInfoOrigin::FindTrackingNodeBasedOnGeo, criterion/geometry_info_origin.cpp:27-58);
2. the coarse nodes of each tracking node's cell neighbourhood, 3^d per node
   (GenerateGridNodeNearTrackingNode); for kInAndOut geometries a flood fill from the geometry's
   origin splits them into the innermost and outermost layers (IdentifyTypeOfLayerByFloodFill,
   FloodFillForInAndOut);
3. each layer grows by the per-level extension counts, layer by layer, in the directions whose
   count is not used up (ExtendGivenNumbOfLayer, ExtendOneLayerGrid,
   FindAllNeighborsWithSpecifiedDirection), with counts from SetNumberOfExtendLayerForGrid
   (grid/grid_manager.cpp:299-379: at least k0IntExtendMin_ = 3, one more on the positive side,
   one less at the tracking level whose neighbourhood layer is already there);
4. the grown region's outermost nodes with an absent neighbour are the coarse level's innermost
   coarse-to-fine layer, and their coarse parents (with their neighbours that straddle the
   region's edge) seed the next coarser level (FindOutmostLayerForFineGrid);
5. node instantiation: the fine-to-coarse layers of each interface from the cells around the coarse
   innermost layer, the coarse outer layer from their coincident nodes
   (InstantiateOverlapLayerOfRefinementInterface, FindOverlappingLayersBasedOnOutermostCoarse,
   IdentifyInterfaceForACell, IdentifyInterfaceNodeOnEdge, IdentifyInterfaceNodeDiagonal,
   OverlapLayerFromHighToLow); every fine node of each region cell whose corners are all in the
   region (NodesBelongToOneCell, FindAllNodesInACellAtOneLevelLower); the background level minus
   the nodes of finer regions, except its coarse-to-fine layers (CheckCoincideBackground,
   InstantiateBackgroundGrid).

Kept as the reference has them, because they decide node sets: the neighbourhood of step 2 is
checked against the domain bounds of the geometry's level, not of the level it is written in
(GenerateGridNodeNearTrackingNode, grid_generation_serial{2d,3d}.cpp); in 2D the parent's
positive-y edge test in step 4 reuses the child's (FindOutmostLayerForFineGrid,
grid_generation_serial2d.cpp:438-447); both coarse-to-fine layers of an inner interface carry the
ghost flag (the fine level's ghost count is applied, grid_manager.cpp:480-483); an extension count
of zero wraps (unsigned) and keeps growing in that direction for as many layers as the largest
count.  Not restated: the MPI partition path (the reference's serial InstantiateGridNodeAllLevel
branch is the one followed here), geometry criteria other than the vertex set, and the
domain-boundary bookkeeping (the solver classifies boundary nodes itself).
"""
from __future__ import annotations

import itertools
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Set, Tuple

import numpy as np

from zllbm_amd import grid as g

Node = Tuple[int, ...]
_EPS = 1e-6             # kEps (defs_libs.h:103)
_EXTEND_MIN = 3         # GridManagerInterface::k0IntExtendMin_ (grid/grid_manager.h:97)
_NUM_F2C, _NUM_C2F = 3, 2                                   # grid/grid_info_interface.h:207-208
_NUM_F2C_GHOST, _NUM_C2F_GHOST = _NUM_F2C // 2 + 1, _NUM_C2F // 2   # :209-210
_MAX_FLOOD_FILL = 9000000   # K0IMaxFloodFill_ (grid/grid_manager.h:308)
_WRAP = 1 << 32          # an extension count past zero (DefAmrLUint arithmetic)


@dataclass
class GeometrySpec:
    """A geometry criterion: its vertices (background units, node i at coordinate i), the level it
    is resolved at and its extension counts per level (GeometryInfoInterface setters
    SetXExtendNegative ... SetInnerExtend, SetGridExtendType, criterion/geometry_info_interface.h)."""
    level: int
    vertices: np.ndarray
    extend_type: str = "out"                 # "out" (kOut) or "in_and_out" (kInAndOut)
    x_neg: Sequence[int] = ()
    x_pos: Sequence[int] = ()
    y_neg: Sequence[int] = ()
    y_pos: Sequence[int] = ()
    z_neg: Sequence[int] = ()
    z_pos: Sequence[int] = ()
    inner: Sequence[int] = ()
    flood_fill_origin: Optional[Sequence[float]] = None   # default: the vertices' mean


@dataclass
class _Interface:
    inner_c2f: Optional[List[Set[Node]]] = None
    outer_c2f: Optional[List[Set[Node]]] = None
    inner_f2c: Optional[List[Set[Node]]] = None
    outer_f2c: Optional[List[Set[Node]]] = None


@dataclass
class Generated:
    grid: g.Grid
    one_lower: List[Set[Node]]               # each level's region in the next coarser indices
    tracking: Dict[int, Set[Node]]           # geometry -> tracking nodes at its level
    interfaces: List[Dict[int, _Interface]] = field(default_factory=list)


class _Gen:
    def __init__(self, dims: int, cells: Sequence[int], max_level: int):
        self.d = dims
        self.cells = tuple(int(c) for c in cells)
        self.max_level = max_level
        self.offsets = [o for o in itertools.product((-1, 0, 1), repeat=dims) if any(o)]
        self.cube = list(itertools.product((-1, 0, 1), repeat=dims))

    # domain bounds of a level (CalDomainBoundsAtGivenLevel, ComputeSFBitsetOnBoundaryAtGivenLevel)
    def lo(self, level: int) -> Node:
        return tuple(1 << level for _ in range(self.d))

    def hi(self, level: int) -> Node:
        return tuple((c + 1) << level for c in self.cells)

    def inside(self, n: Node, level: int) -> bool:
        lo, hi = self.lo(level), self.hi(level)
        return all(lo[a] <= n[a] <= hi[a] for a in range(self.d))

    def neighbors_dir(self, n: Node, neg: Sequence[bool], pos: Sequence[bool]) -> List[Node]:
        """FindAllNeighborsWithSpecifiedDirection: the neighbours whose offset only steps in
        permitted directions (axis and diagonal ones)."""
        out = []
        for o in self.offsets:
            if all((o[a] != -1 or neg[a]) and (o[a] != 1 or pos[a]) for a in range(self.d)):
                out.append(tuple(n[a] + o[a] for a in range(self.d)))
        return out

    def all_neighbors(self, n: Node) -> List[Node]:
        return [tuple(n[a] + o[a] for a in range(self.d)) for o in self.cube]

    # --------------------------------------------------------------------------------------
    def extend_counts(self, geo: GeometrySpec, level: int):
        """SetNumberOfExtendLayerForGrid: (inner neg, inner pos, outer neg, outer pos) per axis."""
        def take(seq):
            return seq[level] if len(seq) > level and seq[level] > _EXTEND_MIN else _EXTEND_MIN
        inner = take(list(geo.inner))
        per_axis = [(list(geo.x_neg), list(geo.x_pos)), (list(geo.y_neg), list(geo.y_pos)),
                    (list(geo.z_neg), list(geo.z_pos))][:self.d]
        return ([inner] * self.d, [inner + 1] * self.d, [take(n) for n, _ in per_axis],
                [take(p) + 1 for _, p in per_axis])

    def near_tracking(self, level: int, tracking: Set[Node]) -> Set[Node]:
        """GenerateGridNodeNearTrackingNode: each tracking node's parent and its 3^d - 1
        neighbours, kept when inside the bounds of the tracking level (level-(l-1) indices
        compared with level-l bounds, as the reference does)."""
        out = set()
        lo, hi = self.lo(level), self.hi(level)
        for t in tracking:
            p = tuple(c >> 1 for c in t)
            for q in self.all_neighbors(p):
                if all(lo[a] <= q[a] <= hi[a] for a in range(self.d)):
                    out.add(q)
        return out

    def flood_fill(self, level: int, origin: Sequence[float], exist: Set[Node]):
        """IdentifyTypeOfLayerByFloodFill + FloodFillForInAndOut at `level`: (outside, inside)."""
        idx = tuple(int(origin[a] * (1 << level) + _EPS) for a in range(self.d))
        idx_max = tuple(((c + 1) << level) - idx[a] for a, c in enumerate(self.cells))
        start = None
        # search -x, -y(, -z), +x, +y(, +z) from the origin for the first existing node; start
        # one step back towards the origin
        for sign, limits in ((-1, idx), (1, idx_max)):
            for a in range(self.d):
                if start is not None:
                    break
                cur = list(idx)
                for _ in range(limits[a]):
                    cur[a] += sign
                    if tuple(cur) in exist:
                        cur[a] -= sign
                        start = tuple(cur)
                        break
        if idx in exist:
            start = idx
        if start is None:
            raise ValueError("no start node for the flood fill at level %d" % level)
        outside, inside = set(), set()
        if start in exist:
            # the origin lies on the layer: every node missing a neighbour is outside
            for n in exist:
                if any(q not in exist for q in self.all_neighbors(n)):
                    outside.add(n)
            return outside, inside
        ring = set()
        for n in exist:
            for q in self.all_neighbors(n):
                if q not in exist:
                    ring.add(q)
        colored = set()
        stack = [start]
        steps = 0
        while stack and steps < _MAX_FLOOD_FILL:
            s = stack.pop()
            steps += 1
            if s in exist:
                inside.add(s)
            elif s in ring and s not in colored:
                colored.add(s)
                stack.extend(q for q in self.all_neighbors(s) if q != s)
        for n in exist:
            if n not in inside and any(q not in exist for q in self.all_neighbors(n)):
                outside.add(n)
        return outside, inside

    def extend_one(self, start: Set[Node], level: int, ext_neg, ext_pos, exist: Set[Node],
                   outmost: Set[Node]) -> Set[Node]:
        """ExtendOneLayerGrid at level-(level-1) indices: the new layer; nodes that stopped
        growing in some direction and still miss a neighbour join `outmost`."""
        lo, hi = self.lo(level - 1), self.hi(level - 1)
        out, tmp = set(), set()
        for n in start:
            on_lo = [n[a] == lo[a] for a in range(self.d)]
            on_hi = [n[a] == hi[a] for a in range(self.d)]
            neg = [not on_lo[a] and ext_neg[a] for a in range(self.d)]
            pos = [not on_hi[a] and ext_pos[a] for a in range(self.d)]
            for q in self.neighbors_dir(n, neg, pos):
                if q not in exist:
                    out.add(q)
                    exist.add(q)
            for a in range(self.d):
                if (not neg[a] and not on_lo[a]) or (not pos[a] and not on_hi[a]):
                    tmp.add(n)
        for n in tmp:
            neg = [n[a] != lo[a] for a in range(self.d)]
            pos = [n[a] != hi[a] for a in range(self.d)]
            if any(q not in exist for q in self.neighbors_dir(n, neg, pos)):
                outmost.add(n)
        return out

    def extend(self, level: int, num_neg, num_pos, start: Set[Node], exist: Set[Node]) -> Set[Node]:
        """ExtendGivenNumbOfLayer: returns the outmost layer."""
        num_neg, num_pos = list(num_neg), list(num_pos)
        layers = max(num_neg + num_pos) if num_neg else 0
        ext_neg, ext_pos = [True] * self.d, [True] * self.d
        outmost: Set[Node] = set()
        cur = set(start)
        for _ in range(layers):
            cur = self.extend_one(cur, level, ext_neg, ext_pos, exist, outmost)
            for a in range(self.d):
                if ext_neg[a]:
                    num_neg[a] = num_neg[a] - 1 if num_neg[a] > 0 else _WRAP
                    if num_neg[a] == 0:
                        ext_neg[a] = False
                if ext_pos[a]:
                    num_pos[a] = num_pos[a] - 1 if num_pos[a] > 0 else _WRAP
                    if num_pos[a] == 0:
                        ext_pos[a] = False
        outmost |= cur
        return outmost

    def outmost_for_fine(self, level: int, outmost: Set[Node], exist: Set[Node], iface: Set[Node],
                         lower: Set[Node], lower_outer: Set[Node]) -> None:
        """FindOutmostLayerForFineGrid: the interface nodes of `outmost` (level-(level-1)
        indices), and the next coarser level's seed nodes."""
        lo, hi = self.lo(level - 1), self.hi(level - 1)
        llo, lhi = tuple(c >> 1 for c in lo), tuple(c >> 1 for c in hi)
        for n in outmost:
            neg = [n[a] != lo[a] for a in range(self.d)]
            pos = [n[a] != hi[a] for a in range(self.d)]
            if all(q in exist for q in self.neighbors_dir(n, neg, pos)):
                continue
            iface.add(n)
            p = tuple(c >> 1 for c in n)
            lower.add(p)
            lneg = [p[a] != llo[a] for a in range(self.d)]
            lpos = [p[a] != lhi[a] for a in range(self.d)]
            if self.d == 2:
                lpos[1] = pos[1]     # the 2D restatement keeps the child's positive-y test
            up = lambda q: tuple(2 * c for c in q)
            lnb = self.neighbors_dir(p, lneg, lpos)
            if up(p) not in exist:
                lower_outer.add(p)
                for q in lnb:
                    if up(q) in exist:
                        lower.add(q)
            else:
                for q in lnb:
                    if up(q) not in exist:
                        lower.add(q)
                        lower_outer.add(q)

    # --------------------------------------------------------------------------------------
    def classify_cell(self, c: Node, iface: Set[Node], exist: Set[Node], layers) -> None:
        """IdentifyInterfaceForACell (first form) with IdentifyInterfaceNodeOnEdge and, in 3D,
        IdentifyInterfaceNodeDiagonal: the fine nodes of the coarse cell at corner c go to the
        inner (0), middle (1) or outer (2) layer by how many of their coarse corners lie on the
        coarse innermost layer."""
        corners = list(itertools.product((0, 1), repeat=self.d))
        if any(tuple(c[a] + k[a] for a in range(self.d)) not in exist for k in corners):
            return
        for t in itertools.product((0, 1, 2), repeat=self.d):
            odd = [a for a in range(self.d) if t[a] == 1]
            f = tuple(2 * c[a] + t[a] for a in range(self.d))
            if len(odd) == self.d and self.d == 3 or (self.d == 2 and len(odd) == 2):
                layers[1].add(f)          # cell centre (2D), body centre (3D): middle layer
                continue
            # the coarse nodes of the edge / face / corner the fine node sits on
            base = [t[a] // 2 if t[a] != 1 else 0 for a in range(self.d)]
            flags = []
            for k in itertools.product((0, 1), repeat=len(odd)):
                q = list(base)
                for j, a in enumerate(odd):
                    q[a] = k[j]
                flags.append(tuple(c[a] + q[a] for a in range(self.d)) in iface)
            if all(flags):
                layers[2].add(f)
            elif not any(flags):
                layers[0].add(f)
            else:
                layers[1].add(f)

    def overlap_layers(self, coarse_innermost: Set[Node], exist: Set[Node]):
        """FindOverlappingLayersBasedOnOutermostCoarse: fine layers [inner, mid, outmost] and the
        coarse outer layer."""
        layers = [set(), set(), set()]
        corner_offsets = list(itertools.product((0, -1), repeat=self.d))
        for n in coarse_innermost:
            for o in corner_offsets:
                self.classify_cell(tuple(n[a] + o[a] for a in range(self.d)), coarse_innermost, exist, layers)
        coarse_outer = {tuple(c >> 1 for c in f) for f in layers[0] if all(c % 2 == 0 for c in f)}
        return layers, coarse_outer


def generate(dims: int, cells: Sequence[int], max_level: int, geometries: Sequence[GeometrySpec]) -> Generated:
    """Generate every level's nodes, flags and interface layers for the given geometries."""
    G = _Gen(dims, cells, max_level)
    one_lower: List[Set[Node]] = [set() for _ in range(max_level + 1)]
    interfaces: List[Dict[int, _Interface]] = [dict() for _ in range(max_level + 1)]
    tracking: Dict[int, Set[Node]] = {}
    for k, geo in enumerate(geometries):
        if not (1 <= geo.level <= max_level):
            raise ValueError("geometry %d: level %d outside 1..%d" % (k, geo.level, max_level))
        verts = np.asarray(geo.vertices, float).reshape(-1, dims)
        tr = set()
        for v in verts:
            n = tuple(int(v[a] * (1 << geo.level) + _EPS) for a in range(dims))
            if not G.inside(n, geo.level):
                raise ValueError("geometry %d: vertex %s outside the domain" % (k, tuple(v)))
            tr.add(n)
        tracking[k] = tr
    innermost: Dict[int, Set[Node]] = {}
    outermost: Dict[int, Set[Node]] = {}
    for level in range(max_level, 0, -1):
        num_inner, num_outer = {}, {}
        for k, geo in enumerate(geometries):
            if geo.level != level:
                continue
            ineg, ipos, oneg, opos = G.extend_counts(geo, level)
            num_outer[k] = ([c - 1 for c in oneg], [c - 1 for c in opos])
            near = G.near_tracking(level, tracking[k])
            if geo.extend_type == "in_and_out":
                num_inner[k] = ([c - 1 for c in ineg], [c - 1 for c in ipos])
                origin = geo.flood_fill_origin
                if origin is None:
                    origin = np.asarray(geo.vertices, float).reshape(-1, dims).mean(0)
                outside, inside = G.flood_fill(level - 1, origin, near)
                innermost[k] = inside
                outermost[k] = outside
            else:
                outermost[k] = near
            one_lower[level] |= near
        inner_cur, outer_cur = {}, {}
        for kind, layer_map, cur, nums in (("inner", innermost, inner_cur, num_inner),
                                           ("outer", outermost, outer_cur, num_outer)):
            for k, start in layer_map.items():
                if k not in nums:
                    ineg, ipos, oneg, opos = G.extend_counts(geometries[k], level)
                    nums[k] = (ineg, ipos) if kind == "inner" else (oneg, opos)
                    if kind == "inner":
                        num_outer.setdefault(k, (oneg, opos))
                    else:
                        num_inner.setdefault(k, (ineg, ipos))
                cur[k] = G.extend(level, nums[k][0], nums[k][1], start, one_lower[level])
        low = level - 1
        for kind, cur, layer_map in (("inner", inner_cur, innermost), ("outer", outer_cur, outermost)):
            for k, out_layer in cur.items():
                layer_map[k] = set()
                info = interfaces[low].setdefault(k, _Interface())
                c2f = [set() for _ in range(_NUM_C2F)]
                if kind == "inner":
                    info.inner_c2f = c2f
                else:
                    info.outer_c2f = c2f
                G.outmost_for_fine(level, out_layer, one_lower[level], c2f[0], one_lower[low], layer_map[k])
    # ---- nodes of every level ----------------------------------------------------------------
    nodes: List[Dict[Node, int]] = [dict() for _ in range(max_level + 1)]

    def add(level: int, n: Node, flag: int) -> None:
        if n in nodes[level]:
            nodes[level][n] |= flag
        elif G.inside(n, level):
            nodes[level][n] = flag

    for level in range(max_level, 0, -1):
        low = level - 1
        for k, info in interfaces[low].items():
            fine = interfaces[level].setdefault(k, _Interface())
            for kind in ("inner", "outer"):
                c2f = getattr(info, kind + "_c2f")
                if not c2f:
                    continue
                f2c, coarse_outer = G.overlap_layers(c2f[0], one_lower[level])
                c2f[1] |= coarse_outer
                setattr(fine, kind + "_f2c", f2c)
                for il, layer in enumerate(f2c):
                    flag = (g.FLAG_F2C0 if il == _NUM_F2C - 1 else 0) | \
                        (g.FLAG_F2C_GHOST if il >= _NUM_F2C - _NUM_F2C_GHOST else 0)
                    for n in layer:
                        add(level, n, flag)
                ghost = _NUM_F2C_GHOST if kind == "inner" else _NUM_C2F_GHOST
                for il, layer in enumerate(c2f):
                    flag = (g.FLAG_C2F0 if il == _NUM_C2F - 1 else 0) | \
                        (g.FLAG_C2F_GHOST if il >= _NUM_C2F - ghost else 0)
                    for n in layer:
                        add(low, n, flag)
    occupied: Set[Node] = set()
    cell_offsets = list(itertools.product((0, 1), repeat=dims))
    fine_offsets = list(itertools.product((0, 1, 2), repeat=dims))
    for level in range(max_level, 0, -1):
        region = one_lower[level]
        sh = level - 1
        for n in region:
            if all(c % (1 << sh) == 0 for c in n):
                occupied.add(tuple(c >> sh for c in n))
            if all(tuple(n[a] + o[a] for a in range(dims)) in region for o in cell_offsets):
                for o in fine_offsets:
                    add(level, tuple(2 * n[a] + o[a] for a in range(dims)), 0)
    for info in interfaces[0].values():
        for c2f in (info.inner_c2f, info.outer_c2f):
            for layer in c2f or []:
                occupied -= layer
    lo, hi = G.lo(0), G.hi(0)
    for n in itertools.product(*[range(lo[a], hi[a] + 1) for a in range(dims)]):
        if n not in occupied:
            add(0, n, 0)
    # ---- into the solver's grid description ----------------------------------------------------
    levels = []
    for level in range(max_level + 1):
        pts = np.array(sorted(nodes[level]), dtype=np.int64).reshape(-1, dims)
        keys = g.morton_encode(dims, *[pts[:, a].astype(np.uint64) for a in range(dims)])
        order = np.argsort(keys)
        flags = np.array([nodes[level][tuple(p)] for p in pts], dtype=np.int32)
        lev = g.Level(keys[order], flags[order])
        merged: Dict[Tuple[int, int], Set[Node]] = {}
        for info in interfaces[level].values():
            for kind, attr in ((g.OUTER_F2C, "outer_f2c"), (g.OUTER_C2F, "outer_c2f"),
                               (g.INNER_F2C, "inner_f2c"), (g.INNER_C2F, "inner_c2f")):
                for il, layer in enumerate(getattr(info, attr) or []):
                    merged.setdefault((kind, il), set()).update(n for n in layer if n in nodes[level])
        for key, layer in merged.items():
            if layer:
                arr = np.array(sorted(layer), dtype=np.int64).reshape(-1, dims)
                lev.layers[key] = np.sort(g.morton_encode(dims, *[arr[:, a].astype(np.uint64) for a in range(dims)]))
        levels.append(lev)
    return Generated(g.Grid(dims, G.cells, levels), one_lower, tracking, interfaces)
