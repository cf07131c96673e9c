"""Helpers for the C++ host driver's tests: the reference's checkpoint record order and files by
its own loops; running the driver is zllbm_amd.host_driver's."""
import numpy as np

from zllbm_amd import grid as g
from zllbm_amd.host_driver import DRIVER, run, ms_per_step, write_closure, write_params  # noqa: F401


def read_checkpoint(path, q):
    """Records of a grid-node checkpoint: header (uint32 count, int32 payload bytes), then per
    node uint64 key, int32 level, int32 flag, q doubles (and, with forces, more doubles)."""
    with open(path, "rb") as fh:
        count = int(np.fromfile(fh, dtype=np.uint32, count=1)[0])
        payload = int(np.fromfile(fh, dtype=np.int32, count=1)[0])
        rec = np.dtype([("key", "<u8"), ("level", "<i4"), ("flag", "<i4"), ("f", "<f8", (q,)),
                        ("force", "<f8", ((payload - 4 - 8 * q) // 8,))])
        data = np.fromfile(fh, dtype=rec, count=count)
        rest = fh.read()
    return count, payload, data, len(rest)


def reference_record_order(grid):
    """(level, key) of every node in the order the reference's checkpoint writer visits them, by
    its own loops: the background cells that hold a node of any level (the level-l key moved to
    the background level), by Morton code ascending; in each, the levels in turn; in a level, the
    2^l per axis nodes of the cell with z outermost and x innermost, where present."""
    dims = grid.dims
    present = [set(int(k) for k in lev.keys) for lev in grid.levels]
    cells = set()
    for l, lev in enumerate(grid.levels):
        cells.update(int(k) >> (dims * l) for k in lev.keys)
    out = []
    for bkey in sorted(cells):
        b = g.morton_decode(dims, np.array([bkey], np.uint64))[0]
        for l in range(grid.max_level + 1):
            m = 1 << l
            for z in (range(m) if dims == 3 else [0]):
                for y in range(m):
                    for x in range(m):
                        c = [int(b[0]) * m + x, int(b[1]) * m + y] + ([int(b[2]) * m + z] if dims == 3 else [])
                        k = int(g.morton_encode(dims, *[np.array([v], np.uint64) for v in c])[0])
                        if k in present[l]:
                            out.append((l, k))
    return out


# layer kinds in the order the reference's interface_nodes writer takes them, with their
# indicators (kIndictor*, mpi/mpi_manager.h:615-616)
KIND_ORDER = ((1, g.INNER_F2C), (2, g.OUTER_F2C), (3, g.INNER_C2F), (4, g.OUTER_C2F))


def layer_vectors(lev):
    """The driver's InterfaceLayerInfo of one level: per kind the list of layers (key sets), as
    long as the largest layer index the grid names."""
    per = {}
    for (kind, layer), keys in lev.layers.items():
        per.setdefault(kind, {})[layer] = set(int(k) for k in keys)
    return {kind: [d.get(i, set()) for i in range(max(d) + 1)] for kind, d in per.items()}


def _cells(grid):
    dims = grid.dims
    cells = {}
    for l, lev in enumerate(grid.levels):
        for k in lev.keys:
            c = cells.setdefault(int(k) >> (dims * l), [0, 0, 0, 0])
            c[0] |= 1 << l
            c[1] += 1
            c[3] += 1 << l
        for kind, vec in layer_vectors(lev).items():
            for layer in vec:
                for k in layer:
                    cells[k >> (dims * l)][2] += 1
    return cells


def reference_loadings(grid):
    """Bytes of the reference's loadings file for one rank, by its own loops
    (ComputeComputationalLoadOnEachRank + WriteCheckPointNodesAtWhichLevels): header {uint32 total
    load, uint32 cells, int32 max level}, then per background cell by code {uint64 code, level
    bits, int32 nodes, int32 interface entries, int32 load}; a level-l node costs 2^l."""
    import struct
    cells = _cells(grid)
    nbytes = (grid.max_level + 1 + 7) // 8
    out = struct.pack("<IIi", sum(c[3] for c in cells.values()), len(cells), grid.max_level)
    for code in sorted(cells):
        bits, nodes, iface, load = cells[code]
        out += struct.pack("<Q", code) + bits.to_bytes(nbytes, "little") + struct.pack("<iii", nodes, iface, load)
    return out


def reference_interface_nodes(grid):
    """Bytes of the reference's interface_nodes file for one rank (WriteCheckPointInterfaceNodes):
    the index table of every layer of the one geometry interface (criterion 1, count 0) of every
    level that has layers, then per background cell, level and node of the cell (z, y, x) one
    {uint64 key, int32 index} per layer kind whose layers hold the node, naming the first."""
    import struct
    dims = grid.dims
    vecs = [layer_vectors(lev) for lev in grid.levels]
    index, table = {}, b""
    for l, vec in enumerate(vecs):
        if not grid.levels[l].layers:
            continue
        for ind, kind in KIND_ORDER:
            for layer in range(len(vec.get(kind, []))):
                index[(l, ind, layer)] = len(index)
                table += struct.pack("<iiiii", l, 1, 0, ind, layer)
    out = struct.pack("<i", len(index)) + table
    cells = _cells(grid)
    for code in sorted(cells):
        b = g.morton_decode(dims, np.array([code], np.uint64))[0]
        for l in range(grid.max_level + 1):
            if not cells[code][0] >> l & 1:
                continue
            m = 1 << l
            for z in (range(m) if dims == 3 else [0]):
                for y in range(m):
                    for x in range(m):
                        c = [int(b[0]) * m + x, int(b[1]) * m + y] + ([int(b[2]) * m + z] if dims == 3 else [])
                        k = int(g.morton_encode(dims, *[np.array([v], np.uint64) for v in c])[0])
                        for ind, kind in KIND_ORDER:
                            for layer, keys in enumerate(vecs[l].get(kind, [])):
                                if k in keys:
                                    out += struct.pack("<Qi", k, index[(l, ind, layer)])
                                    break
    return out
