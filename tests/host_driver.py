"""Helpers that run the C++ host driver (zllbm_amd/zllbm_lbm_driver) on a grid."""
import os
import subprocess

import numpy as np

from zllbm_amd import grid as g

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "zllbm_amd", "zllbm_lbm_driver")
_DIR = {1: "xmin", 2: "xmax", 3: "ymin", 4: "ymax", 5: "zmin", 6: "zmax"}
_SCHEME = {1: "bounce_back", 2: "periodic", 3: "nee"}


def write_closure(path, model):
    """Closure file of the driver: parameters, mean, spread, minimum, maximum, output scale."""
    vals = [np.asarray(a, np.float64).ravel() for a in (model.params, model.mean, model.std, model.fmin, model.fmax)]
    with open(path, "w") as fh:
        for v in vals:
            fh.write(" ".join(repr(float(np.float32(x))) for x in v) + "\n")
        fh.write(repr(float(model.nu_scale)) + "\n")


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


def write_params(path, kw, dims, collision="srt", precision="fp64", closure_path=None, extra=()):
    lines = ["viscosity = %r" % kw["nu0"], "rho = 1.0",
             "velocity = " + " ".join(["0.0"] * dims), "collision = " + collision, "precision = " + precision]
    if kw.get("les") == 2:
        lines += ["les = learned", "cs = %r" % kw.get("cs", 0.1), "closure = " + closure_path]
    elif kw.get("les"):
        lines += ["les = smagorinsky", "cs = %r" % kw.get("cs", 0.1), "les_tau_mode = %d" % kw.get("les_tau_mode", 1)]
    for d, (scheme, val) in kw["bc"].items():
        lines.append("bc.%s = %s %s" % (_DIR[d], _SCHEME[scheme], " ".join(repr(float(v)) for v in val)))
    lines += list(extra)
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def run(grid, kw, steps, workdir, collision="srt", precision="fp64", model=None, checkpoint_in=None,
        checkpoint_out=None, vtu_out=None):
    gpath = os.path.join(workdir, "grid.bin")
    ppath = os.path.join(workdir, "params.txt")
    opath = os.path.join(workdir, "out.bin")
    cpath = os.path.join(workdir, "closure.txt")
    g.write_grid_file(grid, gpath)
    if model is not None:
        write_closure(cpath, model)
    extra = []
    if checkpoint_in:
        extra.append("checkpoint_in = " + checkpoint_in)
    if checkpoint_out:
        extra.append("checkpoint_out = " + checkpoint_out)
    if vtu_out:
        extra.append("vtu_out = " + vtu_out)
    write_params(ppath, kw, grid.dims, collision, precision, cpath, extra)
    proc = subprocess.run([DRIVER, gpath, ppath, str(steps), opath], capture_output=True, text=True, timeout=300)
    if proc.returncode != 0:
        return proc, None
    q = 9 if grid.dims == 2 else 19
    out = []
    with open(opath, "rb") as fh:
        for _ in grid.levels:
            n = int(np.fromfile(fh, dtype=np.int64, count=1)[0])
            out.append(np.fromfile(fh, dtype=np.float64, count=n * q).reshape(n, q))
    return proc, out
