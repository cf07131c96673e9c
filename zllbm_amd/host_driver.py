"""Runs the C++ host driver (zllbm_amd/zllbm_lbm_driver: the plugin classes of
zllbm_amd/host/lbm_hip_solver.h under the AmrManager::TimeMarching loop) on a grid: writes its
grid, parameter, closure and initial-field files and reads back the populations and timing."""
import os
import re
import subprocess

import numpy as np

from zllbm_amd import grid as g

DRIVER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "zllbm_lbm_driver")
_DIR = {1: "xmin", 2: "xmax", 3: "ymin", 4: "ymax", 5: "zmin", 6: "zmax"}
_SCHEME = {1: "bounce_back", 2: "periodic", 3: "nee"}


def write_closure(path, model):
    """Closure file of the driver: parameters, mean, spread, minimum, maximum, output scale."""
    vals = [np.asarray(a, np.float64).ravel() for a in (model.params, model.mean, model.std, model.fmin, model.fmax)]
    with open(path, "w") as fh:
        for v in vals:
            fh.write(" ".join(repr(float(np.float32(x))) for x in v) + "\n")
        fh.write(repr(float(model.nu_scale)) + "\n")


def write_init(path, init):
    """Initial-field file of the driver: per level the node count, the densities and the velocity
    components node by node (key order), as little-endian doubles."""
    with open(path, "wb") as fh:
        for rho, u in init:
            rho = np.ascontiguousarray(rho, dtype="<f8")
            np.array([len(rho)], dtype="<i8").tofile(fh)
            rho.tofile(fh)
            np.ascontiguousarray(u, dtype="<f8").tofile(fh)


def write_ib(path, sets):
    """Immersed-boundary file of the driver: per marker set (level, coordinates (n, dims), areas,
    velocities (n, dims), stencil, stiffness) one record of int32 level, int32 stencil, double
    stiffness, int64 n, then the coordinates and velocities padded to 3 components and the areas."""
    with open(path, "wb") as fh:
        for level, xyz, area, vel, stencil, stiffness in sets:
            area = np.ascontiguousarray(area, dtype="<f8")
            n = len(area)

            def pad3(a):
                a = np.asarray(a, np.float64).reshape(n, -1)
                return np.ascontiguousarray(np.pad(a, ((0, 0), (0, 3 - a.shape[1]))), dtype="<f8")
            np.array([level, stencil], dtype="<i4").tofile(fh)
            np.array([stiffness], dtype="<f8").tofile(fh)
            np.array([n], dtype="<i8").tofile(fh)
            pad3(xyz).tofile(fh)
            area.tofile(fh)
            pad3(vel).tofile(fh)


def write_params(path, kw, dims, collision="srt", precision="fp64", closure_path=None, extra=()):
    """Parameter file of the driver from solver keywords (nu0, bc, les, cs, les_tau_mode)."""
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
        checkpoint_out=None, vtu_out=None, checkpoint_dir_in=None, checkpoint_dir_out=None, time_interp=False,
        graph_replay=None, init=None, warmup=None, ib=None, output=True, extra=(), timeout=300):
    """Runs the driver for `steps` background steps; returns the finished process and, when it
    succeeded and `output` is set, the post-stream populations of every level (key order).  `ib`
    lists marker sets (level, coordinates, areas, velocities, stencil, stiffness)."""
    gpath = os.path.join(workdir, "grid.bin")
    ppath = os.path.join(workdir, "params.txt")
    opath = os.path.join(workdir, "out.bin") if output else "-"
    cpath = os.path.join(workdir, "closure.txt")
    ipath = os.path.join(workdir, "init.bin")
    g.write_grid_file(grid, gpath)
    if model is not None:
        write_closure(cpath, model)
    extra = list(extra)
    for key, val in (("checkpoint_in", checkpoint_in), ("checkpoint_out", checkpoint_out), ("vtu_out", vtu_out),
                     ("checkpoint_dir_in", checkpoint_dir_in), ("checkpoint_dir_out", checkpoint_dir_out)):
        if val:
            extra.append("%s = %s" % (key, val))
    if time_interp:
        extra.append("time_interpolation = 1")
    if graph_replay is not None:
        extra.append("graph_replay = %d" % int(bool(graph_replay)))
    if warmup is not None:
        extra.append("warmup = %d" % warmup)
    if init is not None:
        write_init(ipath, init)
        extra.append("init = " + ipath)
    if ib:
        bpath = os.path.join(workdir, "ib.bin")
        write_ib(bpath, ib)
        extra.append("ib = " + bpath)
    write_params(ppath, kw, grid.dims, collision, precision, cpath, extra)
    proc = subprocess.run([DRIVER, gpath, ppath, str(steps), opath], capture_output=True, text=True, timeout=timeout)
    if proc.returncode != 0 or not output:
        return proc, None
    q = 9 if grid.dims == 2 else 19
    out = []
    with open(opath, "rb") as fh:
        for _ in grid.levels:
            n = int(np.fromfile(fh, dtype=np.int64, count=1)[0])
            out.append(np.fromfile(fh, dtype=np.float64, count=n * q).reshape(n, q))
    return proc, out


def ms_per_step(proc):
    """The driver's timing line: wall milliseconds per background step after the warm-up."""
    m = re.search(r"^ms_per_step (\S+)$", proc.stdout, re.M)
    return float(m.group(1)) if m else None
