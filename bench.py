"""Benchmark of the MI355X lattice Boltzmann time step (one JSON line on stdout).

Workloads (synthetic data, sizes from the headline configurations):
  tgv256  D3Q19 MRT + Smagorinsky (Cs 0.1), 256^3-cell periodic Taylor-Green vortex (u0 0.05)
          with seeded Gaussian noise 1e-3, FP32 shifted populations.
  amr4    D3Q19 MRT + Smagorinsky on a 4-level box-refined periodic grid (32^3 background,
          finest level 256^3-equivalent), FP32 shifted populations.

value = whole-job MLUPS (lattice-node updates per second / 1e6, every node of every level
counted once per update it receives), with populations resident in HBM.
roofline: algorithmic bytes of the fused stream-collide kernel, 2 * Q * 4 bytes per node update,
divided by its average launch time measured with HIP events on the library stream.
cpu_baseline: the FP64 CPU oracle (same algorithm) on a bounded sample on the host cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_workload(name, hip, g):
    from tests import scenarios as sc
    if name == "tgv256":
        n = 256
        grid = g.uniform_grid(3, (n, n, n))
        nu = 0.05 * (n / (2 * np.pi)) / 1600.0     # Re = u0 / (k nu) = 1600
        kw = dict(nu0=nu, bc=sc.all_bc(3, sc.PER), collision=(hip.MRT,), les=hip.LES_SMAGORINSKY,
                  les_tau_mode=hip.LES_TAU_PHYSICAL, cs=0.1, precision=hip.FP32_SHIFTED)
        init = sc.tgv_fields(grid, 0.05, noise=1e-3)
        desc = "D3Q19 MRT-LES (Cs 0.1) TGV 256^3 periodic, u0 0.05 + 1e-3 noise, FP32 shifted"
        return grid, kw, [init], desc
    raise SystemExit("unknown workload " + name)


def cpu_baseline(seconds=15.0):
    from oracle.zlbm_oracle import Oracle, lib
    from zllbm_amd import grid as g
    from tests import scenarios as sc
    threads = os.cpu_count() or 1
    lib().zo_set_threads(threads)
    n = 32
    grid = g.uniform_grid(3, (n, n, n))
    kw = dict(nu0=0.05 * (256 / (2 * np.pi)) / 1600.0, bc=sc.all_bc(3, sc.PER), collision=(1,), les=1,
              les_tau_mode=1, cs=0.1)
    o = Oracle(grid, **kw)
    rho, u = sc.tgv_fields(grid, 0.05, noise=1e-3)
    o.init_macro(0, rho, u)
    o.time_marching(1)
    steps = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        o.time_marching(1)
        steps += 1
    dt = time.perf_counter() - t0
    mlups = grid.levels[0].n * steps / dt / 1e6
    return {"value": round(mlups, 4), "unit": "MLUPS", "cores": threads, "kind": "port",
            "sample": "FP64 oracle, D3Q19 MRT-LES TGV 32^3 periodic (35,937 nodes), %d steps in %.1f s" % (steps, dt)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", default="tgv256")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB mebibytes (for counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "launches": 5}))
        return
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)
    from zllbm_amd import grid as g, hip
    grid, kw, inits, desc = build_workload(args.workload, hip, g)
    log("building", desc)
    s = hip.Solver(grid, **kw).build()
    for lvl, (rho, u) in enumerate(inits):
        s.init_macro(lvl, rho, u)
    nodes = [lev.n for lev in grid.levels]
    updates_per_step = sum(n << l for l, n in enumerate(nodes))
    log("nodes per level", nodes, "aux", [s.info(l)[1] for l in range(len(nodes))])
    s.time_marching(args.warmup)
    s.synchronize()
    s.timing_enable(True)
    t0 = time.perf_counter()
    s.time_marching(args.steps)
    s.synchronize()
    t1 = time.perf_counter()
    k_ms, k_n = s.timing_read()
    s.timing_enable(False)
    wall = t1 - t0
    mlups = updates_per_step * args.steps / wall / 1e6
    q = 19 if grid.dims == 3 else 9
    bytes_per_update = 2 * q * (4 if kw.get("precision", 0) == hip.FP32_SHIFTED else 8)
    avg_launch_s = k_ms / 1e3 / max(k_n, 1)
    # one launch covers one level substep; the finest-level launch dominates for AMR
    launch_nodes = updates_per_step * args.steps / max(k_n, 1)
    achieved = launch_nodes * bytes_per_update / avg_launch_s / 1e9
    rho, u = s.macros(0)
    finite = bool(np.all(np.isfinite(rho)))
    log("wall %.3f s, kernel %.3f ms x %d, finite=%s" % (wall, k_ms, k_n, finite))
    out = {
        "metric": "MLUPS + % HBM roofline, D3Q19 MRT-LES AMR; FP32+learned vs FP64 plain",
        "value": round(mlups, 2), "unit": "MLUPS", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32" if kw.get("precision") == hip.FP32_SHIFTED else "fp64",
        "data": "synthetic",
        "config": {"workload": desc, "nodes_per_level": nodes, "updates_per_step": updates_per_step},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_stream_collide", "avg_launch_ms": round(avg_launch_s * 1e3, 4),
                     "bytes_per_update": bytes_per_update},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
