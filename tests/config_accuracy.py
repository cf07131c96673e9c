"""Velocity and density error of each benchmark configuration, on a GPU box.

    python tests/config_accuracy.py [--configs 1,2,3,4,5]

For every configuration:
* the FP32 shifted-population solver against the FP64 solver, both on the GPU, stepped in lockstep
  from the same initial state at the configuration's full size; likewise the BF16-storage solver
  for the configurations without the learned closure (1 to 3);
* the FP64 GPU solver against the FP64 oracle on the CPU (16 threads) where the oracle finishes
  in a minute or so: configurations 1 and 2 at full size and length, configuration 3's TGV at
  64^3 (same u0, Re and noise) for 100 steps, configuration 4's closure on a 2-level 24^3 box
  for 20 steps, and configuration 5 at full size (5 levels, IB sphere, learned closure) for 10
  background steps.

Throughput: MLUPS (node updates of all levels per background step) of the FP32 and FP64 forms,
from graph-replayed background steps with the populations resident on the GPU.

Errors: velocity relative L2, sqrt(sum |u_a - u_b|^2 / sum |u_b|^2) over the nodes of all levels,
and density RMS difference. One JSON line per configuration on stdout; progress on stderr.
This is a measurement script, not a test (pytest does not collect it).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from zllbm_amd import closure, hip, workloads  # noqa: E402

MODEL = os.path.join(ROOT, "zllbm_amd", "data", "closure_turbulence.npz")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def gpu_solver(W, precision, model=None):
    kw = dict(W.solver)
    if kw.get("les") == hip.LES_MLP:
        kw["mlp_operands"] = hip.MLP_OPERANDS_FP16
    s = hip.Solver(W.grid, precision=precision, **kw)
    ib = getattr(W, "ib", None)
    if ib is not None:
        s.set_ib(ib.level, ib.xyz, ib.area, np.zeros_like(ib.xyz), stencil_dis=2)
    s.build()
    if kw.get("les") == hip.LES_MLP:
        s.mlp_set(model)
    for lvl, (rho, u) in enumerate(W.init):
        s.init_macro(lvl, rho, u)
    return s


def oracle_solver(W, model=None):
    from oracle.zlbm_oracle import Oracle
    kw = dict(W.solver)
    o = Oracle(W.grid, **kw)
    ib = getattr(W, "ib", None)
    if ib is not None:
        o.set_ib(ib.level, ib.xyz, ib.area, np.zeros_like(ib.xyz), stencil_dis=2)
    if kw.get("les") == hip.LES_MLP:
        o.set_mlp(model.params.astype(np.float64), model.mean, model.std, model.fmin, model.fmax, model.nu_scale)
    for lvl, (rho, u) in enumerate(W.init):
        o.init_macro(lvl, rho, u)
    return o


def macros(f, dims):
    """Density and velocity of post-stream populations [n][Q] (FP64, unshifted), the same
    moments for both sides."""
    from oracle.zlbm_oracle import lattice
    cx, cy, cz, _, _ = (np.asarray(v, np.float64) for v in lattice(dims))
    rho = f.sum(1)
    u = np.stack([f @ cx, f @ cy] + ([f @ cz] if dims == 3 else []), axis=1) / rho[:, None]
    return rho, u


def state(sim, W):
    out = [macros(sim.f(l), W.grid.dims) for l in range(W.grid.max_level + 1)]
    return np.concatenate([r for r, _ in out]), np.concatenate([u for _, u in out])


def errors(a, b):
    (ra, ua), (rb, ub) = a, b
    ua, ub = ua[:, :ub.shape[1]], ub[:, :ua.shape[1]]
    return float(np.sqrt(np.sum((ua - ub) ** 2) / np.sum(ub ** 2))), float(np.sqrt(np.mean((ra - rb) ** 2)))


def fp32_vs_fp64(W, steps, every, model=None, precision=hip.FP32_SHIFTED):
    a = gpu_solver(W, precision, model)
    b = gpu_solver(W, hip.FP64, model)
    name = "BF16" if precision == hip.BF16_SHIFTED else "FP32"
    curve = []
    for t in range(every, steps + 1, every):
        a.time_marching(every)
        b.time_marching(every)
        lu, lr = errors(state(a, W), state(b, W))
        curve.append((t, lu, lr))
        log("  %s vs FP64 after %6d steps: L2(u) %.3e  RMS(rho) %.3e" % (name, t, lu, lr))
    return curve


def gpu_vs_oracle(W, steps, model=None):
    from oracle.zlbm_oracle import lib as oracle_lib
    oracle_lib().zo_set_threads(min(16, os.cpu_count() or 1))
    s = gpu_solver(W, hip.FP64, model)
    o = oracle_solver(W, model)
    t0 = time.perf_counter()
    o.time_marching(steps)
    t_cpu = time.perf_counter() - t0
    s.time_marching(steps)
    lu, lr = errors(state(s, W), state(o, W))
    log("  FP64 GPU vs FP64 oracle after %d steps: L2(u) %.3e  RMS(rho) %.3e (oracle %.1f s)" % (steps, lu, lr, t_cpu))
    return {"steps": steps, "l2_u": lu, "rms_rho": lr, "oracle_seconds": round(t_cpu, 1)}


def throughput(W, steps, model=None):
    out = {}
    forms = [("fp32", hip.FP32_SHIFTED), ("fp64", hip.FP64)]
    if W.solver.get("les") != hip.LES_MLP:
        forms.append(("bf16", hip.BF16_SHIFTED))
    for name, prec in forms:
        s = gpu_solver(W, prec, model)
        s.time_marching(10)
        s.synchronize()
        t0 = time.perf_counter()
        s.time_marching(steps)
        s.synchronize()
        wall = time.perf_counter() - t0
        out[name + "_mlups"] = round(W.updates_per_step * steps / wall / 1e6, 1)
        out[name + "_ms_per_step"] = round(wall / steps * 1e3, 4)
        del s
    log("  MLUPS " + ", ".join("%s %.1f" % (k[:4].upper(), v) for k, v in out.items() if k.endswith("mlups")))
    return out


def summary(curve):
    return {"steps": curve[-1][0], "l2_u_final": curve[-1][1], "rms_rho_final": curve[-1][2],
            "l2_u_mean": float(np.mean([c[1] for c in curve])),
            "l2_u_curve": [[t, round(u, 9)] for t, u, _ in curve]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--configs", default="1,2,3,4,5")
    args = ap.parse_args()
    want = {int(c) for c in args.configs.split(",")}
    model = closure.ClosureModel.load(MODEL)
    if 1 in want:
        W = workloads.cavity(128)
        log("config 1:", W.name, W.nodes)
        rec = {"config": 1, "workload": W.name, "throughput": throughput(W, 2000),
               "fp32_vs_fp64": summary(fp32_vs_fp64(W, 10000, 2000)),
               "bf16_vs_fp64": summary(fp32_vs_fp64(W, 10000, 2000, precision=hip.BF16_SHIFTED)),
               "fp64_vs_oracle": gpu_vs_oracle(W, 10000)}
        print(json.dumps(rec), flush=True)
    if 2 in want:
        W = workloads.cylinder_amr()
        log("config 2:", W.name, W.nodes)
        rec = {"config": 2, "workload": W.name, "throughput": throughput(W, 400),
               "fp32_vs_fp64": summary(fp32_vs_fp64(W, 2000, 500)),
               "bf16_vs_fp64": summary(fp32_vs_fp64(W, 2000, 500, precision=hip.BF16_SHIFTED)),
               "fp64_vs_oracle": gpu_vs_oracle(W, 2000)}
        print(json.dumps(rec), flush=True)
    if 3 in want:
        W = workloads.tgv(256)
        log("config 3:", W.name, W.nodes)
        rec = {"config": 3, "workload": W.name, "throughput": throughput(W, 200),
               "fp32_vs_fp64": summary(fp32_vs_fp64(W, 2000, 500)),
               "bf16_vs_fp64": summary(fp32_vs_fp64(W, 2000, 500, precision=hip.BF16_SHIFTED))}
        small = workloads.tgv(64)
        rec["fp64_vs_oracle"] = dict(gpu_vs_oracle(small, 100), workload=small.name)
        print(json.dumps(rec), flush=True)
    if 4 in want:
        W = workloads.decaying_turbulence(32, 4)
        W.solver["les"] = hip.LES_MLP
        log("config 4:", W.name, W.nodes)
        small = workloads.decaying_turbulence(24, 2)
        small.solver["les"] = hip.LES_MLP
        rec = {"config": 4, "workload": W.name, "throughput": throughput(W, 200, model),
               "fp32_vs_fp64": summary(fp32_vs_fp64(W, 1000, 250, model)),
               "fp64_vs_oracle": dict(gpu_vs_oracle(small, 20, model), workload=small.name)}
        print(json.dumps(rec), flush=True)
    if 5 in want:
        W = workloads.sphere_amr()
        log("config 5:", W.name, W.nodes)
        rec = {"config": 5, "workload": W.name, "throughput": throughput(W, 200, model),
               "fp32_vs_fp64": summary(fp32_vs_fp64(W, 400, 100, model)),
               "fp64_vs_oracle": gpu_vs_oracle(W, 10, model)}
        print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
