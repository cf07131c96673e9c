"""Benchmark of the MI355X lattice Boltzmann time step (one JSON line on stdout).

Workloads (synthetic data, zllbm_amd.workloads):
  amr4    (default) D3Q19 MRT-LES decaying turbulence on 4 levels over a periodic 32^3 box
          (256^3-equivalent), refined where |vorticity| exceeds its 80th percentile; FP32 shifted
          populations with the learned closure (matrix-core network with FP16 operands, BF16
          with --mlp-operands bf16; Smagorinsky fallback).
          Compared in the same run with the FP64 Smagorinsky form on the same grid (MLUPS), and
          with the FP64 form of the same closure (velocity and density L2).
  tgv256  D3Q19 MRT + Smagorinsky (Cs 0.1), 256^3-cell periodic Taylor-Green vortex, FP32
          arithmetic on FP32 (default) or BF16 (--populations bf16) shifted populations.
  sphere5 D3Q19 MRT + learned closure, 5 levels over a 32x16x16 channel (512x256x256
          finest-equivalent), direct-forcing immersed sphere of 48 finest cells (7,239 markers).

value = whole-job MLUPS (lattice-node updates per second / 1e6; a level-l node counts 2^l
updates per background step), populations resident in HBM.
roofline: algorithmic bytes of the fused stream-collide kernel on the finest level, 2 * Q * s
bytes per node update (s = 4 for FP32 storage, 2 for BF16), over its average launch time from
HIP events on the library stream.
cpu_baseline: the FP64 CPU oracle (same algorithm, OpenMP over nodes) on a bounded sample.
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
METRIC = "MLUPS + % HBM roofline, D3Q19 MRT-LES AMR; FP32+learned vs FP64 plain"
DEFAULT_CLOSURE = os.path.join(ROOT, "zllbm_amd", "data", "closure_turbulence.npz")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


MLP_OPERANDS = "fp16"

# HBM bytes per launch of the roofline kernel, measured by rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE runs of this script on the same workload and options, with the counters scaled by
# what they read for a plain 1 GiB copy (bytes = 2.0 x FETCH_SIZE + 1.0 x WRITE_SIZE); the
# summaries, with their commands, are the named files under profiles/
PROFILED_TRAFFIC = {
    ("amr4", "mrt", "fp32", "fp16"): (580.5e6, "profiles/k_stream_collide_mlp_amr4.md"),
    ("tgv256", "mrt", "fp32", None): (2884.4e6, "profiles/k_stream_collide_tgv256.md"),
    ("tgv256", "mrt", "bf16", None): (1492.6e6, "profiles/k_stream_collide_bf16_tgv256.md"),
    ("sphere5", "mrt", "fp32", "fp16"): (82.8e6, "profiles/k_ib_sphere5.md"),
}


def make_solver(hip, W, precision, les, model=None, **extra):
    kw = dict(W.solver)
    kw.update(les=les, **extra)
    if les == hip.LES_MLP:
        kw.update(mlp_operands=hip.MLP_OPERANDS_FP16 if MLP_OPERANDS == "fp16" else hip.MLP_OPERANDS_BF16)
    s = hip.Solver(W.grid, precision=precision, **kw)
    ib = getattr(W, "ib", None)
    if ib is not None:
        s.set_ib(ib.level, ib.xyz, ib.area, np.zeros_like(ib.xyz), stencil_dis=2)
    s.build()
    if model is not None and kw.get("les") == hip.LES_MLP:
        s.mlp_set(model)
    for lvl, (rho, u) in enumerate(W.init):
        s.init_macro(lvl, rho, u)
    return s


def warm(s, warmup):
    """`warmup` untimed background steps that leave captured the graph the timed steps replay.
    Replays run two steps each and keep level 0's buffer parity; an odd count therefore starts
    with one direct step, so that the parity the timed region starts from is the one whose graph
    the warm-up replayed (a capture would otherwise land in the timed region)."""
    if warmup % 2:
        s.time_marching(1)
        warmup -= 1
    s.time_marching(warmup)


def cpp_drop_in(hip, W, model, mlp_operands, steps, warmup=4):
    """The same step through the C++ plugin classes under the AmrManager::TimeMarching loop (the
    example driver, a process of its own on the same GPU): wall ms per background step with whole
    background steps replayed as graphs (SetGraphReplay, the default) and with each substep
    launched by its own AdvancingAtCurrentTime call."""
    import shutil
    import tempfile
    from zllbm_amd import host_driver
    kw = dict(W.solver, les=hip.LES_MLP)
    res = {"loop": "AmrManager::TimeMarching over GridInfoLbmHip / SolverLbmHip (zllbm_lbm_driver)",
           "steps": steps, "warmup": warmup}
    if not os.path.exists(host_driver.DRIVER):
        res["skipped"] = "zllbm_amd/zllbm_lbm_driver is not built (__graft_entry__.build())"
        return res
    for replay in (True, False):
        d = tempfile.mkdtemp(prefix="zllbm_driver_")
        try:
            ib = getattr(W, "ib", None)
            sets = [(ib.level, ib.xyz, ib.area, np.zeros_like(ib.xyz), 2, 2.0)] if ib is not None else None
            proc, _ = host_driver.run(W.grid, kw, warmup + steps, d, collision="mrt", precision="fp32", model=model,
                                      graph_replay=replay, init=W.init, warmup=warmup, ib=sets, output=False,
                                      extra=("mlp_operands = " + mlp_operands,), timeout=600)
        finally:
            shutil.rmtree(d, ignore_errors=True)
        ms = host_driver.ms_per_step(proc) if proc.returncode == 0 else None
        if ms is None:
            # reported in the line, not fatal to it: the ctypes measurement above stands alone
            res["error"] = "host driver exit %d: %s" % (proc.returncode, proc.stderr[-500:])
            log("C++ drop-in failed:", res["error"])
            return res
        res["graph_replay_ms_per_step" if replay else "per_substep_ms_per_step"] = round(ms, 4)
        log("C++ drop-in (%s): %.3f ms per step" % ("graph replay" if replay else "per-substep launches", ms))
    return res


def timed(s, W, warmup, steps, kernel_steps=20):
    """Wall time of `steps` background steps (graph replays) after `warmup`, then the finest
    level's fused-kernel time from per-launch events over `kernel_steps` direct-launch steps."""
    warm(s, warmup)
    s.synchronize()
    t0 = time.perf_counter()
    s.time_marching(steps)
    s.synchronize()
    wall = time.perf_counter() - t0
    s.timing_enable(True)
    s.time_marching(kernel_steps)
    k_ms, k_n = s.timing_read(W.grid.max_level)
    s.timing_enable(False)
    return wall, k_ms, k_n, warmup + steps + kernel_steps


def state(s, W):
    rho, u = [], []
    for l in range(W.grid.max_level + 1):
        r, v = s.macros(l)
        rho.append(r)
        u.append(v)
    return np.concatenate(rho), np.concatenate(u)


def l2_curve(hip, W, les, model, steps, every):
    """Relative velocity L2 and density RMS between the FP32 and FP64 forms of the same model,
    stepped in lockstep from the same initial state, every `every` steps."""
    a = make_solver(hip, W, hip.FP32_SHIFTED, les, model)
    b = make_solver(hip, W, hip.FP64, les, model)
    out = []
    for t in range(every, steps + 1, every):
        a.time_marching(every)
        b.time_marching(every)
        ra, ua = state(a, W)
        rb, ub = state(b, W)
        out.append((t, float(np.sqrt(np.sum((ua - ub) ** 2) / np.sum(ub ** 2))), float(np.sqrt(np.mean((ra - rb) ** 2)))))
    return out


def cpu_baseline(model, seconds=15.0, threads=None):
    """FP64 oracle on a bounded sample of the same workload: 2 levels over a 24^3 box, on
    `threads` OpenMP threads (default: every host core)."""
    from oracle.zlbm_oracle import Oracle, lib
    from zllbm_amd import workloads
    threads = threads or os.cpu_count() or 1
    lib().zo_set_threads(threads)
    W = workloads.decaying_turbulence(24, 2, buffer=2, les=2 if model is not None else 1)
    o = Oracle(W.grid, **W.solver)
    if model is not None:
        o.set_mlp(model.params.astype(np.float64), model.mean, model.std, model.fmin, model.fmax, model.nu_scale)
    for l, (rho, u) in enumerate(W.init):
        o.init_macro(l, rho, u)
    o.time_marching(1)
    steps = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        o.time_marching(1)
        steps += 1
    dt = time.perf_counter() - t0
    return {"value": round(W.updates_per_step * steps / dt / 1e6, 4), "unit": "MLUPS", "cores": threads,
            "kind": "port",
            "sample": "FP64 oracle, D3Q19 MRT %s, decaying turbulence on 2 levels over a 24^3 box (%d updates "
                      "per step), %d steps in %.1f s" % ("+ learned closure" if model is not None else "+ Smagorinsky",
                                                         W.updates_per_step, steps, dt)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", default="amr4", choices=["amr4", "tgv256", "sphere5"])
    ap.add_argument("--collision", default="mrt", choices=["mrt", "srt"],
                    help="collision operator of every level (srt: to price the MRT moment transform)")
    ap.add_argument("--populations", default="fp32", choices=["fp32", "bf16"],
                    help="storage of the shifted populations (bf16: tgv256 only; no learned closure)")
    ap.add_argument("--closure", default=DEFAULT_CLOSURE, help="trained closure (zllbm_amd.train_closure)")
    ap.add_argument("--no-compare", action="store_true", help="skip the FP64 runs")
    ap.add_argument("--mlp-operands", default="fp16", choices=["fp16", "bf16"],
                    help="operands of the FP32 solver's matrix-core closure network")
    ap.add_argument("--l2-steps", type=int, default=1000)
    ap.add_argument("--l2-every", type=int, default=100)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--one-stream", action="store_true",
                    help="time marching on one stream (levels do not overlap; for per-kernel traces)")
    ap.add_argument("--no-graphs", action="store_true",
                    help="direct launches instead of replayed graphs of two background steps")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB mebibytes (for counter calibration)")
    args = ap.parse_args()
    global MLP_OPERANDS
    MLP_OPERANDS = args.mlp_operands
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
    if args.populations == "bf16" and args.workload != "tgv256":
        print("bench.py: BF16 populations do not run with the learned closure (use --workload tgv256)", file=sys.stderr)
        sys.exit(2)
    from zllbm_amd import closure, hip, workloads
    out = {"metric": METRIC, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic"}
    model = None
    if args.workload == "amr4":
        W = workloads.decaying_turbulence(32, 4)
        model = closure.ClosureModel.load(args.closure)
        les = hip.LES_MLP
    elif args.workload == "sphere5":
        W = workloads.sphere_amr()
        model = closure.ClosureModel.load(args.closure)
        les = hip.LES_MLP
    else:
        W = workloads.tgv(256)
        les = hip.LES_SMAGORINSKY
    if args.collision == "srt":
        W.solver["collision"] = (hip.SRT,)
    bf16 = args.populations == "bf16"
    log("workload:", W.name, "nodes per level", W.nodes)
    s = make_solver(hip, W, hip.BF16_SHIFTED if bf16 else hip.FP32_SHIFTED, les, model)
    if args.one_stream:
        s.level_streams_enable(False)
    if args.no_graphs:
        s.graphs_enable(False)
    log("aux nodes per level", [s.info(l)[1] for l in range(len(W.nodes))])
    wall, k_ms, k_n, _ = timed(s, W, args.warmup, args.steps)
    mlups = W.updates_per_step * args.steps / wall / 1e6
    q = 19
    bpu = 2 * q * (2 if bf16 else 4)
    finest = W.grid.max_level
    avg_s = k_ms / 1e3 / max(k_n, 1)
    achieved = W.nodes[finest] * bpu / avg_s / 1e9
    log("%s %s: %.3f ms per step, finest-level kernel %.4f ms x %d" % (
        "BF16" if bf16 else "FP32", "learned closure" if model is not None else "Smagorinsky",
        wall / args.steps * 1e3, avg_s * 1e3, k_n))
    kernel = ("k_stream_collide_mlp" if model is not None else "k_stream_collide") + " (level %d)" % finest
    traffic, traffic_src = PROFILED_TRAFFIC.get(
        (args.workload, args.collision, args.populations, args.mlp_operands if model is not None else None), (None, None))
    out.update({"value": round(mlups, 2), "unit": "MLUPS", "ms_per_step": round(wall / args.steps * 1e3, 4),
                "config": {"workload": W.name, "nodes_per_level": W.nodes, "updates_per_step": W.updates_per_step,
                           "collision": args.collision.upper(), "closure": "learned (%s-operand MFMA), Smagorinsky fallback" % args.mlp_operands.upper()
                           if model is not None else "Smagorinsky Cs 0.1",
                           "populations": "BF16 shifted" if bf16 else "FP32 shifted"},
                "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": round(achieved / HBM_PEAK_GBS, 4),
                             # whole step: every update at the kernel's algorithmic bytes, over the step time
                             "frac_step": round(mlups * 1e6 * bpu / (HBM_PEAK_GBS * 1e9), 4), "traffic": traffic,
                             "traffic_unit": "bytes per launch", "traffic_source": traffic_src, "kernel": kernel,
                             "avg_launch_ms": round(avg_s * 1e3, 4), "bytes_per_update": bpu,
                             "updates_per_launch": W.nodes[finest]}})
    if model is not None:
        fb = []
        for l in range(len(W.nodes)):
            fb.append(round(float(np.mean(s.closure_features(l)[2])), 4))
        out["config"]["closure_fallback_fraction"] = fb
    del s
    if not args.no_compare and model is not None:
        c_steps = max(20, args.steps // 2)
        s64 = make_solver(hip, W, hip.FP64, hip.LES_SMAGORINSKY)
        wall64, _, _, _ = timed(s64, W, 5, c_steps, kernel_steps=0)
        del s64
        mlups64 = W.updates_per_step * c_steps / wall64 / 1e6
        log("FP64 Smagorinsky: %.3f ms per step" % (wall64 / c_steps * 1e3))
        s32 = make_solver(hip, W, hip.FP32_SHIFTED, hip.LES_SMAGORINSKY)
        wall32, _, _, _ = timed(s32, W, 5, c_steps, kernel_steps=0)
        del s32
        log("FP32 Smagorinsky: %.3f ms per step" % (wall32 / c_steps * 1e3))
        # the same learned-closure step with the MRT transform on the matrix cores
        sx = make_solver(hip, W, hip.FP32_SHIFTED, les, model, moment_transform=hip.TRANSFORM_MFMA)
        wallx, kx_ms, kx_n, _ = timed(sx, W, 5, c_steps)
        del sx
        log("FP32 learned, MRT transform on the matrix cores: %.3f ms per step, finest-level kernel %.4f ms" % (
            wallx / c_steps * 1e3, kx_ms / max(kx_n, 1)))
        cpp = cpp_drop_in(hip, W, model, args.mlp_operands, c_steps)
        # accuracy: FP32 against FP64, same closure, in lockstep (the learned closure, and for
        # reference the plain Smagorinsky pair, whose difference is FP32 rounding alone)
        curve = l2_curve(hip, W, hip.LES_MLP, model, args.l2_steps, args.l2_every)
        curve_s = l2_curve(hip, W, hip.LES_SMAGORINSKY, None, args.l2_steps, args.l2_every)
        for (t, u1, r1), (_, u2, _) in zip(curve, curve_s):
            log("L2(u) after %4d steps: learned FP32 vs FP64 %.3e, Smagorinsky FP32 vs FP64 %.3e" % (t, u1, u2))
        out["comparison"] = {"fp64_smagorinsky_mlups": round(mlups64, 2),
                             "fp64_smagorinsky_ms_per_step": round(wall64 / c_steps * 1e3, 4),
                             "speedup_fp32_learned_vs_fp64_smagorinsky": round(mlups / mlups64, 3),
                             "fp32_smagorinsky_mlups": round(W.updates_per_step * c_steps / wall32 / 1e6, 2),
                             "mrt_transform_matrix_cores": {
                                 "ms_per_step": round(wallx / c_steps * 1e3, 4),
                                 "finest_kernel_ms": round(kx_ms / max(kx_n, 1), 4),
                                 "vector_units_finest_kernel_ms": round(avg_s * 1e3, 4),
                                 "kernel": "k_stream_collide_mlp (level %d), v_mfma_f32_32x32x2f32 transform" % finest},
                             "cpp_drop_in": cpp,
                             "l2_u_vs_fp64_mean": float(np.mean([c[1] for c in curve])),
                             "l2_u_vs_fp64_final": curve[-1][1], "l2_rho_vs_fp64_final": curve[-1][2],
                             "l2_u_curve": [[t, round(u, 8)] for t, u, _ in curve],
                             "l2_u_curve_smagorinsky_fp32_vs_fp64": [[t, round(u, 8)] for t, u, _ in curve_s],
                             "l2_reference": "FP64 form of the same closure on the GPU (pinned to the CPU "
                                             "oracle by the parity tests), stepped in lockstep from the same "
                                             "initial state"}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(model)
        one = cpu_baseline(model, seconds=10.0, threads=1)
        out["cpu_baseline"]["one_thread"] = {"value": one["value"], "unit": one["unit"], "cores": 1,
                                             "sample": one["sample"]}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
