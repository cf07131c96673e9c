"""Train the learned sub-grid closure of the decaying-turbulence workload on the GPU.

A periodic box at the finest level's resolution (base * 2^(levels-1) cells, Smagorinsky Cs 0.05)
is started from the workload's Fourier modes and run in FP32, once at the finest level's
viscosity and once at twice it; snapshots of its velocity and density are box-filtered into
a-priori samples (zllbm_amd.closure.filtered_samples): by 2, 4 and 8 in the first run (the
resolutions and relaxation times of the three coarser levels) and by 2 in the second (the
finest level's relaxation time).  The
network is trained by the device kernels (zlh_mlp_train) on all snapshots but the last, which is
held out to compare its error with Smagorinsky's.

    python -m zllbm_amd.train_closure --out zllbm_amd/data/closure_turbulence.npz
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np


def log(*a):
    print(*a, flush=True)


def main(argv=None):
    from zllbm_amd import closure, grid as g, hip, workloads
    from zllbm_amd.fields import fourier_modes
    ap = argparse.ArgumentParser()
    ap.add_argument("--base", type=int, default=32)
    ap.add_argument("--levels", type=int, default=4)
    ap.add_argument("--snapshots", type=int, default=4)
    ap.add_argument("--every", type=int, default=256, help="fine steps between snapshots")
    ap.add_argument("--per-ratio", type=int, default=200000, help="samples kept per snapshot and ratio")
    ap.add_argument("--train-steps", type=int, default=4000)
    ap.add_argument("--form", choices=("fp32", "bf16-mfma"), default="fp32",
                    help="gradient kernel: vector units (FP32) or matrix cores (BF16 operands)")
    ap.add_argument("--out", required=True)
    args = ap.parse_args(argv)
    t0 = time.time()
    fine = args.base << (args.levels - 1)
    nu0 = 1e-4 * (1 << (args.levels - 1))              # background viscosity of the workload
    nu_fine = nu0 * (1 << (args.levels - 1))            # lattice viscosity at the finest level
    modes = fourier_modes((args.base,) * 3)
    grid = g.uniform_grid(3, (fine,) * 3)
    # run 1 at the finest level's viscosity: filters 2, 4, 8 give levels L-2, L-3, L-4;
    # run 2 at twice that viscosity: filter 2 gives the finest level's relaxation time
    runs = [(nu_fine, [r for r in (2, 4, 8) if (1 << (args.levels - 1)) >= r]), (2.0 * nu_fine, [2])]
    rng = np.random.default_rng(20261202)
    data = [[] for _ in range(args.snapshots)]
    for nu_run, ratios in runs:
        xyz = grid.coords(0).astype(np.float64) - 1.0
        u0 = modes.velocity(xyz / (1 << (args.levels - 1)))
        del xyz
        s = hip.Solver(grid, precision=hip.FP32_SHIFTED, collision=(hip.MRT,), nu0=nu_run,
                       bc=workloads.periodic_bc(3), les=hip.LES_SMAGORINSKY, cs=0.05).build()
        s.init_macro(0, np.ones(len(u0)), u0)
        del u0
        for snap in range(args.snapshots):
            s.time_marching(args.every)
            rho, u = s.macros(0)
            uf = closure.unique_field(grid, u)
            rf = closure.unique_field(grid, rho)
            del u, rho
            for r in ratios:
                tau0 = 3.0 * nu_run / r + 0.5
                x, t, pi = closure.filtered_samples(uf, r, tau0, rf)
                if len(t) > args.per_ratio:
                    keep = rng.choice(len(t), args.per_ratio, replace=False)
                    x, t, pi = x[keep], t[keep], pi[keep]
                data[snap].append((r, tau0, x, t, pi))
            log("nu %.4g snapshot %d (fine step %d): %s samples, %.0f s" % (
                nu_run, snap, (snap + 1) * args.every, [len(p[3]) for p in data[snap]], time.time() - t0))
        del s
    s = hip.Solver(grid, precision=hip.FP32_SHIFTED, nu0=nu_fine, bc=workloads.periodic_bc(3))
    train = [p for per in data[:-1] for p in per]
    x = np.concatenate([p[2] for p in train])
    t = np.concatenate([p[3] for p in train])
    form = hip.MLP_BF16_MFMA if args.form == "bf16-mfma" else hip.MLP_FP32
    model, hist = closure.fit_model(s, x, t, steps=args.train_steps, batch=4096, lr=3e-3, form=form)
    log("trained on %d samples: loss %.4f -> %.4f, %.0f s" % (len(t), hist[:50].mean(), hist[-50:].mean(),
                                                             time.time() - t0))
    s.mlp_set(model)
    report = {}
    for r, tau0, xh, th, pih in data[-1]:
        key = "tau0 %.4f" % tau0
        inside = np.all((xh >= model.fmin) & (xh <= model.fmax), axis=1)
        pred = s.mlp_eval(model.normalise(xh)) * model.nu_scale
        smag = closure.smagorinsky_nu_t(pih, tau0, 0.1)
        rms = lambda a: float(np.sqrt(np.mean((a[inside] - th[inside]) ** 2)))
        report[key] = {"filter": r, "in_range": float(inside.mean()), "rms_mlp": rms(pred),
                          "rms_smagorinsky_cs0.1": rms(smag), "rms_target": float(np.sqrt(np.mean(th ** 2)))}
    log(json.dumps(report))
    np.savez(args.out, params=model.params, mean=model.mean, std=model.std, fmin=model.fmin, fmax=model.fmax,
             nu_scale=np.float64(model.nu_scale), report=json.dumps(report), fine=fine)
    log("wrote", args.out)


if __name__ == "__main__":
    sys.exit(main())
