"""Throughput of closure training on a GPU box: Adam steps per second with the vector-unit (FP32)
and the matrix-core (BF16 operands) gradient kernels, on 2^20 synthetic samples.

    python tests/training_speed.py [--steps-scale 1.0]

One line per (form, batch) on stdout. This is a measurement script, not a test (pytest does not
collect it).
"""
import argparse
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from zllbm_amd import closure, hip  # noqa: E402
from tests import scenarios as sc  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps-scale", type=float, default=1.0)
    args = ap.parse_args()
    grid, kw = sc.periodic_tgv(3, 4, nu=0.02)
    s = hip.Solver(grid, **kw)
    rng = np.random.default_rng(0)
    n = 1 << 20
    X = rng.normal(size=(n, 9)).astype(np.float32)
    T = np.logaddexp(0.0, X @ rng.normal(size=9) * 0.5).astype(np.float32)
    p0 = closure.init_params(5)
    for form, name in ((hip.MLP_FP32, "fp32 vector units"), (hip.MLP_BF16_MFMA, "bf16 matrix cores")):
        for batch, steps in ((4096, 400), (65536, 100), (n, 20)):
            steps = max(2, int(steps * args.steps_scale))
            s.mlp_train(p0, X, T, 2, batch, 1e-3, form=form)
            t0 = time.perf_counter()
            _, h = s.mlp_train(p0, X, T, steps, batch, 1e-3, form=form)
            dt = time.perf_counter() - t0
            print("%-18s batch %7d: %.3f ms per Adam step, %.1f M samples/s, loss %.4e -> %.4e (%d steps; "
                  "includes one upload of the %d samples)"
                  % (name, batch, dt / steps * 1e3, batch * steps / dt / 1e6, h[0], h[-1], steps, n), flush=True)


if __name__ == "__main__":
    main()
