"""The factored D3Q19 MRT transforms of the device kernels (zllbm_amd/csrc/mrt_d3q19.hpp) against
the integer moment matrix of the same lattice (Lattice<3>::moment, zllbm_amd/csrc/lattice.h),
compiled for the host with g++: m = M f and f = M^T y for random vectors, to FP64 rounding."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

HARNESS = r"""
#define __host__
#define __device__
#define __forceinline__ inline
#include <cmath>
#include <cstdio>
#include <random>
#include "lattice.h"
#define ZLH_HD inline
#include "mrt_d3q19.hpp"
using namespace zlh;
int main() {
    std::mt19937_64 g(7);
    std::uniform_real_distribution<double> U(-1, 1);
    double fwd = 0, back = 0, scale = 0;
    for (int t = 0; t < 2000; ++t) {
        double f[19], m[19], y[19], fb[19];
        for (int i = 0; i < 19; ++i) { f[i] = U(g); y[i] = U(g); }
        if (t % 3 == 0) f[t % 19] = 0.;   /* include zero entries */
        mrt19_moments(f, m);
        mrt19_populations(y, fb);
        for (int k = 0; k < 19; ++k) {
            double r = 0;
            for (int i = 0; i < 19; ++i) r += Lattice<3>::moment(k, i) * f[i];
            fwd = std::fmax(fwd, std::fabs(r - m[k]));
            scale = std::fmax(scale, std::fabs(r));
        }
        for (int i = 0; i < 19; ++i) {
            double r = 0;
            for (int k = 0; k < 19; ++k) r += Lattice<3>::moment(k, i) * y[k];
            back = std::fmax(back, std::fabs(r - fb[i]));
        }
    }
    std::printf("%.6e %.6e %.6e\n", fwd, back, scale);
    return 0;
}
"""


def test_factored_transforms_equal_the_moment_matrix(tmp_path):
    src = tmp_path / "mrt_check.cpp"
    exe = tmp_path / "mrt_check"
    src.write_text(HARNESS)
    subprocess.run(["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "zllbm_amd", "csrc"), "-o", str(exe),
                    str(src)], check=True, capture_output=True, text=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    fwd, back, scale = map(float, out)
    assert scale > 10.0                       # rows with coefficients up to 30 were exercised
    assert fwd < 1e-13 * scale
    assert back < 1e-13 * scale
