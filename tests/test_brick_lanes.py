"""The pull's lane arithmetic of the device kernels (zllbm_amd/csrc/lattice.h): stepping one
node along an axis inside a brick on the interleaved (Morton) lane bits, without
de-interleaving, against de-interleave / step modulo the edge / re-interleave, for every lane,
axis and step of the 2D (8x8) and 3D (4x4x4) bricks; compiled for the host with g++."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

HARNESS = r"""
#define __host__
#define __device__
#define __forceinline__ inline
#include <cstdio>
#include "lattice.h"
using namespace zlh;
template <int D> int check() {
    using B = Brick<D>;
    constexpr int E = B::edge;
    int bad = 0;
    for (int lane = 0; lane < 64; ++lane)
        for (int a = 0; a < D; ++a) {
            const uint32_t m4 = lane_axis_mask4<D>(a);
            const uint32_t own4 = 4u * (uint32_t)lane & m4;
            int c[3] = {B::lane_coord(lane, 0), B::lane_coord(lane, 1), D == 3 ? B::lane_coord(lane, 2) : 0};
            for (int step = -1; step <= 1; step += 2) {
                int s[3] = {0, 0, 0};
                s[a] = (c[a] + step + E) % E;
                const uint32_t want = 4u * (uint32_t)B::lane_of(s[0], s[1], s[2]);
                if (lane_axis_step4(own4, m4, step) != want) ++bad;
            }
            /* own bits of the axis: the lane with the other coordinates zeroed */
            int o[3] = {0, 0, 0};
            o[a] = c[a];
            if (own4 != 4u * (uint32_t)B::lane_of(o[0], o[1], o[2])) ++bad;
            if ((own4 == 0u) != (c[a] == 0) || (own4 == m4) != (c[a] == E - 1)) ++bad;
        }
    return bad;
}
int main() { std::printf("%d %d\n", check<2>(), check<3>()); return 0; }
"""


def test_lane_steps_match_the_morton_layout(tmp_path):
    src = tmp_path / "lanes.cpp"
    exe = tmp_path / "lanes"
    src.write_text(HARNESS)
    subprocess.run(["g++", "-O1", "-std=c++17", "-I" + os.path.join(ROOT, "zllbm_amd", "csrc"), "-o", str(exe),
                    str(src)], check=True, capture_output=True, text=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    assert out == ["0", "0"]
