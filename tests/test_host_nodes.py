"""Host-side node information of the C++ drop-in (no device call): NodeInfoCoarse2fine /
NodeInfoFine2Coarse on node mirrors, against the reference's closed forms (feq of the source's
macros plus the tau-ratio-scaled non-equilibrium part, lbm_grid_interface.cpp:570-610,
lbm_collision_models.cpp:142-151, 244-267), and the node record sizes
(lbm_grid_interface.cpp:737-749, lbm_grid_interface_mpi.cpp:77-90)."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r"""
#include <cstdio>
#include <memory>
#include "lbm_hip_solver.h"
using namespace rootproject;
using namespace rootproject::lbmproject;
int main(int argc, char **argv) {
    const int dims = argv[1][0] - '0';
    std::shared_ptr<amrproject::SolverInterface> solver =
        dims == 2 ? SolverCreatorLbmD2Q9Hip().CreateSolver() : SolverCreatorLbmD3Q19Hip().CreateSolver();
    auto *lbm = dynamic_cast<SolverLbmHip *>(solver.get());
    lbm->SetDefaultViscosity(0.013);
    lbm->SetNumForces(dims);
    std::vector<std::shared_ptr<amrproject::GridInfoInterface>> gi;
    for (int l = 0; l < 3; ++l) {
        gi.push_back(solver->ptr_grid_info_creator_->CreateGridInfo());
        gi[l]->SetGridLevel(l);
        gi[l]->SetPtrSolver(solver);
        gi[l]->InitialGridInfo(dims);
    }
    const int q = dims == 2 ? 9 : 19;
    GridNodeLbmHip src, out;
    for (int i = 0; i < q; ++i) {
        double v;
        if (std::scanf("%lf", &v) != 1) return 3;
        src.f_.push_back(v);
    }
    double r, u[3];
    if (std::scanf("%lf %lf %lf %lf", &r, &u[0], &u[1], &u[2]) != 4) return 3;
    src.rho_ = r;
    src.velocity_.assign(u, u + dims);
    gi[1]->NodeInfoCoarse2fine(src, &out);      // coarse level 1 -> fine level 2
    for (double v : out.f_) std::printf("%.17g ", v);
    std::printf("\n");
    gi[2]->NodeInfoFine2Coarse(src, &out);      // fine level 2 -> coarse level 1
    for (double v : out.f_) std::printf("%.17g ", v);
    std::printf("\n%d %d\n", gi[1]->GetSizeOfGridNodeInfoForCheckPoint(),
                gi[1]->GetSizeOfGridNodeInfoForMpiCommunication());
    return 0;
}
"""


def _lattice(dims):
    from oracle.zlbm_oracle import lattice
    cx, cy, cz, w, _ = lattice(dims)
    return np.stack([cx, cy, cz], 1).astype(np.float64), np.asarray(w, np.float64)


def _feq(c, w, rho, u):
    cu = c @ u
    return w * rho * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * u @ u)


@pytest.fixture(scope="module")
def program(tmp_path_factory):
    d = tmp_path_factory.mktemp("hostnodes")
    src = d / "nodes.cpp"
    src.write_text(PROGRAM)
    exe = str(d / "nodes")
    pkg = os.path.join(ROOT, "zllbm_amd")
    subprocess.run(["g++", "-O1", "-std=c++17", "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(pkg, "host"),
                    str(src), "-o", exe, "-L" + pkg, "-lzllbm_host", "-lzllbm_hip", "-Wl,-rpath," + pkg],
                   check=True)
    return exe


@pytest.mark.parametrize("dims", [2, 3])
def test_node_info_between_levels(program, dims):
    c, w = _lattice(dims)
    rng = np.random.default_rng(dims)
    u_f = np.array([0.03, -0.02, 0.01 if dims == 3 else 0.0])
    f = _feq(c, w, 1.02, u_f) + rng.normal(0, 1e-3, len(w))
    rho_node, u_node = 0.98, np.array([0.011, 0.004, -0.006 if dims == 3 else 0.0])
    inp = " ".join(repr(float(v)) for v in list(f) + [rho_node] + list(u_node)) + "\n"
    out = subprocess.run([program, str(dims)], input=inp, capture_output=True, text=True, check=True).stdout.split("\n")
    c2f = np.array(out[0].split(), np.float64)
    f2c = np.array(out[1].split(), np.float64)
    sizes = [int(v) for v in out[2].split()]
    nu = 0.013
    tau = lambda l: 3 * nu * 2 ** l + 0.5
    # coarse to fine: macros from the coarse populations, ratio tau_f / (2 tau_c) of levels 1 -> 2
    rho = f.sum()
    u = (c.T @ f) / rho
    feq = _feq(c, w, rho, u)
    np.testing.assert_allclose(c2f, feq + 0.5 * tau(2) / tau(1) * (f - feq), rtol=0, atol=1e-15)
    # fine to coarse: the fine node's own rho, velocity, ratio 2 tau_c / tau_f
    feq = _feq(c, w, rho_node, u_node)
    np.testing.assert_allclose(f2c, feq + 2 * tau(1) / tau(2) * (f - feq), rtol=0, atol=1e-15)
    q = len(w)
    assert sizes == [4 + 8 * (q + dims), 8 * (q + 1 + dims + dims)]
