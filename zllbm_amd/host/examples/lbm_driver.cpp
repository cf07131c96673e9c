// lbm_driver.cpp -- runs the LBM time step through the reference-style plugin interfaces:
// a grid manager with one GridInfoLbmHip per level holding the node maps and interface layers,
// a SolverLbmHip created by its creator and configured with the reference's setters, and the
// AmrManager::TimeMarching loop.  Mirrors the setup sequence of the reference's example drivers
// (examples/lbm/immersed_boundary/cylinder2d/cylinder2d.cpp:14-111).
//
// usage: zllbm_lbm_driver <grid file> <parameter file> <background steps> <output file>
//   grid file:  "ZLGRID1" + dims, levels, index bounds, then per level keys, flags and layers
//   parameters: "key = value" lines (viscosity, rho, velocity, collision, les, cs, les_tau_mode,
//               precision = <fp64|fp32|bf16>, closure = <file>, mlp_operands = <fp16|bf16>,
//               checkpoint_in = <file> (start from it), checkpoint_out = <file> (write it at the end),
//               checkpoint_dir_in / checkpoint_dir_out = <folder> (the same with the reference's
//               checkpoint folder: loadings, grid_nodes, interface_nodes),
//               vtu_out = <prefix> (VTU files of every level at the end),
//               bc.xmin ... bc.zmax = <bounce_back|periodic|nee> [u v w],
//               time_interpolation = <0|1>, graph_replay = <0|1> (default 1: SetGraphReplay),
//               init = <file> (initial density and velocity: per level an int64 node count, then
//               the densities and the velocity components node by node, as doubles, in key order),
//               warmup = <steps> (of the background steps, the first ones, left out of the timing),
//               ib = <file> (immersed-boundary markers, SetImmersedBoundary: records of int32 level,
//               int32 stencil, double stiffness, int64 n, then n x 3 coordinates, n areas and
//               n x 3 velocities, as doubles))
//   stdout:     the solver method, "ms_per_step <t>" (wall time of the background steps after the
//               warm-up, device synchronised at both ends) and "done <steps> steps"
//   closure file (les = learned): whitespace-separated numbers, the 1409 network parameters,
//               then 9 each of input mean, spread, training minimum and maximum, then the output
//               scale
//   output:     per level the node count and the post-stream populations in key order ("-": none)
#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include "lbm_hip_solver.h"

using namespace rootproject;
using namespace rootproject::lbmproject;

namespace {
template <typename T>
bool rd(std::ifstream &in, T *v, size_t n = 1) {
    in.read(reinterpret_cast<char *>(v), sizeof(T) * n);
    return bool(in);
}

amrproject::EDomainBoundaryDirection dir_of(const std::string &s) {
    if (s == "xmin") return amrproject::EDomainBoundaryDirection::kBoundaryXMin;
    if (s == "xmax") return amrproject::EDomainBoundaryDirection::kBoundaryXMax;
    if (s == "ymin") return amrproject::EDomainBoundaryDirection::kBoundaryYMin;
    if (s == "ymax") return amrproject::EDomainBoundaryDirection::kBoundaryYMax;
    if (s == "zmin") return amrproject::EDomainBoundaryDirection::kBoundaryZMin;
    if (s == "zmax") return amrproject::EDomainBoundaryDirection::kBoundaryZMax;
    throw ZlhError("unknown boundary " + s);
}
}  // namespace

int main(int argc, char **argv) {
    if (argc != 5) {
        std::fprintf(stderr, "usage: %s <grid> <params> <steps> <out>\n", argv[0]);
        return 2;
    }
    try {
        std::ifstream in(argv[1], std::ios::binary);
        char magic[8];
        if (!rd(in, magic, 8) || std::strncmp(magic, "ZLGRID1", 7) != 0) throw ZlhError("bad grid file");
        int32_t dims, nlev;
        uint32_t dmin[3], dmax[3];
        rd(in, &dims); rd(in, &nlev); rd(in, dmin, 3); rd(in, dmax, 3);
        // solver (SolverCreatorLbmD2Q9 / D3Q19) and grid manager
        std::shared_ptr<amrproject::SolverInterface> solver =
            dims == 2 ? SolverCreatorLbmD2Q9Hip().CreateSolver() : SolverCreatorLbmD3Q19Hip().CreateSolver();
        auto *lbm = dynamic_cast<SolverLbmHip *>(solver.get());
        amrproject::GridManagerInterface gm(dims);
        amrproject::SFBitsetAuxInterface sfbitset_aux;
        gm.SetMaxLevel(nlev - 1);
        gm.SetIndexRangeOfBackgroundNode(std::vector<DefAmrLUint>(dmin, dmin + dims),
                                         std::vector<DefAmrLUint>(dmax, dmax + dims));
        gm.vec_ptr_solver_.push_back(solver);
        solver->SetPtrToGridManager(&gm);
        std::vector<std::vector<uint64_t>> level_keys(nlev);
        for (int l = 0; l < nlev; ++l) {
            auto gi = solver->ptr_grid_info_creator_->CreateGridInfo();
            gi->SetGridLevel(l);
            gi->SetPtrSolver(solver);
            gi->SetPtrToParentGridManager(&gm);
            gi->SetPtrSFBitsetAux(&sfbitset_aux);
            gi->InitialGridInfo(dims);
            int64_t n;
            rd(in, &n);
            std::vector<uint64_t> keys(n);
            std::vector<int32_t> flags(n);
            rd(in, keys.data(), n);
            rd(in, flags.data(), n);
            for (int64_t i = 0; i < n; ++i) {
                auto node = gi->GridNodeCreator();
                node->flag_status_ = flags[i];
                gi->map_grid_node_.emplace(DefSFBitset(keys[i]), std::move(node));
            }
            level_keys[l] = keys;
            int32_t nlayers;
            rd(in, &nlayers);
            auto info = std::make_shared<amrproject::InterfaceLayerInfo>();
            for (int k = 0; k < nlayers; ++k) {
                int32_t kind, layer;
                int64_t m;
                rd(in, &kind); rd(in, &layer); rd(in, &m);
                std::vector<uint64_t> lk(m);
                rd(in, lk.data(), m);
                std::vector<DefMap<DefInt>> *v = kind == ZLH_OUTER_F2C ? &info->vec_outer_fine2coarse_
                                               : kind == ZLH_OUTER_C2F ? &info->vec_outer_coarse2fine_
                                               : kind == ZLH_INNER_F2C ? &info->vec_inner_fine2coarse_
                                                                       : &info->vec_inner_coarse2fine_;
                if ((int)v->size() <= layer) v->resize(layer + 1);
                for (uint64_t key : lk) (*v)[layer].emplace(DefSFBitset(key), 0);
            }
            if (nlayers) gi->map_ptr_interface_layer_info_[{amrproject::ECriterionType::kGeometry, 0}] = info;
            gm.vec_ptr_grid_info_.push_back(gi);
        }
        if (!in) throw ZlhError("truncated grid file");
        // parameters (ReadAndSetupSolverParameters keys, lbm_solver_interface.cpp:49-117)
        int warmup = 0;
        std::ifstream pin(argv[2]);
        std::string line;
        while (std::getline(pin, line)) {
            if (line.empty() || line[0] == '#') continue;
            const size_t eq = line.find('=');
            if (eq == std::string::npos) continue;
            std::string key = line.substr(0, eq);
            key.erase(key.find_last_not_of(" \t") + 1);
            std::istringstream val(line.substr(eq + 1));
            if (key == "viscosity") { double v; val >> v; lbm->SetDefaultViscosity(v); }
            else if (key == "rho") { double v; val >> v; lbm->SetDefaultDensity(v); }
            else if (key == "velocity") { std::vector<double> u(dims); for (auto &x : u) val >> x; lbm->SetDefaultVelocity(u); }
            else if (key == "collision") {
                std::string t; int l = 0;
                while (val >> t) lbm->SetCollisionOperator(l++, t == "mrt" ? ELbmCollisionOperatorType::kLbmMrt
                                                                           : ELbmCollisionOperatorType::kLbmSrt);
            }
            else if (key == "les") { std::string t; val >> t; if (t != "none") lbm->SetLesModel(t); }
            else if (key == "cs") { double v; val >> v; lbm->SetSmagorinskyConstant(v); }
            else if (key == "les_tau_mode") { int v; val >> v; lbm->SetLesTauMode(v); }
            else if (key == "precision") {
                std::string t; val >> t;
                lbm->SetPrecision(t == "fp32" ? ZLH_FP32_SHIFTED : t == "bf16" ? ZLH_BF16_SHIFTED : ZLH_FP64);
            }
            else if (key == "time_interpolation") { int v; val >> v; lbm->SetTimeInterpolation(v != 0); }
            else if (key == "graph_replay") { int v; val >> v; lbm->SetGraphReplay(v != 0); }
            else if (key == "ib") {
                // records {int32 level, int32 stencil, double stiffness, int64 n, n x 3 coordinates,
                // n areas, n x 3 velocities} until the end of the file
                std::string path; val >> path;
                std::ifstream bin(path, std::ios::binary);
                int32_t level, stencil;
                double stiffness;
                int64_t n;
                while (rd(bin, &level) && rd(bin, &stencil) && rd(bin, &stiffness) && rd(bin, &n)) {
                    std::vector<double> xyz((size_t)n * 3), area((size_t)n), vel((size_t)n * 3);
                    if (n < 0 || !rd(bin, xyz.data(), xyz.size()) || !rd(bin, area.data(), area.size()) ||
                        !rd(bin, vel.data(), vel.size()))
                        throw ZlhError("immersed-boundary file " + path + " is short");
                    std::vector<std::array<DefReal, 3>> x((size_t)n), v((size_t)n);
                    for (int64_t i = 0; i < n; ++i) {
                        x[i] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
                        v[i] = {vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]};
                    }
                    lbm->SetImmersedBoundary(level, x, area, v, stencil, stiffness);
                }
            }
            else if (key == "warmup") { val >> warmup; }
            else if (key == "init") {
                std::string path; val >> path;
                std::ifstream iin(path, std::ios::binary);
                for (int l = 0; l < nlev; ++l) {
                    int64_t n;
                    if (!rd(iin, &n) || n != (int64_t)level_keys[l].size())
                        throw ZlhError("init file " + path + ": level " + std::to_string(l) + " node count");
                    std::vector<double> rho(n), u((size_t)n * dims);
                    if (!rd(iin, rho.data(), rho.size()) || !rd(iin, u.data(), u.size()))
                        throw ZlhError("init file " + path + " is short");
                    lbm->SetInitialMacros(l, rho, u);
                }
            }
            else if (key == "mlp_operands") {
                std::string t; val >> t;
                lbm->SetMlpOperands(t == "bf16" ? ZLH_MLP_OPERANDS_BF16 : ZLH_MLP_OPERANDS_FP16);
            }
            else if (key == "closure") {
                std::string path; val >> path;
                std::ifstream cin(path);
                auto take = [&](size_t n) {
                    std::vector<float> v(n);
                    for (auto &x : v)
                        if (!(cin >> x)) throw ZlhError("closure file " + path + " is short");
                    return v;
                };
                const std::vector<float> p = take(1409), mean = take(9), sd = take(9), lo = take(9), hi = take(9);
                double scale;
                if (!(cin >> scale)) throw ZlhError("closure file " + path + " is short");
                lbm->SetLearnedClosure(p, mean, sd, lo, hi, scale);
            }
            else if (key.rfind("bc.", 0) == 0) {
                std::string scheme; val >> scheme;
                std::vector<double> v; double x; while (val >> x) v.push_back(x);
                ELbmBoundaryConditionScheme s = scheme == "bounce_back" ? ELbmBoundaryConditionScheme::kBounceBack
                                              : scheme == "periodic" ? ELbmBoundaryConditionScheme::kPeriodic
                                              : scheme == "nee" ? ELbmBoundaryConditionScheme::kNonEqExtrapolation
                                                                : ELbmBoundaryConditionScheme::kUndefined;
                lbm->SetDomainBoundaryCondition(dir_of(key.substr(3)), s, v);
            }
        }
        std::string checkpoint_in, checkpoint_out, checkpoint_dir_in, checkpoint_dir_out, vtu_out;
        {
            std::ifstream pin2(argv[2]);
            while (std::getline(pin2, line)) {
                std::istringstream ls(line);
                std::string key, eq, val;
                if (ls >> key >> eq >> val && eq == "=") {
                    if (key == "checkpoint_in") checkpoint_in = val;
                    if (key == "checkpoint_out") checkpoint_out = val;
                    if (key == "checkpoint_dir_in") checkpoint_dir_in = val;
                    if (key == "checkpoint_dir_out") checkpoint_dir_out = val;
                    if (key == "vtu_out") vtu_out = val;
                }
            }
        }
        solver->SolverInitial();
        if (!checkpoint_in.empty()) lbm->ReadCheckPointGridNodes(checkpoint_in);
        if (!checkpoint_dir_in.empty()) lbm->ReadCheckPoint(checkpoint_dir_in);
        std::cout << solver->GetSolverMethod() << std::endl;
        const int steps = std::atoi(argv[3]);
        if (warmup < 0 || warmup > steps) throw ZlhError("warm-up of " + std::to_string(warmup) + " steps out of " +
                                                         std::to_string(steps));
        for (int t = 0; t < warmup; ++t) TimeMarching(&gm, (DefAmrLUint)t);
        lbm->BuildDevice();
        if (zlh_synchronize(lbm->Context()) != ZLH_OK) throw ZlhError("zlh_synchronize failed");
        const auto t0 = std::chrono::steady_clock::now();
        for (int t = warmup; t < steps; ++t) TimeMarching(&gm, (DefAmrLUint)t);
        if (zlh_synchronize(lbm->Context()) != ZLH_OK) throw ZlhError("zlh_synchronize failed");
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (steps > warmup) std::cout << "ms_per_step " << ms / (steps - warmup) << std::endl;
        if (!checkpoint_out.empty()) lbm->WriteCheckPointGridNodes(checkpoint_out);
        if (!checkpoint_dir_out.empty()) lbm->WriteCheckPoint(checkpoint_dir_out);
        if (!vtu_out.empty()) lbm->WriteVtu(vtu_out);
        std::ofstream out;
        if (std::strcmp(argv[4], "-") != 0) out.open(argv[4], std::ios::binary);
        for (int l = 0; l < nlev && out.is_open(); ++l) {
            lbm->CopyPopulationsToHost(l);
            const int64_t n = (int64_t)level_keys[l].size();
            out.write(reinterpret_cast<const char *>(&n), sizeof(n));
            for (uint64_t key : level_keys[l]) {
                auto *node = dynamic_cast<GridNodeLbmHip *>(gm.vec_ptr_grid_info_[l]->map_grid_node_.at(DefSFBitset(key)).get());
                out.write(reinterpret_cast<const char *>(node->f_.data()), sizeof(double) * node->f_.size());
            }
        }
        std::cout << "done " << steps << " steps" << std::endl;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
