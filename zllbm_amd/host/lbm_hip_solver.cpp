// lbm_hip_solver.cpp -- SolverLbmHip / GridInfoLbmHip: the reference's solver-manager and
// grid-information plugin calls translated to the C ABI of the device library.
#include "lbm_hip_solver.h"
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <tuple>

namespace rootproject {
namespace lbmproject {

namespace {
SolverLbmHip *solver_of(const amrproject::GridInfoInterface &g) {
    auto sp = g.GetPtrToSolver().lock();
    auto *s = dynamic_cast<SolverLbmHip *>(sp.get());
    if (!s) throw ZlhError("grid information has no LBM HIP solver");
    return s;
}

// index along axis a of a Morton key (x at bit dims i, y at dims i + 1, z at dims i + 2;
// SFBitsetComputeIndices, grid/morton_aux.cpp:20-119, 336-351)
uint32_t morton_axis(uint64_t key, int dims, int a) {
    uint32_t v = 0;
    for (int i = 0; a + dims * i < 64; ++i) v |= (uint32_t)((key >> (a + dims * i)) & 1u) << i;
    return v;
}

// Morton key of an index triple (SFBitsetEncoding, grid/morton_aux.cpp:274-286)
uint64_t morton_key(int dims, uint32_t x, uint32_t y, uint32_t z) {
    uint64_t k = 0;
    const uint32_t c[3] = {x, y, z};
    for (int i = 0; dims * i < 64; ++i)
        for (int a = 0; a < dims && a + dims * i < 64; ++a) k |= (uint64_t)((c[a] >> i) & 1u) << (a + dims * i);
    return k;
}

// checkpoint record order of one node: background cell, level, then z, y, x inside the cell
struct CheckpointOrder {
    uint64_t cell;
    int32_t level;
    uint32_t z, y, x;
    uint64_t key;
    bool operator<(const CheckpointOrder &o) const {
        return std::tie(cell, level, z, y, x) < std::tie(o.cell, o.level, o.z, o.y, o.x);
    }
};
}  // namespace

// ---------------------------------------------------------------------------------------------
// GridInfoLbmHip
// ---------------------------------------------------------------------------------------------
// GridInfoLbmInteface::AdvancingAtCurrentTime (lbm_grid_interface.cpp:104-153): the IB hook,
// collide + stream and the domain boundaries run inside RunSolverOnGivenGrid on the device; the
// coarse/fine exchange follows through InformationFromGridOfDifferentLevel.
void GridInfoLbmHip::AdvancingAtCurrentTime(const amrproject::ETimeSteppingScheme time_scheme,
                                            const DefInt time_step_level, const DefReal time_step_current,
                                            amrproject::MpiManager *const, amrproject::CriterionManager *const) {
    SolverLbmHip *s = solver_of(*this);
    if (s->AdvanceWithinBackgroundStep(i_level_, time_step_level)) return;
    if (ptr_sfbitset_aux_ == nullptr) throw ZlhError("pointer to SFBitsetAux is nullptr");
    s->RunSolverOnGivenGrid(time_scheme, time_step_level, time_step_current, *ptr_sfbitset_aux_, this);
    s->InformationFromGridOfDifferentLevel(time_step_level, *ptr_sfbitset_aux_, this);
}

std::unique_ptr<amrproject::GridNode> GridInfoLbmHip::GridNodeCreator() const {
    auto n = std::make_unique<GridNodeLbmHip>();
    if (auto sp = ptr_solver_.lock()) {
        auto *s = dynamic_cast<SolverLbmHip *>(sp.get());
        if (s) n->rho_ = s->GetDefaultDensity();
    }
    return n;
}

namespace {
// D2Q9 / D3Q19 directions in the reference's order (lbm_d2q9.h:53-61, lbm_d3q19.h:73-85) and
// their weights
struct HostLattice {
    int q;
    int c[19][3];
    DefReal w[19];
};
HostLattice host_lattice(int dims) {
    HostLattice L{};
    if (dims == 2) {
        const int c[9][2] = {{0, 0}, {1, 0}, {0, 1}, {-1, 0}, {0, -1}, {1, 1}, {-1, 1}, {-1, -1}, {1, -1}};
        L.q = 9;
        for (int i = 0; i < 9; ++i) {
            L.c[i][0] = c[i][0];
            L.c[i][1] = c[i][1];
            const int c2 = c[i][0] * c[i][0] + c[i][1] * c[i][1];
            L.w[i] = c2 == 0 ? 4. / 9. : (c2 == 1 ? 1. / 9. : 1. / 36.);
        }
    } else {
        const int c[19][3] = {{0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1},
                              {1, 1, 0}, {-1, 1, 0}, {1, -1, 0}, {-1, -1, 0}, {1, 0, 1}, {-1, 0, 1}, {1, 0, -1},
                              {-1, 0, -1}, {0, 1, 1}, {0, -1, 1}, {0, 1, -1}, {0, -1, -1}};
        L.q = 19;
        for (int i = 0; i < 19; ++i) {
            for (int a = 0; a < 3; ++a) L.c[i][a] = c[i][a];
            const int c2 = c[i][0] * c[i][0] + c[i][1] * c[i][1] + c[i][2] * c[i][2];
            L.w[i] = c2 == 0 ? 1. / 3. : (c2 == 1 ? 1. / 18. : 1. / 36.);
        }
    }
    return L;
}

// compressible equilibrium (CalFeq2D/3DCompressible, lbm_solver_interface.cpp:374-435)
void host_feq(const HostLattice &L, DefReal rho, const std::vector<DefReal> &u, std::vector<DefReal> *feq) {
    feq->resize(L.q);
    const DefReal usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    for (int q = 0; q < L.q; ++q) {
        const DefReal cu = L.c[q][0] * u[0] + L.c[q][1] * u[1] + L.c[q][2] * u[2];
        (*feq)[q] = L.w[q] * rho * (1. + 3. * cu + 4.5 * cu * cu - 1.5 * usq);
    }
}

// equilibrium part plus the scaled non-equilibrium part of a source node; the macros from the
// source's populations without force (CalMacro2D/3DCompressible, lbm_solver_interface.cpp:477-530)
// or its own rho_, velocity_ (those of its last collision)
void rescaled_populations(int dims, const GridNodeLbmHip &src, DefReal ratio, GridNodeLbmHip *dst, bool macros_from_f) {
    const HostLattice L = host_lattice(dims);
    if ((int)src.f_.size() != L.q) throw ZlhError("node information: the source node has no populations");
    std::vector<DefReal> feq, u(3, 0.);
    DefReal rho = 0.;
    if (macros_from_f) {
        for (int q = 0; q < L.q; ++q) {
            rho += src.f_[q];
            for (int a = 0; a < 3; ++a) u[a] += L.c[q][a] * src.f_[q];
        }
        for (int a = 0; a < 3; ++a) u[a] /= rho;
    } else {
        rho = src.rho_;
        for (int a = 0; a < dims && a < (int)src.velocity_.size(); ++a) u[a] = src.velocity_[a];
    }
    host_feq(L, rho, u, &feq);
    u.resize(dims);
    dst->rho_ = rho;
    dst->velocity_ = u;
    dst->f_.resize(feq.size());
    for (size_t q = 0; q < feq.size(); ++q) dst->f_[q] = feq[q] + ratio * (src.f_[q] - feq[q]);
}

// relaxation time of a level (CalRelaxationTime, lbm_collision_models.cpp:105-120): the level's
// viscosity nu0 / dt in lattice units of that level, tau = 3 nu / dt + 1/2
DefReal level_tau(DefReal nu0, DefInt level) { return 3. * nu0 * (DefReal)(1u << level) + 0.5; }
}  // namespace

int GridInfoLbmHip::TransferInfoFromCoarseGrid(const amrproject::SFBitsetAuxInterface &, const DefInt,
                                               const amrproject::GridInfoInterface &) {
    SolverLbmHip *s = solver_of(*this);
    if (!s->Context()) s->BuildDevice();
    const zlh_status st = zlh_transfer_from_coarse(s->Context(), i_level_);
    if (st != ZLH_OK) throw ZlhError(std::string("zlh_transfer_from_coarse: ") + zlh_last_error(s->Context()));
    return 0;
}

int GridInfoLbmHip::TransferInfoToCoarseGrid(const amrproject::SFBitsetAuxInterface &, const DefInt,
                                             amrproject::GridInfoInterface *const) {
    SolverLbmHip *s = solver_of(*this);
    if (!s->Context()) s->BuildDevice();
    const zlh_status st = zlh_transfer_to_coarse(s->Context(), i_level_);
    if (st != ZLH_OK) throw ZlhError(std::string("zlh_transfer_to_coarse: ") + zlh_last_error(s->Context()));
    return 0;
}

// called on the coarse grid's information (the reference's coarse_grid_info.NodeInfoCoarse2fine):
// macros from the coarse populations, ratio tau_f / (2 tau_c) of this (coarse) level
void GridInfoLbmHip::NodeInfoCoarse2fine(const amrproject::GridNode &coarse_node, amrproject::GridNode *const ptr_fine_node) const {
    const auto &c = dynamic_cast<const GridNodeLbmHip &>(coarse_node);
    auto *f = dynamic_cast<GridNodeLbmHip *>(ptr_fine_node);
    if (!f) throw ZlhError("NodeInfoCoarse2fine: the fine node is not an LBM node");
    const DefReal nu = solver_of(*this)->GetDefaultViscosity();
    rescaled_populations(dims_, c, 0.5 * level_tau(nu, i_level_ + 1) / level_tau(nu, i_level_), f, true);
}

// called on the fine grid's information: rho_, velocity_ of the fine node's last collision, ratio
// 2 tau_c / tau_f of this (fine) level
void GridInfoLbmHip::NodeInfoFine2Coarse(const amrproject::GridNode &fine_node, amrproject::GridNode *const ptr_coarse_node) const {
    const auto &f = dynamic_cast<const GridNodeLbmHip &>(fine_node);
    auto *c = dynamic_cast<GridNodeLbmHip *>(ptr_coarse_node);
    if (!c) throw ZlhError("NodeInfoFine2Coarse: the coarse node is not an LBM node");
    const DefReal nu = solver_of(*this)->GetDefaultViscosity();
    rescaled_populations(dims_, f, 2. * level_tau(nu, i_level_ - 1) / level_tau(nu, i_level_), c, false);
}

int GridInfoLbmHip::GetSizeOfGridNodeInfoForCheckPoint() const {
    const SolverLbmHip *s = solver_of(*this);
    return (int)(sizeof(DefInt) + sizeof(DefReal) * (s->k0NumQ_ + s->GetNumForces()));
}

int GridInfoLbmHip::GetSizeOfGridNodeInfoForMpiCommunication() const {
    const SolverLbmHip *s = solver_of(*this);
    return (int)(sizeof(DefReal) * (s->k0NumQ_ + 1 + s->GetSolverDim() + s->GetNumForces()));
}

void GridInfoLbmHip::SetupOutputVariables() {
    const SolverLbmHip *s = solver_of(*this);
    output_variables_.clear();
    for (const std::string &name : output_variable_name_) {
        if (name != "rho" && name != "velocity") continue;
        auto v = std::make_unique<amrproject::OutputNodeVariableInfoInterface>();
        v->output_name_ = name;
        v->variable_dims_ = name == "rho" ? 1 : s->GetSolverDim();
        output_variables_.push_back(std::move(v));
    }
}

void GridInfoLbmHip::SetNodeVariablesAsZeros(amrproject::GridNode *const ptr_node) {
    auto *n = dynamic_cast<GridNodeLbmHip *>(ptr_node);
    if (!n) return;
    n->rho_ = 0.;
    std::fill(n->velocity_.begin(), n->velocity_.end(), 0.);
    std::fill(n->f_.begin(), n->f_.end(), 0.);
}

// ---------------------------------------------------------------------------------------------
// SolverLbmHip
// ---------------------------------------------------------------------------------------------
SolverLbmHip::SolverLbmHip(DefInt dims, DefInt num_q) : k0NumQ_(num_q) {
    k0SolverDims_ = dims;
    name_ = dims == 2 ? "LbmD2Q9" : "LbmD3Q19";
    solver_type_ = name_;
    k0Velocity_.assign(dims, 0.);
    ptr_grid_info_creator_ = std::make_unique<GridInfoLbmHipCreator>();
}

SolverLbmHip::~SolverLbmHip() {
    if (ctx_) zlh_destroy(ctx_);
}

std::string SolverLbmHip::GetSolverMethod() {
    return "Solver LBM D" + std::to_string(k0SolverDims_) + "Q" + std::to_string(k0NumQ_) +
           " Model is active (MI355X).";
}

void SolverLbmHip::Check(zlh_status s, const char *what) const {
    if (s != ZLH_OK) {
        std::string msg = std::string(what) + " failed (" + std::to_string(s) + "): " + zlh_last_error(ctx_);
        std::fprintf(stderr, "%s\n", msg.c_str());
        throw ZlhError(msg);
    }
}

void SolverLbmHip::SetLesModel(const std::string &type) {
    if (type == "smagorinsky") les_ = ZLH_LES_SMAGORINSKY;
    else if (type == "learned") les_ = ZLH_LES_MLP;
    else throw ZlhError("LES model " + type + " is not defined, please choose from: smagorinsky, learned");
}

void SolverLbmHip::SetLearnedClosure(const std::vector<float> &params, const std::vector<float> &mean,
                                     const std::vector<float> &std_dev, const std::vector<float> &fmin,
                                     const std::vector<float> &fmax, const DefReal nu_scale) {
    constexpr size_t kParams = 32 * 9 + 32 + 32 * 32 + 32 + 32 + 1, kFeatures = 9;
    if (params.size() != kParams || mean.size() != kFeatures || std_dev.size() != kFeatures ||
        fmin.size() != kFeatures || fmax.size() != kFeatures)
        throw ZlhError("learned closure: expected 1409 parameters and 9 values per feature table");
    closure_ = Closure{params, mean, std_dev, fmin, fmax, nu_scale};
}

void SolverLbmHip::SetCollisionOperator(const DefInt i_level, const ELbmCollisionOperatorType type) {
    if (static_cast<DefInt>(collision_type_.size()) <= i_level) collision_type_.resize(i_level + 1, type);
    collision_type_[i_level] = type;
}

void SolverLbmHip::SetDomainBoundaryCondition(const amrproject::EDomainBoundaryDirection which,
                                              const ELbmBoundaryConditionScheme scheme,
                                              const std::vector<DefReal> &values) {
    BoundarySetting &b = bc_[static_cast<int>(which)];
    b.boundary_scheme_ = scheme;
    b.SetValues(values);
}

void SolverLbmHip::SetImmersedBoundary(DefInt level, const std::vector<std::array<DefReal, 3>> &xyz,
                                       const std::vector<DefReal> &area,
                                       const std::vector<std::array<DefReal, 3>> &velocity, DefInt stencil_dis,
                                       DefReal stiffness) {
    Ib &m = ib_[level];
    m.xyz.clear();
    m.vel.clear();
    for (auto &p : xyz) m.xyz.insert(m.xyz.end(), p.begin(), p.end());
    for (auto &p : velocity) m.vel.insert(m.vel.end(), p.begin(), p.end());
    m.area = area;
    m.stencil = stencil_dis;
    m.stiffness = stiffness;
    bool_forces_ = true;
}

// SolverLbmInterface::SolverInitial (lbm_solver_interface.cpp:19-44): parameter checks; the
// device state needs the mesh and is built on the first time step (or by BuildDevice)
void SolverLbmHip::SolverInitial() {
    if (k0LbmViscosity_ < 1e-6) std::fprintf(stderr, "LBM viscosity is less than predefined kEps\n");
    if (static_cast<DefInt>(k0Velocity_.size()) != k0SolverDims_)
        throw ZlhError("size of initial velocity is not equal to solver dimension");
}

void SolverLbmHip::BuildDevice() {
    if (ctx_) return;
    if (!ptr_grid_manager_) throw ZlhError("Grid manager is not assigned for solver.");
    amrproject::GridManagerInterface &gm = *ptr_grid_manager_;
    zlh_solver_desc d{};
    d.dims = k0SolverDims_;
    d.max_level = gm.GetMaxLevel();
    d.precision = precision_;
    d.compressible = 1;
    d.nu0 = k0LbmViscosity_;
    d.rho0 = k0Rho_;
    for (int a = 0; a < k0SolverDims_ && a < (int)k0Velocity_.size(); ++a) d.u0[a] = k0Velocity_[a];
    d.les = les_;
    d.mlp_operands = mlp_operands_;
    d.les_tau_mode = les_tau_mode_;
    d.cs = cs_;
    d.has_force = bool_forces_ ? 1 : 0;
    for (int a = 0; a < (int)k0ConstForce_.size() && a < 3; ++a) d.const_force[a] = k0ConstForce_[a];
    // InstantiateCollisionOperator (collision_model/lbm_collision_models.cpp:60-73): SRT by default,
    // the last given type for levels beyond the list
    for (int l = 0; l < 16; ++l) {
        ELbmCollisionOperatorType t = collision_type_.empty() ? ELbmCollisionOperatorType::kLbmSrt
                                      : collision_type_[std::min<size_t>(l, collision_type_.size() - 1)];
        d.collision[l] = t == ELbmCollisionOperatorType::kLbmMrt ? ZLH_MRT : ZLH_SRT;
    }
    for (auto &kv : bc_) {
        d.bc_scheme[kv.first] = static_cast<int>(kv.second.boundary_scheme_);
        for (int a = 0; a < 3; ++a) d.bc_value[kv.first][a] = kv.second.values_[a];
    }
    const std::vector<DefAmrLUint> lo = gm.GetMinIndexOfBackgroundNodeArrAsVec(),
                                   hi = gm.GetMaxIndexOfBackgroundNodeArrAsVec();
    for (int a = 0; a < k0SolverDims_; ++a) {
        if (a >= (int)lo.size() || a >= (int)hi.size())
            throw ZlhError("background index range not set on the grid manager");
        d.dom_min[a] = lo[a];
        d.dom_max[a] = hi[a];
    }
    zlh_status s = zlh_create(&d, &ctx_);
    if (s != ZLH_OK) {
        ctx_ = nullptr;
        throw ZlhError("zlh_create failed (" + std::to_string(s) + ")");
    }
    Check(zlh_set_time_interpolation(ctx_, time_interpolation_ ? 1 : 0), "zlh_set_time_interpolation");
    for (int l = 0; l <= gm.GetMaxLevel(); ++l) {
        amrproject::GridInfoInterface &gi = *gm.vec_ptr_grid_info_.at(l);
        std::vector<std::pair<uint64_t, int32_t>> nodes;
        nodes.reserve(gi.map_grid_node_.size());
        for (auto &kv : gi.map_grid_node_) nodes.emplace_back(kv.first.to_ullong(), kv.second->flag_status_);
        std::sort(nodes.begin(), nodes.end());
        std::vector<uint64_t> keys(nodes.size());
        std::vector<int32_t> flags(nodes.size());
        for (size_t i = 0; i < nodes.size(); ++i) { keys[i] = nodes[i].first; flags[i] = nodes[i].second; }
        Check(zlh_set_level_nodes(ctx_, l, keys.data(), flags.data(), (int64_t)keys.size()), "zlh_set_level_nodes");
        // interface layers: every interface's maps of one kind and layer are merged
        std::map<std::pair<int, int>, std::vector<uint64_t>> layers;
        for (auto &it : gi.map_ptr_interface_layer_info_) {
            const amrproject::InterfaceLayerInfo &li = *it.second;
            auto add = [&](int kind, const std::vector<DefMap<DefInt>> &v) {
                for (size_t layer = 0; layer < v.size() && layer < 3; ++layer)
                    for (auto &kv : v[layer]) layers[{kind, (int)layer}].push_back(kv.first.to_ullong());
            };
            add(ZLH_OUTER_F2C, li.vec_outer_fine2coarse_);
            add(ZLH_OUTER_C2F, li.vec_outer_coarse2fine_);
            add(ZLH_INNER_F2C, li.vec_inner_fine2coarse_);
            add(ZLH_INNER_C2F, li.vec_inner_coarse2fine_);
        }
        for (auto &kv : layers) {
            std::sort(kv.second.begin(), kv.second.end());
            kv.second.erase(std::unique(kv.second.begin(), kv.second.end()), kv.second.end());
            Check(zlh_set_interface_layer(ctx_, l, kv.first.first, kv.first.second, kv.second.data(),
                                          (int64_t)kv.second.size()),
                  "zlh_set_interface_layer");
        }
    }
    for (auto &kv : ib_) {
        const Ib &m = kv.second;
        Check(zlh_ib_set_markers(ctx_, kv.first, m.xyz.data(), m.area.data(), m.vel.data(), (int64_t)m.area.size(),
                                 m.stencil, m.stiffness),
              "zlh_ib_set_markers");
    }
    Check(zlh_build(ctx_), "zlh_build");
    if (les_ == ZLH_LES_MLP) {
        if (closure_.params.empty()) throw ZlhError("learned closure selected without SetLearnedClosure");
        Check(zlh_mlp_set(ctx_, closure_.params.data(), closure_.mean.data(), closure_.std_dev.data(),
                          closure_.fmin.data(), closure_.fmax.data(), closure_.nu_scale),
              "zlh_mlp_set");
    }
    // initial populations: node mirrors with f_ set, else feq(rho0, u0)
    Check(zlh_init_feq(ctx_), "zlh_init_feq");
    for (int l = 0; l <= gm.GetMaxLevel(); ++l) {
        amrproject::GridInfoInterface &gi = *gm.vec_ptr_grid_info_.at(l);
        std::vector<std::pair<uint64_t, const GridNodeLbmHip *>> nodes;
        bool any = false;
        for (auto &kv : gi.map_grid_node_) {
            auto *n = dynamic_cast<const GridNodeLbmHip *>(kv.second.get());
            nodes.emplace_back(kv.first.to_ullong(), n);
            if (n && (int)n->f_.size() == k0NumQ_) any = true;
        }
        if (!any) continue;
        std::sort(nodes.begin(), nodes.end(),
                  [](const auto &a, const auto &b) { return a.first < b.first; });
        std::vector<double> f(nodes.size() * k0NumQ_);
        std::vector<double> feq(k0NumQ_);
        for (size_t i = 0; i < nodes.size(); ++i) {
            const GridNodeLbmHip *n = nodes[i].second;
            if (!n || (int)n->f_.size() != k0NumQ_)
                throw ZlhError("level " + std::to_string(l) + ": populations set on some nodes only");
            std::copy(n->f_.begin(), n->f_.end(), f.begin() + i * k0NumQ_);
        }
        Check(zlh_upload_populations(ctx_, l, f.data()), "zlh_upload_populations");
    }
    for (auto &kv : init_macros_)
        Check(zlh_init_macros(ctx_, kv.first, kv.second.first.data(), kv.second.second.data()), "zlh_init_macros");
}

void SolverLbmHip::SetGraphReplay(const bool on) {
    if (cycle_pos_ != 0) throw ZlhError("graph replay switched inside a background step");
    graph_replay_ = on;
}

void SolverLbmHip::SetInitialMacros(DefInt level, const std::vector<DefReal> &rho, const std::vector<DefReal> &velocity) {
    if (!ptr_grid_manager_ || level < 0 || level > ptr_grid_manager_->GetMaxLevel())
        throw ZlhError("initial macros: bad level " + std::to_string(level));
    const size_t n = ptr_grid_manager_->vec_ptr_grid_info_.at(level)->map_grid_node_.size();
    if (rho.size() != n || velocity.size() != n * (size_t)k0SolverDims_)
        throw ZlhError("initial macros of level " + std::to_string(level) + ": expected " + std::to_string(n) +
                       " densities and " + std::to_string(n * k0SolverDims_) + " velocity components");
    init_macros_[level] = {rho, velocity};
    if (ctx_) Check(zlh_init_macros(ctx_, level, rho.data(), velocity.data()), "zlh_init_macros");
}

bool SolverLbmHip::AdvanceWithinBackgroundStep(DefInt level, DefInt time_step_level) {
    if (!graph_replay_) return false;
    if (!ctx_) BuildDevice();
    const DefInt max_level = ptr_grid_manager_->GetMaxLevel();
    if (cycle_order_.empty()) {
        amrproject::MultiTimeSteppingC2F scheme(max_level);
        cycle_order_ = scheme.k0TimeSteppingOrder_;
    }
    if (cycle_pos_ == 0) {
        if (level != 0 || time_step_level != 1)
            throw ZlhError("graph replay: a background step begins with level 0's substep 1, not level " +
                           std::to_string(level) + "'s substep " + std::to_string(time_step_level));
        Check(zlh_time_marching(ctx_, 1), "zlh_time_marching");
        cycle_tsl_.assign(max_level + 1, 0);
    } else if (level != cycle_order_[cycle_pos_]) {
        throw ZlhError("graph replay: substep of level " + std::to_string(level) + " where the coarse-to-fine order has level " +
                       std::to_string(cycle_order_[cycle_pos_]));
    }
    if (time_step_level != ++cycle_tsl_[level])
        throw ZlhError("graph replay: level " + std::to_string(level) + " substep " + std::to_string(time_step_level) +
                       " where " + std::to_string(cycle_tsl_[level]) + " is next");
    if (++cycle_pos_ == cycle_order_.size()) cycle_pos_ = 0;
    return true;
}

// SolverLbmInterface::RunSolverOnGivenGrid (lbm_solver_interface.cpp:246-262)
void SolverLbmHip::RunSolverOnGivenGrid(const amrproject::ETimeSteppingScheme, const DefInt time_step_level,
                                        const DefReal, const amrproject::SFBitsetAuxInterface &,
                                        amrproject::GridInfoInterface *const ptr_grid_info) {
    if (!ctx_) BuildDevice();
    Check(zlh_step_level(ctx_, ptr_grid_info->GetGridLevel(), time_step_level), "zlh_step_level");
}

// SolverLbmInterface::InformationFromGridOfDifferentLevel (lbm_solver_interface.cpp:224-238): after
// the fine level's even substeps, the fine grid's TransferInfoFromCoarseGrid, then its
// TransferInfoToCoarseGrid, with the coarse grid's information; with time interpolation also
// the (time-interpolated) TransferInfoFromCoarseGrid after odd substeps
void SolverLbmHip::InformationFromGridOfDifferentLevel(const DefInt time_step_level,
                                                       const amrproject::SFBitsetAuxInterface &sfbitset_aux,
                                                       amrproject::GridInfoInterface *const ptr_grid_info) {
    const DefInt i_level = ptr_grid_info->GetGridLevel();
    if (i_level == 0) return;
    amrproject::GridInfoInterface *coarse = ptr_grid_manager_->vec_ptr_grid_info_.at(i_level - 1).get();
    if (time_step_level % 2 == 0) {
        ptr_grid_info->TransferInfoFromCoarseGrid(sfbitset_aux, amrproject::NodeBitStatus::kNodeStatusCoarse2FineGhost_,
                                                  *coarse);
        ptr_grid_info->TransferInfoToCoarseGrid(sfbitset_aux, amrproject::NodeBitStatus::kNodeStatusCoarse2FineGhost_,
                                                coarse);
    } else if (time_interpolation_) {
        ptr_grid_info->TransferInfoFromCoarseGrid(sfbitset_aux, amrproject::NodeBitStatus::kNodeStatusCoarse2FineGhost_,
                                                  *coarse);
    }
}

void SolverLbmHip::CopyPopulationsToHost(DefInt level) {
    if (!ctx_) throw ZlhError("device state not built");
    amrproject::GridInfoInterface &gi = *ptr_grid_manager_->vec_ptr_grid_info_.at(level);
    std::vector<uint64_t> keys;
    keys.reserve(gi.map_grid_node_.size());
    for (auto &kv : gi.map_grid_node_) keys.push_back(kv.first.to_ullong());
    std::sort(keys.begin(), keys.end());
    const size_t n = keys.size();
    std::vector<double> f(n * k0NumQ_), rho(n), u(n * k0SolverDims_);
    Check(zlh_download_populations(ctx_, level, f.data()), "zlh_download_populations");
    Check(zlh_download_macros(ctx_, level, rho.data(), u.data()), "zlh_download_macros");
    for (size_t i = 0; i < n; ++i) {
        auto *node = dynamic_cast<GridNodeLbmHip *>(gi.map_grid_node_.at(DefSFBitset(keys[i])).get());
        if (!node) continue;
        node->f_.assign(f.begin() + i * k0NumQ_, f.begin() + (i + 1) * k0NumQ_);
        node->rho_ = rho[i];
        node->velocity_.assign(u.begin() + i * k0SolverDims_, u.begin() + (i + 1) * k0SolverDims_);
    }
}

void SolverLbmHip::WriteCheckPointGridNodes(const std::string &file_name) {
    if (!ctx_) BuildDevice();
    amrproject::GridManagerInterface &gm = *ptr_grid_manager_;
    const int dims = k0SolverDims_;
    std::vector<CheckpointOrder> order;
    for (DefInt l = 0; l <= gm.GetMaxLevel(); ++l) {
        CopyPopulationsToHost(l);
        const uint32_t mask = (1u << l) - 1u;
        for (auto &kv : gm.vec_ptr_grid_info_.at(l)->map_grid_node_) {
            const uint64_t key = kv.first.to_ullong();
            order.push_back({key >> (dims * l), (int32_t)l, dims == 3 ? morton_axis(key, dims, 2) & mask : 0u,
                             morton_axis(key, dims, 1) & mask, morton_axis(key, dims, 0) & mask, key});
        }
    }
    std::sort(order.begin(), order.end());
    if (order.size() > 0xffffffffull) throw ZlhError("checkpoint: more nodes than the record count (uint32) holds");
    const int32_t nforce = bool_forces_ ? k0NumForces_ : 0;
    const int32_t payload = (int32_t)(sizeof(DefInt) + sizeof(double) * (k0NumQ_ + nforce));
    std::ofstream out(file_name, std::ios::binary);
    if (!out) throw ZlhError("cannot write checkpoint " + file_name);
    const uint32_t count = (uint32_t)order.size();   // DefAmrLUint
    out.write(reinterpret_cast<const char *>(&count), sizeof(count));
    out.write(reinterpret_cast<const char *>(&payload), sizeof(payload));
    // node forces of every level in key order (GridNodeLbm::force_)
    std::vector<std::vector<uint64_t>> level_keys(gm.GetMaxLevel() + 1);
    std::vector<std::vector<double>> level_force(gm.GetMaxLevel() + 1);
    if (nforce) {
        if (nforce > 3) throw ZlhError("checkpoint: more than 3 node forces");
        for (DefInt l = 0; l <= gm.GetMaxLevel(); ++l) {
            for (auto &kv : gm.vec_ptr_grid_info_.at(l)->map_grid_node_) level_keys[l].push_back(kv.first.to_ullong());
            std::sort(level_keys[l].begin(), level_keys[l].end());
            level_force[l].resize(level_keys[l].size() * 3);
            Check(zlh_download_node_forces(ctx_, l, level_force[l].data()), "zlh_download_node_forces");
        }
    }
    for (const CheckpointOrder &r : order) {
        const auto *ptr = dynamic_cast<const GridNodeLbmHip *>(
            gm.vec_ptr_grid_info_.at(r.level)->map_grid_node_.at(DefSFBitset(r.key)).get());
        if (!ptr || (int)ptr->f_.size() != k0NumQ_) throw ZlhError("checkpoint: a node is not an LBM node");
        const GridNodeLbmHip &node = *ptr;
        const int32_t flag = node.flag_status_;
        out.write(reinterpret_cast<const char *>(&r.key), sizeof(uint64_t));
        out.write(reinterpret_cast<const char *>(&r.level), sizeof(int32_t));
        out.write(reinterpret_cast<const char *>(&flag), sizeof(int32_t));
        out.write(reinterpret_cast<const char *>(node.f_.data()), sizeof(double) * k0NumQ_);
        if (nforce) {
            const auto &ks = level_keys[r.level];
            const size_t i = (size_t)(std::lower_bound(ks.begin(), ks.end(), r.key) - ks.begin());
            out.write(reinterpret_cast<const char *>(level_force[r.level].data() + 3 * i), sizeof(double) * nforce);
        }
    }
    if (!out) throw ZlhError("checkpoint " + file_name + ": write failed");
}

namespace {
template <typename V>
void put(std::vector<char> &b, const V &v) {
    const char *p = reinterpret_cast<const char *>(&v);
    b.insert(b.end(), p, p + sizeof(V));
}

// the nodes of a level inside one background cell in the grid_nodes record order (z, y, x)
std::vector<uint64_t> cell_keys(int dims, uint64_t cell, int level) {
    const uint32_t n = 1u << level;
    const uint32_t base[3] = {morton_axis(cell, dims, 0) << level, morton_axis(cell, dims, 1) << level,
                              dims == 3 ? morton_axis(cell, dims, 2) << level : 0u};
    std::vector<uint64_t> out;
    for (uint32_t z = 0; z < (dims == 3 ? n : 1u); ++z)
        for (uint32_t y = 0; y < n; ++y)
            for (uint32_t x = 0; x < n; ++x) out.push_back(morton_key(dims, base[0] + x, base[1] + y, base[2] + z));
    return out;
}

// the four layer kinds of an interface in the order the reference writes them, with their
// indicators (kIndictor*, mpi/mpi_manager.h:615-616)
std::vector<std::pair<int32_t, const std::vector<DefMap<DefInt>> *>> layer_kinds(const amrproject::InterfaceLayerInfo &li) {
    return {{1, &li.vec_inner_fine2coarse_}, {2, &li.vec_outer_fine2coarse_},
            {3, &li.vec_inner_coarse2fine_}, {4, &li.vec_outer_coarse2fine_}};
}

struct CellLoad {
    std::vector<uint8_t> levels;
    int64_t nodes = 0, interface_nodes = 0, load = 0;
};

std::map<uint64_t, CellLoad> cell_loads(const amrproject::GridManagerInterface &gm, int dims) {
    std::map<uint64_t, CellLoad> cells;
    const size_t nbytes = (size_t)(gm.GetMaxLevel() + 1 + 7) / 8;
    for (DefInt l = 0; l <= gm.GetMaxLevel(); ++l) {
        const auto &gi = *gm.vec_ptr_grid_info_.at(l);
        for (const auto &kv : gi.map_grid_node_) {
            CellLoad &c = cells[kv.first.to_ullong() >> (dims * l)];
            if (c.levels.empty()) c.levels.assign(nbytes, 0);
            c.levels[l / 8] |= (uint8_t)(1u << (l % 8));
            ++c.nodes;
            c.load += (int64_t)1 << l;
        }
        for (const auto &it : gi.map_ptr_interface_layer_info_)
            for (const auto &kind : layer_kinds(*it.second))
                for (const auto &layer : *kind.second)
                    for (const auto &kv : layer) {
                        if (!gi.map_grid_node_.count(kv.first))
                            throw ZlhError("checkpoint: an interface node is not a grid node");
                        ++cells.at(kv.first.to_ullong() >> (dims * l)).interface_nodes;
                    }
    }
    return cells;
}

std::vector<char> read_file(const std::string &name) {
    std::ifstream in(name, std::ios::binary);
    if (!in) throw ZlhError("cannot read checkpoint " + name);
    return std::vector<char>((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
}

void write_file(const std::string &name, const std::vector<char> &b) {
    std::ofstream out(name, std::ios::binary);
    if (!out) throw ZlhError("cannot write checkpoint " + name);
    out.write(b.data(), (std::streamsize)b.size());
    if (!out) throw ZlhError("checkpoint " + name + ": write failed");
}
}  // namespace

std::vector<char> SolverLbmHip::CheckPointLoadings() const {
    const amrproject::GridManagerInterface &gm = *ptr_grid_manager_;
    const auto cells = cell_loads(gm, k0SolverDims_);
    int64_t total = 0;
    for (const auto &c : cells) total += c.second.load;
    if (total > 0xffffffffll || cells.size() > 0xffffffffull)
        throw ZlhError("checkpoint: load or cell count beyond the loadings header's uint32");
    std::vector<char> b;
    put(b, (uint32_t)total);
    put(b, (uint32_t)cells.size());
    put(b, (int32_t)gm.GetMaxLevel());
    for (const auto &c : cells) {
        if (c.second.nodes > INT32_MAX || c.second.interface_nodes > INT32_MAX || c.second.load > INT32_MAX)
            throw ZlhError("checkpoint: a cell's counts exceed int32");
        put(b, c.first);
        b.insert(b.end(), c.second.levels.begin(), c.second.levels.end());
        put(b, (int32_t)c.second.nodes);
        put(b, (int32_t)c.second.interface_nodes);
        put(b, (int32_t)c.second.load);
    }
    return b;
}

std::vector<char> SolverLbmHip::CheckPointInterfaceNodes() const {
    const amrproject::GridManagerInterface &gm = *ptr_grid_manager_;
    const int dims = k0SolverDims_;
    // index table: every layer of every interface of every level, in order
    std::map<std::tuple<int32_t, int32_t, int32_t, int32_t, int32_t>, int32_t> index;
    std::vector<char> table;
    int32_t nidx = 0;
    for (DefInt l = 0; l <= gm.GetMaxLevel(); ++l)
        for (const auto &it : gm.vec_ptr_grid_info_.at(l)->map_ptr_interface_layer_info_)
            for (const auto &kind : layer_kinds(*it.second))
                for (int32_t layer = 0; layer < (int32_t)kind.second->size(); ++layer) {
                    const auto entry = std::make_tuple((int32_t)l, (int32_t)it.first.first, (int32_t)it.first.second,
                                                       kind.first, layer);
                    index[entry] = nidx++;
                    put(table, std::get<0>(entry));
                    put(table, std::get<1>(entry));
                    put(table, std::get<2>(entry));
                    put(table, std::get<3>(entry));
                    put(table, std::get<4>(entry));
                }
    std::vector<char> b;
    put(b, nidx);
    b.insert(b.end(), table.begin(), table.end());
    for (const auto &c : cell_loads(gm, dims))
        for (DefInt l = 0; l <= gm.GetMaxLevel(); ++l) {
            if (!(c.second.levels[l / 8] >> (l % 8) & 1u)) continue;
            const auto &gi = *gm.vec_ptr_grid_info_.at(l);
            for (uint64_t key : cell_keys(dims, c.first, l))
                for (const auto &it : gi.map_ptr_interface_layer_info_)
                    for (const auto &kind : layer_kinds(*it.second))
                        for (int32_t layer = 0; layer < (int32_t)kind.second->size(); ++layer)
                            if (kind.second->at(layer).count(DefSFBitset(key))) {
                                put(b, key);
                                put(b, index.at(std::make_tuple((int32_t)l, (int32_t)it.first.first,
                                                                (int32_t)it.first.second, kind.first, layer)));
                                break;
                            }
        }
    return b;
}

void SolverLbmHip::WriteCheckPoint(const std::string &folder) {
    std::error_code ec;
    std::filesystem::create_directories(folder, ec);
    if (ec) throw ZlhError("cannot create checkpoint folder " + folder + ": " + ec.message());
    write_file(folder + "/loadings", CheckPointLoadings());
    WriteCheckPointGridNodes(folder + "/grid_nodes");
    write_file(folder + "/interface_nodes", CheckPointInterfaceNodes());
}

void SolverLbmHip::ReadCheckPoint(const std::string &folder) {
    if (read_file(folder + "/loadings") != CheckPointLoadings())
        throw ZlhError("checkpoint " + folder + ": loadings do not describe this grid");
    if (read_file(folder + "/interface_nodes") != CheckPointInterfaceNodes())
        throw ZlhError("checkpoint " + folder + ": interface_nodes do not describe this grid's interfaces");
    ReadCheckPointGridNodes(folder + "/grid_nodes");
}

void SolverLbmHip::ReadCheckPointGridNodes(const std::string &file_name) {
    amrproject::GridManagerInterface &gm = *ptr_grid_manager_;
    std::ifstream in(file_name, std::ios::binary);
    if (!in) throw ZlhError("cannot read checkpoint " + file_name);
    uint32_t count = 0;
    int32_t payload = 0;
    in.read(reinterpret_cast<char *>(&count), sizeof(count));
    in.read(reinterpret_cast<char *>(&payload), sizeof(payload));
    const int32_t nforce = bool_forces_ ? k0NumForces_ : 0;
    if (!in || payload != (int32_t)(sizeof(DefInt) + sizeof(double) * (k0NumQ_ + nforce)))
        throw ZlhError("checkpoint " + file_name + ": payload size does not match this solver's lattice and forces");
    size_t nodes = 0;
    for (DefInt l = 0; l <= gm.GetMaxLevel(); ++l) nodes += gm.vec_ptr_grid_info_.at(l)->map_grid_node_.size();
    if (count != nodes) throw ZlhError("checkpoint " + file_name + ": node count differs from the grid");
    std::vector<char> rec(sizeof(uint64_t) + sizeof(int32_t) + payload);
    std::vector<std::vector<uint64_t>> seen(gm.GetMaxLevel() + 1);
    for (uint32_t i = 0; i < count; ++i) {
        if (!in.read(rec.data(), (std::streamsize)rec.size())) throw ZlhError("checkpoint " + file_name + " is truncated");
        uint64_t key;
        int32_t level, flag;
        std::memcpy(&key, rec.data(), sizeof(key));
        std::memcpy(&level, rec.data() + 8, sizeof(level));
        std::memcpy(&flag, rec.data() + 12, sizeof(flag));
        if (level < 0 || level > gm.GetMaxLevel()) throw ZlhError("checkpoint " + file_name + ": bad level");
        auto &map = gm.vec_ptr_grid_info_.at(level)->map_grid_node_;
        auto it = map.find(DefSFBitset(key));
        if (it == map.end()) throw ZlhError("checkpoint " + file_name + ": node not in the grid");
        auto *node = dynamic_cast<GridNodeLbmHip *>(it->second.get());
        if (!node) throw ZlhError("checkpoint " + file_name + ": node is not an LBM node");
        seen[level].push_back(key);
        node->flag_status_ = flag;
        node->f_.resize(k0NumQ_);
        std::memcpy(node->f_.data(), rec.data() + 16, sizeof(double) * k0NumQ_);
    }
    // with the count equal to the grid's, a repeated record means a node without one
    for (auto &v : seen) {
        std::sort(v.begin(), v.end());
        if (std::adjacent_find(v.begin(), v.end()) != v.end())
            throw ZlhError("checkpoint " + file_name + ": a node has two records");
    }
    if (!ctx_) return;   // BuildDevice uploads the mirrors
    for (DefInt l = 0; l <= gm.GetMaxLevel(); ++l) {
        auto &map = gm.vec_ptr_grid_info_.at(l)->map_grid_node_;
        std::vector<std::pair<uint64_t, const GridNodeLbmHip *>> nodes_l;
        for (auto &kv : map) nodes_l.emplace_back(kv.first.to_ullong(), dynamic_cast<const GridNodeLbmHip *>(kv.second.get()));
        std::sort(nodes_l.begin(), nodes_l.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
        std::vector<double> f(nodes_l.size() * k0NumQ_);
        for (size_t i = 0; i < nodes_l.size(); ++i) {
            if ((int)nodes_l[i].second->f_.size() != k0NumQ_) throw ZlhError("checkpoint misses a node");
            std::copy(nodes_l[i].second->f_.begin(), nodes_l[i].second->f_.end(), f.begin() + i * k0NumQ_);
        }
        Check(zlh_upload_populations(ctx_, l, f.data()), "zlh_upload_populations");
    }
    for (auto &kv : init_macros_)
        Check(zlh_init_macros(ctx_, kv.first, kv.second.first.data(), kv.second.second.data()), "zlh_init_macros");
}

void SolverLbmHip::WriteVtu(const std::string &prefix) {
    if (!ctx_) BuildDevice();
    amrproject::GridManagerInterface &gm = *ptr_grid_manager_;
    const int dims = k0SolverDims_;
    std::vector<std::string> pieces;
    for (DefInt l = 0; l <= gm.GetMaxLevel(); ++l) {
        CopyPopulationsToHost(l);
        auto &map = gm.vec_ptr_grid_info_.at(l)->map_grid_node_;
        std::vector<uint64_t> keys;
        keys.reserve(map.size());
        for (auto &kv : map) keys.push_back(kv.first.to_ullong());
        std::sort(keys.begin(), keys.end());
        auto index_of = [&](uint64_t k) -> int64_t {
            auto it = std::lower_bound(keys.begin(), keys.end(), k);
            return (it != keys.end() && *it == k) ? (int64_t)(it - keys.begin()) : -1;
        };
        // cells: corner order of VTK_PIXEL (x fastest, then y) and VTK_VOXEL (then z)
        std::vector<int64_t> conn;
        const int ncorner = dims == 3 ? 8 : 4;
        for (uint64_t k : keys) {
            const uint32_t x = morton_axis(k, dims, 0), y = morton_axis(k, dims, 1);
            const uint32_t z = dims == 3 ? morton_axis(k, dims, 2) : 0u;
            int64_t c[8];
            bool full = true;
            for (int v = 0; v < ncorner && full; ++v) {
                c[v] = index_of(morton_key(dims, x + (v & 1), y + ((v >> 1) & 1), z + ((v >> 2) & 1)));
                full = c[v] >= 0;
            }
            if (full) conn.insert(conn.end(), c, c + ncorner);
        }
        const size_t ncell = conn.size() / ncorner;
        const std::string file = prefix + "_level" + std::to_string(l) + ".vtu";
        std::FILE *fp = std::fopen(file.c_str(), "w");
        if (!fp) throw ZlhError("cannot write " + file);
        std::fprintf(fp, "<?xml version=\"1.0\"?>\n<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n");
        std::fprintf(fp, " <UnstructuredGrid>\n  <Piece NumberOfPoints=\"%zu\" NumberOfCells=\"%zu\">\n", keys.size(), ncell);
        std::fprintf(fp, "    <PointData>\n      <DataArray type=\"Int32\" Name=\"flag\" format=\"ascii\">\n");
        for (uint64_t k : keys) std::fprintf(fp, " %d", (int)map.at(DefSFBitset(k))->flag_status_);
        std::fprintf(fp, "\n      </DataArray>\n      <DataArray type=\"Float32\" Name=\"rho\" format=\"ascii\">\n");
        std::vector<const GridNodeLbmHip *> nodes;
        nodes.reserve(keys.size());
        for (uint64_t k : keys) {
            const auto *n = dynamic_cast<const GridNodeLbmHip *>(map.at(DefSFBitset(k)).get());
            if (!n || (int)n->velocity_.size() < dims) {
                std::fclose(fp);
                throw ZlhError("VTU output: a node of level " + std::to_string(l) + " is not an LBM node");
            }
            nodes.push_back(n);
        }
        for (const GridNodeLbmHip *n : nodes) std::fprintf(fp, " %.9g", (double)(float)n->rho_);
        std::fprintf(fp, "\n      </DataArray>\n      <DataArray type=\"Float32\" Name=\"velocity\" NumberOfComponents=\"%d\" format=\"ascii\">\n", dims);
        for (const GridNodeLbmHip *n : nodes)
            for (int a = 0; a < dims; ++a) std::fprintf(fp, " %.9g", (double)(float)n->velocity_[a]);
        std::fprintf(fp, "\n      </DataArray>\n    </PointData>\n");
        std::fprintf(fp, "   <Points>\n    <DataArray type=\"Float32\" NumberOfComponents=\"3\" format=\"ascii\">\n");
        const double h = 1.0 / (double)(1u << l);
        for (uint64_t k : keys)
            std::fprintf(fp, " %.9g %.9g %.9g", (double)(float)(morton_axis(k, dims, 0) * h),
                         (double)(float)(morton_axis(k, dims, 1) * h),
                         dims == 3 ? (double)(float)(morton_axis(k, dims, 2) * h) : 0.);
        std::fprintf(fp, "\n    </DataArray>\n   </Points>\n   <Cells>\n");
        std::fprintf(fp, "    <DataArray type=\"Int64\" Name=\"connectivity\" format=\"ascii\">\n");
        for (int64_t c : conn) std::fprintf(fp, " %lld", (long long)c);
        std::fprintf(fp, "\n    </DataArray>\n    <DataArray type=\"Int64\" Name=\"offsets\" format=\"ascii\">\n");
        for (size_t i = 1; i <= ncell; ++i) std::fprintf(fp, " %lld", (long long)(i * ncorner));
        std::fprintf(fp, "\n    </DataArray>\n    <DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\">\n");
        for (size_t i = 0; i < ncell; ++i) std::fprintf(fp, " %d", dims == 3 ? 11 : 8);
        std::fprintf(fp, "\n    </DataArray>\n   </Cells>\n  </Piece>\n </UnstructuredGrid>\n</VTKFile>\n");
        if (std::fclose(fp) != 0) throw ZlhError("write failed: " + file);
        pieces.push_back(file.substr(file.find_last_of('/') + 1));
    }
    const std::string pfile = prefix + ".pvtu";
    std::FILE *fp = std::fopen(pfile.c_str(), "w");
    if (!fp) throw ZlhError("cannot write " + pfile);
    std::fprintf(fp, "<?xml version=\"1.0\"?>\n<VTKFile type=\"PUnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n");
    std::fprintf(fp, " <PUnstructuredGrid GhostLevel=\"0\">\n  <PPointData>\n");
    std::fprintf(fp, "   <PDataArray type=\"Int32\" Name=\"flag\"/>\n   <PDataArray type=\"Float32\" Name=\"rho\"/>\n");
    std::fprintf(fp, "   <PDataArray type=\"Float32\" Name=\"velocity\" NumberOfComponents=\"%d\"/>\n  </PPointData>\n", dims);
    std::fprintf(fp, "  <PPoints>\n   <PDataArray type=\"Float32\" NumberOfComponents=\"3\"/>\n  </PPoints>\n");
    for (const std::string &p : pieces) std::fprintf(fp, "  <Piece Source=\"%s\"/>\n", p.c_str());
    std::fprintf(fp, " </PUnstructuredGrid>\n</VTKFile>\n");
    if (std::fclose(fp) != 0) throw ZlhError("write failed: " + pfile);
}

// AmrManager::TimeMarching (amr_manager.cpp:279-300)
void TimeMarching(amrproject::GridManagerInterface *gm, DefAmrLUint time_step_background) {
    amrproject::MultiTimeSteppingC2F scheme(gm->GetMaxLevel());
    std::vector<DefInt> time_step_level(gm->GetMaxLevel() + 1, 0);
    for (DefInt i_level : scheme.k0TimeSteppingOrder_) {
        ++time_step_level[i_level];
        const DefReal t = scheme.GetCurrentTimeStep(i_level, time_step_background, time_step_level[i_level]);
        amrproject::GridInfoInterface &g = *gm->vec_ptr_grid_info_.at(i_level);
        g.SetUpGridAtBeginningOfTimeStep(time_step_level[i_level]);
        g.AdvancingAtCurrentTime(amrproject::ETimeSteppingScheme::kMultiSteppingC2F, time_step_level[i_level], t,
                                 nullptr, nullptr);
    }
}

}  // namespace lbmproject
}  // namespace rootproject
