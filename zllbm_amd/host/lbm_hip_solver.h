// lbm_hip_solver.h -- LBM solver plugin whose time step runs on the MI355X through the C ABI
// of include/zllbm_hip.h.  Mirrors the reference's solver manager SolverLbmInterface and grid
// information GridInfoLbmInteface (lbm_project/lbm_interface.h:140-474): the same setters, the
// same plugin virtuals, and the node maps kept as host mirrors filled on request.
//
// Interface declarations: by default those of amr_interface.h (this directory).  With
// ZLLBM_HIP_REFERENCE_HEADERS defined, the reference's own headers instead, found on its include
// paths (source/, source/amr_project, source/lbm_project), so that the plugin links into the
// reference's AmrManager with one declaration of each interface class.  The code below uses only
// members that both declare alike.
#pragma once
#include <array>
#include <map>
#include <memory>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef ZLLBM_HIP_REFERENCE_HEADERS
#include "grid/grid_manager.h"                              // GridManagerInterface, MultiTimeSteppingC2F
#include "grid/grid_info_interface.h"                       // GridInfoInterface, GridNode
#include "solver_info_interface.h"                          // SolverInterface
#include "collision_model/lbm_collision_models.h"           // ELbmCollisionOperatorType
#include "boundary_conditions/lbm_boundary_conditions.h"    // ELbmBoundaryConditionScheme
#include "les_models/lbm_les_models.h"                      // ELbmLesModelType
#else
#include "amr_interface.h"
#endif
#include "zllbm_hip.h"

namespace rootproject {
namespace lbmproject {

#ifndef ZLLBM_HIP_REFERENCE_HEADERS
// collision_model/lbm_collision_models.h:21-25, boundary_conditions/lbm_boundary_conditions.h:22-27,
// les_models/lbm_les_models.h:20-23
enum class ELbmCollisionOperatorType { kLbmSrt, kLbmMrt, kUndefined };
enum class ELbmBoundaryConditionScheme { kUndefined = 0, kBounceBack = 1, kPeriodic = 2, kNonEqExtrapolation = 3 };
enum class ELbmLesModelType { kSmagorinsky, kUndefined };
#endif

// Fatal errors of the device library surface as this exception (the reference's
// LogManager::LogError terminates; callers that want that behaviour let it propagate).
class ZlhError : public std::runtime_error {
 public:
    using std::runtime_error::runtime_error;
};

// host mirror of one node (GridNodeLbm, lbm_interface.h:34-194)
struct GridNodeLbmHip : public amrproject::GridNode {
    DefReal rho_ = 1.;
    std::vector<DefReal> velocity_, f_, force_;
};

// per-direction boundary setting (BoundaryConditionLbmInterface::SetValues)
struct BoundarySetting {
    ELbmBoundaryConditionScheme boundary_scheme_ = ELbmBoundaryConditionScheme::kUndefined;
    std::array<DefReal, 3> values_{0., 0., 0.};
    void SetValues(const std::vector<DefReal> &values) {
        for (size_t i = 0; i < values.size() && i < 3; ++i) values_[i] = values[i];
    }
};

class SolverLbmHip;

class GridInfoLbmHip : public amrproject::GridInfoInterface {
 public:
    std::map<amrproject::EDomainBoundaryDirection, std::unique_ptr<BoundarySetting>> domain_boundary_condition_;
    void AdvancingAtCurrentTime(const amrproject::ETimeSteppingScheme time_scheme, const DefInt time_step_level,
                                const DefReal time_step_current, amrproject::MpiManager *const ptr_mpi_manager,
                                amrproject::CriterionManager *const ptr_criterion_manager) override;
    std::unique_ptr<amrproject::GridNode> GridNodeCreator() const override;
    void SetNodeVariablesAsZeros(amrproject::GridNode *const ptr_node) override;
    void InitialGridInfo(const DefInt dims) override { dims_ = dims; }
    // TransferInfoFromCoarseGrid / TransferInfoToCoarseGrid (lbm_grid_interface.cpp:408-564) on
    // the device: zlh_transfer_from_coarse / zlh_transfer_to_coarse of this (fine) level; 0 on
    // success as in the reference
    int TransferInfoFromCoarseGrid(const amrproject::SFBitsetAuxInterface &sfbitset_aux, const DefInt node_flag,
                                   const amrproject::GridInfoInterface &grid_info_coarse) override;
    int TransferInfoToCoarseGrid(const amrproject::SFBitsetAuxInterface &sfbitset_aux, const DefInt node_flag,
                                 amrproject::GridInfoInterface *const ptr_grid_info_coarse) override;
    // NodeInfoCoarse2fine / NodeInfoFine2Coarse (lbm_grid_interface.cpp:570-610) on the host
    // mirrors (GridNodeLbmHip): feq of the source's macros plus the tau-ratio-scaled
    // non-equilibrium part (PostStreamCoarse2Fine / PostStreamFine2Coarse,
    // lbm_collision_models.cpp:244-267); the device kernels do the same for the time step
    void NodeInfoCoarse2fine(const amrproject::GridNode &coarse_node, amrproject::GridNode *const ptr_fine_node) const override;
    void NodeInfoFine2Coarse(const amrproject::GridNode &fine_node, amrproject::GridNode *const ptr_coarse_node) const override;
    // node record sizes (lbm_grid_interface.cpp:737-749, lbm_grid_interface_mpi.cpp:77-90)
    int GetSizeOfGridNodeInfoForMpiCommunication() const override;
    int GetSizeOfGridNodeInfoForCheckPoint() const override;
    // names of the output variables (GridInfoLbmInteface::output_variable_name_, lbm_interface.h:262)
    std::set<std::string> output_variable_name_ = {"rho", "velocity"};
    // the names in output_variable_name_ ("rho", "velocity") as output variables
    // (SetupOutputVariables, lbm_grid_interface.cpp:614-660); WriteVtu writes them
    void SetupOutputVariables() override;

 private:
    DefInt dims_ = 0;
};

class GridInfoLbmHipCreator : public amrproject::GridInfoCreatorInterface {
 public:
    std::shared_ptr<amrproject::GridInfoInterface> CreateGridInfo() const override {
        return std::make_shared<GridInfoLbmHip>();
    }
};

class SolverLbmHip : public amrproject::SolverInterface {
 public:
    SolverLbmHip(DefInt dims, DefInt num_q);
    ~SolverLbmHip() override;

    // SolverLbmInterface setters (lbm_interface.h:296-319)
    void SetDefaultViscosity(const DefReal v) { k0LbmViscosity_ = v; }
    void SetDefaultDensity(const DefReal rho) { k0Rho_ = rho; }
    void SetDefaultVelocity(const std::vector<DefReal> &u) { k0Velocity_ = u; }
    void SetConstantForce(const std::vector<DefReal> &force) { bool_forces_ = true; k0ConstForce_ = force; }
    void SetNumForces(const DefInt n) { bool_forces_ = true; k0NumForces_ = n; }
    DefReal GetDefaultViscosity() const { return k0LbmViscosity_; }
    // GetNumForces (lbm_interface.h:308-314): the node force count when forces are on
    DefInt GetNumForces() const { return bool_forces_ ? k0NumForces_ : 0; }
    DefReal GetDefaultDensity() const { return k0Rho_; }
    // LES (SetLesModel, les_models/lbm_les_models.cpp:17-41) with the Smagorinsky constant and the
    // relaxation-time mode exposed
    void SetLesModel(const ELbmLesModelType type) {
        les_ = type == ELbmLesModelType::kSmagorinsky ? ZLH_LES_SMAGORINSKY : ZLH_LES_NONE;
    }
    // "smagorinsky" or "learned" (the learned sub-grid closure of zlh_mlp_set, which has no
    // counterpart in the reference's ELbmLesModelType)
    void SetLesModel(const std::string &type);
    // the network of the learned closure (zlh_mlp_set: 1409 parameters, then 9 each of input mean,
    // spread, training minimum and maximum) and its output scale; Smagorinsky (physical tau mode,
    // the constant of SetSmagorinskyConstant) is its fallback outside the training range
    void SetLearnedClosure(const std::vector<float> &params, const std::vector<float> &mean,
                           const std::vector<float> &std_dev, const std::vector<float> &fmin,
                           const std::vector<float> &fmax, DefReal nu_scale);
    // operands of the FP32 solver's matrix-core network: ZLH_MLP_OPERANDS_FP16 (default) or _BF16
    void SetMlpOperands(const int operands) { mlp_operands_ = operands; }
    void SetSmagorinskyConstant(const DefReal cs) { cs_ = cs; }
    void SetLesTauMode(const int mode) { les_tau_mode_ = mode; }
    // collision operator per level (SetCollisionOperator, collision_model/lbm_collision_models.cpp:78-90)
    void SetCollisionOperator(const DefInt i_level, const ELbmCollisionOperatorType type);
    // SetDomainBoundaryCondition (boundary_conditions/lbm_boundary_conditions.cpp:20-52) for every level
    void SetDomainBoundaryCondition(const amrproject::EDomainBoundaryDirection which,
                                    const ELbmBoundaryConditionScheme scheme, const std::vector<DefReal> &values = {});
    // storage form of the populations on the device
    void SetPrecision(const int precision) { precision_ = precision; }
    // refill fine ghost layers after odd substeps too, from the mean of the coarse level's states
    // before and after its last step (zlh_set_time_interpolation; off is the reference's schedule)
    void SetTimeInterpolation(const bool on) { time_interpolation_ = on; }
    // whole background steps as one graph replay (default on): the first AdvancingAtCurrentTime
    // of a background step (level 0) runs every level's substeps and exchanges of that step
    // through zlh_time_marching, and the calls for the finer levels' substeps that follow, in
    // the coarse-to-fine order (MultiTimeSteppingC2F), only check that order.  Off: each call
    // launches its own substep and exchanges.  Results are identical; between the calls of one
    // background step the device state is that of its end.
    void SetGraphReplay(const bool on);
    // initial macroscopic fields of a level (density, velocity with dims components per node),
    // node by node in ascending Morton key order; the populations start at their equilibrium
    // (zlh_init_macros) once the device is built
    void SetInitialMacros(DefInt level, const std::vector<DefReal> &rho, const std::vector<DefReal> &velocity);
    // immersed-boundary markers of one level (coordinates in background units)
    void SetImmersedBoundary(DefInt level, const std::vector<std::array<DefReal, 3>> &xyz,
                             const std::vector<DefReal> &area, const std::vector<std::array<DefReal, 3>> &velocity,
                             DefInt stencil_dis = 2, DefReal stiffness = 2.);

    const DefInt k0NumQ_;
    bool bool_forces_ = false;
    bool time_interpolation_ = false;

    // SolverInterface
    std::string GetSolverMethod() override;
    void SolverInitial() override;
    void RunSolverOnGivenGrid(const amrproject::ETimeSteppingScheme time_scheme, const DefInt time_step_level,
                              const DefReal time_step_current, const amrproject::SFBitsetAuxInterface &sfbitset_aux,
                              amrproject::GridInfoInterface *const ptr_grid_info) override;
    void InformationFromGridOfDifferentLevel(const DefInt time_step_level,
                                             const amrproject::SFBitsetAuxInterface &sfbitset_aux,
                                             amrproject::GridInfoInterface *const ptr_grid_info) override;

    // device state: built from the grid manager's node maps and interface layers on first use,
    // populations taken from the node mirrors' f_ when set, else feq(rho0, u0)
    void BuildDevice();
    // copy the device populations of a level into the node mirrors (GridNodeLbmHip::f_, rho_, velocity_)
    void CopyPopulationsToHost(DefInt level);
    // checkpoint of every level's nodes in the record format of the reference's grid_nodes file
    // for one rank (MpiManager::WriteCheckPointGridNodes, mpi/mpi_checkpoint.cpp:272-390):
    // header {uint32 node count, int32 payload bytes}, then per node {uint64 Morton key, int32
    // level, payload}, payload = GridNodeLbm::CopyANodeToBufferForCheckpoint
    // (lbm_interface.h:57-67): int32 flag, Q doubles of post-stream populations, and, with forces,
    // GetNumForces() doubles.  Records are ordered as the reference writes them: by background
    // cell (Morton code ascending), then level, then z, y, x inside the cell
    // (SFBitsetHigherLevelInACell, grid/sfbitset_aux.cpp:140-156, 1958-1985).  The force block
    // holds the node forces (the immersed-boundary force of the last forcing step on its
    // supports, zero elsewhere).
    void WriteCheckPointGridNodes(const std::string &file_name);
    // checkpoint folder for one rank, as IoManager::WriteCheckPointData writes it
    // (io/io_checkpoint.cpp:27-51): <folder>/loadings, <folder>/grid_nodes and
    // <folder>/interface_nodes
    void WriteCheckPoint(const std::string &folder);
    // restart from such a folder (IoManager::ReadCheckPointData, io/io_checkpoint.cpp:59-166, on
    // a grid already set): its loadings and interface_nodes must be those of this grid, byte for
    // byte, and grid_nodes restores the populations (ReadCheckPointGridNodes)
    void ReadCheckPoint(const std::string &folder);
    // the loadings file (ComputeComputationalLoadOnEachRank + WriteCheckPointNodesAtWhichLevels,
    // mpi/mpi_checkpoint.cpp:39-190): header {uint32 total load, uint32 background cells, int32
    // max level}, then per background cell that holds a node of any level, by Morton code:
    // {uint64 code, level bit field of (max level + 8) / 8 bytes, int32 nodes, int32 interface-
    // layer entries, int32 load}; a level-l node costs 2^l (grid_manager.cpp:146)
    std::vector<char> CheckPointLoadings() const;
    // the interface_nodes file (WriteCheckPointInterfaceNodes, mpi/mpi_checkpoint.cpp:1214-1410):
    // int32 index count, the index table {int32 level, int32 criterion, int32 criterion count,
    // int32 layer kind (1 inner fine-to-coarse, 2 outer fine-to-coarse, 3 inner coarse-to-fine,
    // 4 outer coarse-to-fine), int32 layer} of every layer of every interface of every level,
    // then per interface node {uint64 key, int32 index} in the grid_nodes order, one record per
    // interface and layer kind naming the first layer that holds the node
    std::vector<char> CheckPointInterfaceNodes() const;
    // ReadCheckPointGridNodes (mpi/mpi_checkpoint.cpp:392-520) for one rank: every record's flag
    // and populations go to the node mirrors (ReadANodeFromBufferForCheckpoint), and then to the
    // device (uploaded now if it is built, else by BuildDevice).  Every node of the grid must
    // have a record, and every record a node.
    void ReadCheckPointGridNodes(const std::string &file_name);
    // VTU output of every level (VtkWriterManager::WriteVtuAll, io/vtk_writer_manager.cpp:220-560,
    // with the LBM variables of GridInfoLbmInteface::SetupOutputVariables / OutputOneVariable,
    // lbm_grid_interface.cpp:614-732): <prefix>_level<l>.vtu per level, an UnstructuredGrid in
    // ASCII with point data flag (Int32), rho and velocity (Float32, dims components), points in
    // background units (Float32, 3 components), and a VTK_PIXEL / VTK_VOXEL cell at every node
    // whose cell's other corners are nodes of the level; <prefix>.pvtu lists the level files.
    // Nodes are written in Morton key order.
    void WriteVtu(const std::string &prefix);
    zlh_ctx *Context() const { return ctx_; }
    // AdvancingAtCurrentTime's graph-replay route: true when the call's substep is (or was, by
    // the background step's first call) advanced as part of a whole background step
    bool AdvanceWithinBackgroundStep(DefInt level, DefInt time_step_level);

 private:
    void Check(zlh_status s, const char *what) const;
    DefReal k0LbmViscosity_ = 1., k0Rho_ = 1.;
    std::vector<DefReal> k0Velocity_, k0ConstForce_;
    DefInt k0NumForces_ = 0;
    int les_ = ZLH_LES_NONE;
    DefReal cs_ = 0.16;   // the reference's fixed constant (lbm_les_models.h:49) unless set
    int les_tau_mode_ = ZLH_LES_TAU_PHYSICAL;
    int precision_ = ZLH_FP64;
    int mlp_operands_ = ZLH_MLP_OPERANDS_FP16;
    struct Closure {
        std::vector<float> params, mean, std_dev, fmin, fmax;
        double nu_scale = 0.;
    } closure_;
    std::vector<ELbmCollisionOperatorType> collision_type_;
    std::map<int, BoundarySetting> bc_;
    struct Ib {
        std::vector<double> xyz, area, vel;
        int stencil = 2;
        double stiffness = 2.;
    };
    std::map<DefInt, Ib> ib_;
    std::map<DefInt, std::pair<std::vector<DefReal>, std::vector<DefReal>>> init_macros_;
    bool graph_replay_ = true;
    std::vector<DefInt> cycle_order_;   // levels of one background step's substeps, in order
    size_t cycle_pos_ = 0;              // next substep of the current background step (0: none begun)
    std::vector<DefInt> cycle_tsl_;     // substeps of each level seen in the current background step
    zlh_ctx *ctx_ = nullptr;
};

class SolverCreatorLbmD2Q9Hip final : public amrproject::SolverCreatorInterface {
 public:
    std::shared_ptr<amrproject::SolverInterface> CreateSolver() const override {
        return std::make_shared<SolverLbmHip>(2, 9);
    }
};
class SolverCreatorLbmD3Q19Hip final : public amrproject::SolverCreatorInterface {
 public:
    std::shared_ptr<amrproject::SolverInterface> CreateSolver() const override {
        return std::make_shared<SolverLbmHip>(3, 19);
    }
};

// AmrManager::TimeMarching (amr_project/amr_manager.cpp:279-300) over a grid manager whose
// levels hold GridInfoLbmHip instances
void TimeMarching(amrproject::GridManagerInterface *grid_manager, DefAmrLUint time_step_background);

}  // namespace lbmproject
}  // namespace rootproject
