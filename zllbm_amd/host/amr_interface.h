// amr_interface.h -- the grid-manager and solver-plugin interfaces on the LBM time-step path,
// re-declared with the reference's names and signatures so that a solver written against
// them drops into the same call sites:
//   SolverInterface / SolverCreatorInterface      (amr_project/solver_info_interface.h:29-72)
//   GridNode, InterfaceLayerInfo, GridInfoInterface, GridInfoCreatorInterface
//                                                 (amr_project/grid/grid_info_interface.h:58-410)
//   MultiTimeSteppingC2F, NodeBitStatus           (amr_project/grid/grid_manager.h:40-73)
// Only the members the time step reads or writes are declared; mesh generation, MPI, IO and
// the geometry criteria stay with the caller.
#pragma once
#include <bitset>
#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

namespace rootproject {
using DefReal = double;
using DefInt = int32_t;
using DefUint = uint32_t;
using DefSizet = size_t;
using DefAmrLUint = uint32_t;
using DefSFCodeToUint = uint64_t;
static constexpr DefInt kSFBitsetBit = 64;
using DefSFBitset = std::bitset<kSFBitsetBit>;
// identity hash of the Morton code (defs_libs.h:37-43)
class HashFunc {
 public:
    DefSFCodeToUint operator()(const DefSFBitset &b) const noexcept { return b.to_ullong(); }
};
template <typename value>
using DefMap = std::unordered_map<DefSFBitset, value, HashFunc>;
static constexpr DefInt kXIndex = 0, kYIndex = 1, kZIndex = 2;

namespace amrproject {
enum class ETimeSteppingScheme { kMultiSteppingC2F = 1 };
enum class EDomainBoundaryDirection {
    kUndefined = 0, kBoundaryXMin = 1, kBoundaryXMax = 2, kBoundaryYMin = 3, kBoundaryYMax = 4,
    kBoundaryZMin = 5, kBoundaryZMax = 6
};
// ECriterionType (criterion/criterion_numerates.h:15-18); its value is written into checkpoints
enum class ECriterionType : int { kUndefined = 0, kGeometry = 1 };

struct NodeBitStatus {
    inline static DefInt kNodeStatus0_ = 0;
    inline static DefInt kNodeStatusCoarse2Fine0_ = 1 << 1;
    inline static DefInt kNodeStatusCoarse2FineGhost_ = 1 << 2;
    inline static DefInt kNodeStatusFine2Coarse0_ = 1 << 3;
    inline static DefInt kNodeStatusFine2CoarseGhost_ = 1 << 4;
    inline static DefInt kNodeStatusMpiPartitionOuter_ = 1 << 5;
    inline static DefInt kNodeStatusMpiPartitionInner_ = 1 << 6;
    inline static DefInt kNodeStatusMpiInterpInner_ = 1 << 7;
};

class SFBitsetAuxInterface {
 public:
    virtual ~SFBitsetAuxInterface() = default;
};
class MpiManager;
class CriterionManager;
class InputParser;
class GridManagerInterface;
class GridInfoInterface;
class SolverInterface;

struct GridNode {
    DefInt flag_status_ = 0;
    virtual ~GridNode() {}
};

// one output variable of a grid's nodes (grid/grid_info_interface.h:93-101): its name and
// component count
struct OutputNodeVariableInfoInterface {
    DefInt variable_dims_ = 1;
    std::string output_name_;
    virtual void WriteNodeVariable(const GridNode &grid_node) {}
    virtual ~OutputNodeVariableInfoInterface() {}
};

class InterfaceLayerInfo {
 public:
    std::vector<DefMap<DefInt>> vec_outer_coarse2fine_, vec_outer_fine2coarse_;
    std::vector<DefMap<DefInt>> vec_inner_coarse2fine_, vec_inner_fine2coarse_;
};

class GridInfoInterface {
 protected:
    DefInt i_level_ = 0;
    std::weak_ptr<SolverInterface> ptr_solver_;
    SFBitsetAuxInterface *ptr_sfbitset_aux_ = nullptr;
    GridManagerInterface *ptr_parent_grid_manager_ = nullptr;
    DefInt k0NumFine2CoarseLayer_ = 3;
    DefInt k0NumCoarse2FineLayer_ = 2;

 public:
    DefInt GetGridLevel() const { return i_level_; }
    void SetGridLevel(const DefInt i_level) { i_level_ = i_level; }
    void SetPtrSolver(const std::weak_ptr<SolverInterface> &ptr_solver) { ptr_solver_ = ptr_solver; }
    std::weak_ptr<SolverInterface> GetPtrToSolver() const { return ptr_solver_; }
    void SetPtrSFBitsetAux(SFBitsetAuxInterface *const p) { ptr_sfbitset_aux_ = p; }
    SFBitsetAuxInterface *GetPtrSFBitsetAux() const { return ptr_sfbitset_aux_; }
    void SetPtrToParentGridManager(GridManagerInterface *p) { ptr_parent_grid_manager_ = p; }
    GridManagerInterface *GetPtrToParentGridManager() const { return ptr_parent_grid_manager_; }
    DefInt GetNumFine2CoarseLayer() const { return k0NumFine2CoarseLayer_; }
    DefInt GetNumCoarse2FineLayer() const { return k0NumCoarse2FineLayer_; }

    virtual void AdvancingAtCurrentTime(const ETimeSteppingScheme time_scheme, const DefInt time_step_level,
                                        const DefReal time_step_current, MpiManager *const ptr_mpi_manager,
                                        CriterionManager *const ptr_criterion_manager) {}
    virtual void InitialGridInfo(const DefInt dims) {}
    virtual std::unique_ptr<GridNode> GridNodeCreator() const { return std::make_unique<GridNode>(); }
    virtual void SetNodeVariablesAsZeros(GridNode *const ptr_node) {}
    virtual void InitialNotComputeNodeFlag() {}
    virtual void SetUpGridAtBeginningOfTimeStep(const DefInt time_step) {}
    // communication between grids of different refinement levels (grid_info_interface.h:296-299)
    virtual int TransferInfoFromCoarseGrid(const SFBitsetAuxInterface &sfbitset_aux, const DefInt node_flag,
                                           const GridInfoInterface &grid_info_coarse) { return -1; }
    virtual int TransferInfoToCoarseGrid(const SFBitsetAuxInterface &sfbitset_aux, const DefInt node_flag,
                                         GridInfoInterface *const ptr_grid_info_coarse) { return -1; }
    // node information across levels (grid_info_interface.h:325-326)
    virtual void NodeInfoCoarse2fine(const GridNode &coarse_node, GridNode *const ptr_fine_node) const {}
    virtual void NodeInfoFine2Coarse(const GridNode &fine_node, GridNode *const ptr_coarse_node) const {}
    // node record sizes (grid_info_interface.h:376-377)
    virtual int GetSizeOfGridNodeInfoForMpiCommunication() const { return sizeof(DefInt); }
    virtual int GetSizeOfGridNodeInfoForCheckPoint() const { return GetSizeOfGridNodeInfoForMpiCommunication(); }
    // output variables (grid_info_interface.h:278-279)
    std::vector<std::unique_ptr<OutputNodeVariableInfoInterface>> output_variables_;
    virtual void SetupOutputVariables() {}

    std::map<std::pair<ECriterionType, DefInt>, std::shared_ptr<InterfaceLayerInfo>> map_ptr_interface_layer_info_;
    DefMap<std::unique_ptr<GridNode>> map_grid_node_{};
    static constexpr DefInt kFlagInsideDomain_ = 0, kFlagOutsideDomain_ = -1, kFlagXMinBoundary_ = 1,
        kFlagXMaxBoundary_ = 2, kFlagYMinBoundary_ = 4, kFlagYMaxBoundary_ = 8, kFlagZMinBoundary_ = 16,
        kFlagZMaxBoundary_ = 32;
    std::vector<DefSFBitset> k0VecBitsetDomainMin_, k0VecBitsetDomainMax_;
    std::vector<DefMap<DefInt>> domain_boundary_min_, domain_boundary_max_;
    virtual ~GridInfoInterface() {}
};

class GridInfoCreatorInterface {
 public:
    virtual std::shared_ptr<GridInfoInterface> CreateGridInfo() const = 0;
    virtual ~GridInfoCreatorInterface() {}
};

class SolverInterface {
 protected:
    DefInt k0SolverDims_ = 0;
    std::string solver_type_, name_;
    GridManagerInterface *ptr_grid_manager_ = nullptr;

 public:
    void SetSolverDims(const DefInt k0SolverDims) { k0SolverDims_ = k0SolverDims; }
    void SetName(const std::string &name) { name_ = name; }
    void SetPtrToGridManager(GridManagerInterface *const ptr_grid_manager) { ptr_grid_manager_ = ptr_grid_manager; }
    DefInt GetSolverDim() const { return k0SolverDims_; }
    std::string GetSolverType() const { return solver_type_; }
    std::string GetName() const { return name_; }
    GridManagerInterface *GetPtrToParentGridManager() const { return ptr_grid_manager_; }

    std::unique_ptr<GridInfoCreatorInterface> ptr_grid_info_creator_;
    virtual std::string GetSolverMethod() = 0;
    virtual void SolverInitial() = 0;
    virtual void RunSolverOnGivenGrid(const ETimeSteppingScheme time_scheme, const DefInt time_step_level,
                                      const DefReal time_step_current, const SFBitsetAuxInterface &sfbitset_aux,
                                      GridInfoInterface *const ptr_grid_info) = 0;
    virtual void InformationFromGridOfDifferentLevel(const DefInt time_step_current,
                                                     const SFBitsetAuxInterface &sfbitset_aux,
                                                     GridInfoInterface *const ptr_grid_info) {}
    virtual void ReadAndSetupSolverParameters(InputParser *const ptr_input_parser) {}
    virtual ~SolverInterface() = default;
};

class SolverCreatorInterface {
 public:
    virtual std::shared_ptr<SolverInterface> CreateSolver() const = 0;
    virtual ~SolverCreatorInterface() {}
};

// The parts of GridManagerInterface the time step reads (grid/grid_manager.h:80-160): dimensions,
// levels, the per-level grid information, the solvers and the background node index range.  The
// reference's class is abstract (GridManager2D / GridManager3D implement the index range); this
// one stores the range, set with SetIndexRangeOfBackgroundNode.
class GridManagerInterface {
 public:
    const DefInt k0GridDims_;
    std::vector<std::shared_ptr<SolverInterface>> vec_ptr_solver_;
    std::vector<std::shared_ptr<GridInfoInterface>> vec_ptr_grid_info_;
    explicit GridManagerInterface(DefInt dims) : k0GridDims_(dims) {}
    void SetMaxLevel(const DefInt max_level) { k0MaxLevel_ = max_level; }
    DefInt GetMaxLevel() const { return k0MaxLevel_; }
    virtual std::vector<DefAmrLUint> GetMinIndexOfBackgroundNodeArrAsVec() const { return min_index_; }
    virtual std::vector<DefAmrLUint> GetMaxIndexOfBackgroundNodeArrAsVec() const { return max_index_; }
    void SetIndexRangeOfBackgroundNode(const std::vector<DefAmrLUint> &lo, const std::vector<DefAmrLUint> &hi) {
        min_index_ = lo;
        max_index_ = hi;
    }
    virtual ~GridManagerInterface() {}

 protected:
    DefInt k0MaxLevel_ = 0;

 private:
    std::vector<DefAmrLUint> min_index_, max_index_;
};

// Coarse-to-fine multi time stepping (grid/grid_manager.cpp:997-1024): a level step is
// followed by two substeps of the next finer level.
class MultiTimeSteppingC2F {
 public:
    std::vector<DefInt> k0TimeSteppingOrder_;
    explicit MultiTimeSteppingC2F(const DefInt max_level) { Append(0, max_level); }
    DefReal GetCurrentTimeStep(const DefInt i_level, const DefAmrLUint time_step_background,
                               const DefInt time_step_level) const {
        return time_step_background + static_cast<DefReal>(time_step_level) / static_cast<DefReal>(1u << i_level);
    }

 private:
    void Append(DefInt level, DefInt max_level) {
        k0TimeSteppingOrder_.push_back(level);
        if (level < max_level) { Append(level + 1, max_level); Append(level + 1, max_level); }
    }
};

}  // namespace amrproject
}  // namespace rootproject
