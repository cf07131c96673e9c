// state.hpp -- per-level device state of the time step (brick layout, deferred-node arrays,
// boundary, transfer and immersed-boundary tables).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <map>
#include <unordered_map>
#include <vector>
#include "kernels.hpp"

namespace zlh {

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) {
            if (p) (void)hipFree(p);
            p = o.p; bytes = o.bytes;
            o.p = nullptr; o.bytes = 0;
        }
        return *this;
    }
    ~DevBuf() { if (p) (void)hipFree(p); }
    template <typename X> X *as() const { return static_cast<X *>(p); }
};

// Coarse-to-fine interpolation and fine-to-coarse copy of one refined level; the tables live on
// the fine level's state (TransferInfoFromCoarseGrid / TransferInfoToCoarseGrid,
// lbm_grid_interface.cpp:408-564).
// one interpolation workgroup: its ghost nodes [t0, t1) of c2f_target, their contributor words
// [c0, c1) and the rows [r0, r1) of group_rows it stages
struct GatherGroup {
    int32_t t0, t1, c0, c1, r0, r1, pad0, pad1;
};

struct TransferTables {
    int64_t n_c2f = 0;          // fine ghost nodes interpolated from the coarse level
    int64_t n_src = 0;          // coarse nodes feeding the interpolation
    int64_t n_f2c = 0;          // coarse nodes copied from the fine level
    DevBuf c2f_target;          // [n_c2f] fine aux index, gather order
    DevBuf c2f_begin;           // [n_c2f + 1] offsets into the contributor list
    DevBuf c2f_contrib;         // contributor word {LDS offset of the row, FP16 weight bits}
    int64_t n_groups = 0;       // gather workgroups
    int64_t staged_rows = 0;    // rows staged by all workgroups
    int64_t n_contrib = 0;      // contributors of all ghost nodes
    int64_t n_fallback = 0;     // contributors read from a fine fallback node
    size_t lds_bytes = 0;       // rows plus contributor words of the largest group
    DevBuf groups;              // [n_groups] GatherGroup
    DevBuf group_rows;          // rows of src_f each group stages, first use order
    int64_t n_fb = 0;           // fine nodes standing in for missing coarse sources
    DevBuf src_coarse_ai;       // [n_src] coarse aux index of each source row (ascending)
    DevBuf fb_fine_ai;          // [n_fb] fine aux index of each fallback row (ascending)
    DevBuf src_f;               // [n_src + n_fb][Q] rescaled coarse rows, then fallback copies
                                // (storage form), as of the last synchronised transfer
    DevBuf src_mid;             // the same rows of a time-interpolated refill (coarse rows: mean
                                // of src_f and the current ones)
    DevBuf f2c_coarse_ai;       // [n_f2c]
    DevBuf f2c_fine_ai;         // [n_f2c] fine aux index of the coincident node
    bool f2c_apart = false;     // no fine node is both a gather target and a fine-to-coarse
                                // source: the gather and the copy may share one launch
};

// Direct-forcing immersed boundary of one level (FsiImmersedBoundary::CalculateBodyForce,
// immersed_boundary.cpp:35-384); the marker stencils are fixed, so weights are precomputed.
struct IbTables {
    int64_t n_markers = 0;
    int support = 0;            // (2 * stencil)^dims entries per marker
    int64_t n_reset = 0;
    DevBuf node_ai;             // [n_markers][support] aux index, -1 for unused entries
    DevBuf weight;              // [n_markers][support] delta weight (double)
    DevBuf reset_ai;            // [n_reset] aux indices of every supported node, ascending
    DevBuf spread_begin;        // [n_reset + 1] offsets into the spread arrays per supported node
    DevBuf spread_marker;       // marker of each spread term, ascending per node
    DevBuf spread_weight;       // its delta weight times the marker's area / (dt dx^(dims-1))
    DevBuf node_u;              // [naux][3] velocity recomputed at the reset (double)
    DevBuf marker_vel;          // [n_markers][3]
    DevBuf marker_force;        // [n_markers][3] Lagrangian force
    double stiffness = 2.;
};

struct IbMarkers {
    std::vector<double> xyz, area, vel;
    int stencil_dis = 2;
    double stiffness = 2.;
    int64_t n = 0;
};

// host-side plans (node indices in key order) from which the device tables are laid out
struct TransferPlan {
    std::vector<int64_t> tgt;          // fine node of each interpolated ghost node
    std::vector<int64_t> begin;        // contributor offsets per target
    std::vector<int64_t> contrib;      // >= 0 coarse node index, < 0: -(fine node index) - 1
    std::vector<double> coeff;
    std::vector<int64_t> f2c_coarse, f2c_fine;
    int64_t incomplete = 0;            // stencils with a contributor absent on both levels
};

struct IbPlan {
    std::vector<int64_t> node;         // [n_markers][support] node index or -1
    std::vector<double> weight;
    std::vector<double> scale;
    int support = 0;
    // periodic domains: duplicates of stencil nodes on periodic planes and the table entry
    // (marker x support + entry) whose weight each takes in the spread
    std::vector<int64_t> dup_node, dup_entry;
};

// Roles of a node on the deferred ("aux") path.  Its post-stream values are kept in the aux array
// (kAuxRecorded) for the boundary, transfer and forcing kernels to read or change; its collision
// waits for the next substep's aux collision where they change after the fused kernel: after
// every substep (kAuxDeferAlways: domain-boundary nodes, immersed-boundary supports, fine-to-
// coarse targets and sources, nodes with an absent pull source) or only after the substeps that
// the transfer into the level follows (kAuxDeferOnTransfer: coarse-to-fine ghost targets).  A
// recorded node that is not deferred collides in the fused kernel like a bulk node.
constexpr uint8_t kAuxRecorded = 1, kAuxDeferOnTransfer = 2, kAuxDeferAlways = 4;

struct LevelState {
    // host copies
    std::vector<uint64_t> keys;
    std::vector<int32_t> flags;
    std::map<std::pair<int, int>, std::vector<uint64_t>> layers;
    IbMarkers ib;
    bool have = false;
    // sizes and host maps
    int64_t n = 0, nbrick = 0, naux = 0;
    std::vector<uint64_t> bkeys;
    std::vector<int32_t> node_slot_h, node_aux_h;
    std::vector<uint32_t> missing_h;
    std::vector<uint8_t> aux_mark;   // kAux* roles per node (0: bulk)
    // aux nodes of this level that the transfer from the coarser level writes (ghost targets) and
    // reads (fallback rows, fine-to-coarse sources), and whether the transfer into the next finer
    // level conflicts with them (writes a node the first touches, or reads a node it writes)
    std::vector<int32_t> xfer_writes, xfer_reads;
    bool xfer_overlap = false;
    // device
    DevBuf keys_d, bkeys_d, node_slot_d, slot_node_d, occ_d, nbr_d, auxmask_d, auxbase_d, aux_missing_d,
        aux_slot_d, node_aux_d, f_d[2], aux_f_d, aux_macro_d, aux_force_d,
        aux_macro_flag_d;   // [naux] 1: the aux collision records rho, u (fine-to-coarse sources)
    // boundary tables per direction: aux index of node, partner aux index, slot
    DevBuf bc_ai_d[7], bc_partner_d[7], bc_slot_d[7];
    DevBuf bc_u_d[7];   // [n][3] per-node boundary velocity of a profiled face (empty: uniform)
    int64_t bc_n[7] = {0, 0, 0, 0, 0, 0, 0};
    std::vector<int64_t> bnode[7], bpart[7];
    TransferTables xfer;
    IbTables ibt;
    TransferPlan xplan;
    IbPlan ibplan;
    // deferred nodes per kind of substep (0: no transfer into the level follows, 1: one does): a
    // per-brick lane mask for the fused kernel and the list of their aux indices for the aux
    // collision that follows
    DevBuf defmask_d[2], deflist_d[2];
    int64_t ndef[2] = {0, 0};
    int cur = 0;        // buffer read by the next fused kernel
    bool fresh = true;  // no step since upload
    // nodes the next aux collision takes: kPendingAll (every recorded node: after an upload), or
    // the deferred nodes of the last substep's kind (kPendingPlain, kPendingTransfer)
    int pending = 0;
    int last_tsl = 0;   // substep index of the last step (a transfer into the level needs it even)
    StepParams sp{};
};

constexpr int kPendingAll = 0, kPendingPlain = 1, kPendingTransfer = 2;

}  // namespace zlh
