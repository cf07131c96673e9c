// amr_ib_types.hpp -- device tables of the coarse/fine transfer and the immersed boundary.
#pragma once
#include <cstdint>

namespace zlh {

// Coarse-to-fine interpolation and fine-to-coarse copy of one refined level (tables live on
// the fine level's state).
struct TransferTables {
    int64_t n_c2f = 0;             // fine ghost nodes interpolated from the coarse level
    int64_t n_src = 0;             // coarse nodes feeding the interpolation
    int64_t n_f2c = 0;             // coarse nodes copied from the fine level
    void *c2f_target = nullptr;    // [n_c2f] fine aux index
    void *c2f_begin = nullptr;     // [n_c2f + 1] CSR offsets into the contributor list
    void *c2f_src = nullptr;       // contributors: >= 0 coarse source row, < 0 -(fine aux index) - 1
    void *c2f_coeff = nullptr;     // contributor weights (double)
    void *src_coarse_ai = nullptr; // [n_src] coarse aux index of each source row
    void *src_f = nullptr;         // [Q][n_src] rescaled coarse populations (double)
    void *src_macro = nullptr;     // [n_src][4] rho, u of the coarse sources
    void *f2c_coarse_ai = nullptr; // [n_f2c] coarse aux index
    void *f2c_fine_ai = nullptr;   // [n_f2c] fine aux index of the coincident node
    int64_t n_contrib = 0;
};

struct IbTables {
    int64_t n_markers = 0;
    int support = 0;               // stencil nodes per marker ((2 * stencil)^dims)
    void *node_ai = nullptr;       // [n_markers][support] aux index, -1 absent
    void *weight = nullptr;        // [n_markers][support] delta weight (double)
    void *reset_ai = nullptr;      // aux indices of every supported node
    int64_t n_reset = 0;
    void *marker_vel = nullptr;    // [n_markers][3]
    void *marker_scale = nullptr;  // [n_markers] area / (dt dx^{d-1})
    double *marker_force_d = nullptr;  // [n_markers][3] Lagrangian force
    void *node_u = nullptr;        // [n_aux][3] velocity of supported nodes (double)
    double stiffness = 2.;
};

}  // namespace zlh
