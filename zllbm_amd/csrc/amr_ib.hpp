// amr_ib.hpp -- coarse/fine transfer and immersed-boundary tables and launches.
// Included by zlh_api.hip after the level state is defined.
#pragma once

namespace zlh {

static thread_local std::string g_xfer_err;
inline const char *transfer_error() { return g_xfer_err.c_str(); }

// marks the nodes of `level` that the transfers read or write (filled in below)
inline zlh_status mark_transfer_nodes(const zlh_solver_desc &d, std::vector<LevelState> &lv, int level,
                                      std::vector<uint8_t> &aux) {
    LevelState &L = lv[level];
    for (auto &kv : L.layers)
        if (!kv.second.empty()) { g_xfer_err = "refinement interfaces not supported yet"; return ZLH_EINVAL; }
    return ZLH_OK;
}

inline zlh_status mark_ib_nodes(const zlh_solver_desc &d, const std::vector<uint64_t> &keys,
                                const std::vector<double> &xyz, int64_t n, int stencil, int level,
                                std::vector<uint8_t> &aux) {
    if (n) { g_xfer_err = "immersed boundary not supported yet"; return ZLH_EINVAL; }
    return ZLH_OK;
}

inline zlh_status build_transfer_tables(const zlh_solver_desc &d, std::vector<LevelState> &lv, int level,
                                        hipStream_t stream) {
    return ZLH_OK;
}

inline zlh_status build_ib_tables(const zlh_solver_desc &d, LevelState &L, int level, hipStream_t stream) {
    return ZLH_OK;
}

inline zlh_status launch_ib(const zlh_solver_desc &d, LevelState &L, int level, hipStream_t stream) {
    return ZLH_OK;
}

inline zlh_status launch_transfers(const zlh_solver_desc &d, std::vector<LevelState> &lv, int level,
                                   hipStream_t stream) {
    return ZLH_OK;
}

}  // namespace zlh
