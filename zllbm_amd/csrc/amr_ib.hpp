// amr_ib.hpp -- coarse/fine transfer and immersed-boundary forcing: host plans, device tables
// and kernels.  Included by zlh_api.hip after its allocation helpers.
#pragma once

namespace zlh {

// ----------------------------------------------------------------------------------------------
// kernels
// ----------------------------------------------------------------------------------------------
template <int D>
__device__ __forceinline__ void macro_feq_double(const double *f, int compressible, double &rho, double *u,
                                                 double *feq) {
    using L = Lattice<D>;
    constexpr int Q = L::Q;
    double r = f[0], ux = 0, uy = 0, uz = 0;
    for (int q = 1; q < Q; ++q) {
        r += f[q];
        ux += L::cx[q] * f[q];
        uy += L::cy[q] * f[q];
        if (D == 3) uz += L::cz[q] * f[q];
    }
    if (compressible) { ux /= r; uy /= r; if (D == 3) uz /= r; }
    rho = r; u[0] = ux; u[1] = uy; u[2] = uz;
    const double usq = ux * ux + uy * uy + (D == 3 ? uz * uz : 0.);
    for (int q = 0; q < Q; ++q) {
        double cu = ux * L::cx[q] + uy * L::cy[q];
        if (D == 3) cu += uz * L::cz[q];
        feq[q] = compressible ? L::w[q] * r * (1. + 3. * cu + 4.5 * cu * cu - 1.5 * usq)
                              : L::w[q] * (r + 3. * cu + 4.5 * cu * cu - 1.5 * usq);
    }
}

// NodeInfoCoarse2fine (lbm_grid_interface.cpp:570-588) + PostStreamCoarse2Fine
// (lbm_collision_models.cpp:244-251) for every coarse node that feeds an interpolation stencil.
// The rescaled populations are stored node-major, [n_src][Q], in the populations' storage form
// (FP64, or FP32 shifted by w), so that each stencil contributor is one contiguous read.
template <int D, typename T, bool SHIFT>
__global__ void k_c2f_sources(int64_t n_src, const int32_t *src_ai, const T *aux_f, int64_t naux, double tau_ratio,
                              int compressible, T *src_f) {
    constexpr int Q = Lattice<D>::Q;
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_src) return;
    const int64_t j = src_ai[r];
    double f[Q], feq[Q], rho, u[3];
    for (int q = 0; q < Q; ++q) f[q] = double(aux_f[q * naux + j]) + (SHIFT ? Lattice<D>::w[q] : 0.);
    macro_feq_double<D>(f, compressible, rho, u, feq);
    for (int q = 0; q < Q; ++q)
        src_f[r * Q + q] = T(feq[q] + tau_ratio * (f[q] - feq[q]) - (SHIFT ? Lattice<D>::w[q] : 0.));
}

// InterpolationLagrangian{2D,3D} (grid/grid_interpolation.cpp:224-304, 433-525): each ghost node
// is the weighted sum of its contributors, accumulated in stencil order.  A workgroup takes a
// group of consecutive (Morton-ordered, so neighbouring) ghost nodes: it first copies the union
// of their coarse source rows into LDS with coalesced loads (neighbours share most of their
// 4^d stencils), then each ghost node is summed from LDS by four threads of 5 (D2Q9: 5 and 4)
// directions.  A contributor is one 8-byte word {LDS row or fine fallback, FP32 weight}: the
// tensor-product Lagrange weights are dyadic rationals, exact in FP32.
constexpr int kGatherDirs = 5;
template <int D, typename T, bool SHIFT>
__global__ __launch_bounds__(256) void k_c2f_gather(int64_t n_tgt, int group, int max_rows,
                                                    const int32_t *__restrict__ tgt_ai,
                                                    const int32_t *__restrict__ begin,
                                                    const int2 *__restrict__ contrib,
                                                    const int32_t *__restrict__ rows_begin,
                                                    const int32_t *__restrict__ rows, const T *__restrict__ src_f,
                                                    T *__restrict__ aux_f, int64_t naux) {
    constexpr int Q = Lattice<D>::Q;
    constexpr int NQ = kGatherDirs;
    constexpr int PARTS = (Q + NQ - 1) / NQ;
    extern __shared__ unsigned char lds_raw[];
    T *rowbuf = reinterpret_cast<T *>(lds_raw);                                  // [max_rows][Q]
    int2 *cbuf = reinterpret_cast<int2 *>(lds_raw + (((size_t)max_rows * Q * sizeof(T) + 15) & ~(size_t)15));
    const int64_t gi = blockIdx.x;
    const int64_t t_first = gi * group;
    const int64_t t_end = t_first + group < n_tgt ? t_first + group : n_tgt;
    const int32_t c0 = begin[t_first], nc = begin[t_end] - c0;
    const int32_t r0 = rows_begin[gi], nr = rows_begin[gi + 1] - r0;
    // staging: every load of a thread is issued before its first LDS write (a load followed by
    // its store, repeated, would wait once per element)
    for (int rb = 0; rb < nr; rb += 256) {
        const int r = rb + (int)threadIdx.x;
        if (r < nr) {
            const T *row = src_f + (int64_t)rows[r0 + r] * Q;
            T v[Q];
#pragma unroll
            for (int q = 0; q < Q; ++q) v[q] = row[q];
#pragma unroll
            for (int q = 0; q < Q; ++q) rowbuf[r * Q + q] = v[q];
        }
    }
    for (int cb = 0; cb < nc; cb += 4 * 256) {
        int2 w[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int e = cb + j * 256 + (int)threadIdx.x;
            w[j] = contrib[c0 + (e < nc ? e : nc - 1)];
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int e = cb + j * 256 + (int)threadIdx.x;
            if (e < nc) cbuf[e] = w[j];
        }
    }
    __syncthreads();
    const int local = threadIdx.x / PARTS;
    const int q0 = (threadIdx.x - local * PARTS) * NQ;
    const int64_t t = t_first + local;
    if (local >= group || t >= n_tgt) return;
    double acc[NQ];
#pragma unroll
    for (int i = 0; i < NQ; ++i) acc[i] = 0.;
    const int32_t k1 = begin[t + 1] - c0;
    for (int32_t k = begin[t] - c0; k < k1; ++k) {
        const int2 cw = cbuf[k];
        const double c = (double)__int_as_float(cw.y);
#pragma unroll
        for (int i = 0; i < NQ; ++i) {
            if (q0 + i >= Q) break;
            const int q = q0 + i;
            const double v = cw.x >= 0 ? double(rowbuf[cw.x * Q + q]) : double(aux_f[q * naux + (-(int64_t)cw.x - 1)]);
            acc[i] = acc[i] + c * (v + (SHIFT ? Lattice<D>::w[q] : 0.));
        }
    }
    const int64_t ai = tgt_ai[t];
#pragma unroll
    for (int i = 0; i < NQ; ++i)
        if (q0 + i < Q) aux_f[(q0 + i) * naux + ai] = T(acc[i] - (SHIFT ? Lattice<D>::w[q0 + i] : 0.));
}

// NodeInfoFine2Coarse (lbm_grid_interface.cpp:594-610) + PostStreamFine2Coarse
// (lbm_collision_models.cpp:259-267): rho, u of the fine node's last collision
template <int D, typename T, bool SHIFT>
__global__ void k_f2c(int64_t n, const int32_t *coarse_ai, const int32_t *fine_ai, const T *fine_f, int64_t naux_f,
                      const double *fine_macro, double tau_ratio, int compressible, T *coarse_f, int64_t naux_c) {
    using L = Lattice<D>;
    constexpr int Q = L::Q;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int64_t i = fine_ai[p], j = coarse_ai[p];
    const double rho = fine_macro[i * 4], ux = fine_macro[i * 4 + 1], uy = fine_macro[i * 4 + 2];
    const double uz = D == 3 ? fine_macro[i * 4 + 3] : 0.;
    const double usq = ux * ux + uy * uy + (D == 3 ? uz * uz : 0.);
    for (int q = 0; q < Q; ++q) {
        double cu = ux * L::cx[q] + uy * L::cy[q];
        if (D == 3) cu += uz * L::cz[q];
        const double feq = compressible ? L::w[q] * rho * (1. + 3. * cu + 4.5 * cu * cu - 1.5 * usq)
                                        : L::w[q] * (rho + 3. * cu + 4.5 * cu * cu - 1.5 * usq);
        const double ff = double(fine_f[q * naux_f + i]) + (SHIFT ? L::w[q] : 0.);
        coarse_f[q * naux_c + j] = T(feq + tau_ratio * (ff - feq) - (SHIFT ? L::w[q] : 0.));
    }
}

// immersed boundary, step 1: zero the force of every supported node and recompute its velocity
// with the remaining (constant) force (the lazy reset, immersed_boundary.cpp:76-85, 284-296)
template <int D, typename T, typename C, bool SHIFT>
__global__ void k_ib_reset(int64_t n, const int32_t *reset_ai, const T *aux_f, int64_t naux, double *aux_force,
                           double *node_u, StepParams p) {
    constexpr int Q = Lattice<D>::Q;
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int64_t ai = reset_ai[k];
    aux_force[ai * 3] = 0.;
    aux_force[ai * 3 + 1] = 0.;
    aux_force[ai * 3 + 2] = 0.;
    C f[Q];
    for (int q = 0; q < Q; ++q) f[q] = C(aux_f[q * naux + ai]);
    C F[3] = {C(p.const_force[0]), C(p.const_force[1]), C(p.const_force[2])};
    C rho, drho, u[3];
    macro<D, SHIFT, C>(f, F, C(p.dt), 1, p.compressible, rho, drho, u);
    node_u[ai * 3] = double(u[0]);
    node_u[ai * 3 + 1] = double(u[1]);
    node_u[ai * 3 + 2] = double(u[2]);
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// immersed boundary, steps 2-3: one wavefront per marker; lanes hold the stencil nodes.
// Velocity interpolation by a wavefront reduction, Lagrangian force F = k (U - u_ref), and the
// FP64 atomic spread F delta area / (dt dx^(d-1)) (DirectForcingScheme{2D,3D},
// immersed_boundary.cpp:108-364)
template <int D>
__global__ __launch_bounds__(256) void k_ib_markers(int64_t n_markers, int support, const int32_t *node_ai,
                                                    const double *weight, const double *node_u,
                                                    const double *marker_vel, const double *marker_scale,
                                                    double stiffness, double *marker_force, double *aux_force) {
    const int64_t m = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (m >= n_markers) return;
    int32_t ai = -1;
    double w = 0.;
    if (lane < support) {
        ai = node_ai[m * support + lane];
        w = weight[m * support + lane];
    }
    double u[3] = {0., 0., 0.};
    if (ai >= 0)
        for (int a = 0; a < D; ++a) u[a] = node_u[(int64_t)ai * 3 + a] * w;
    double fm[3] = {0., 0., 0.};
    for (int a = 0; a < D; ++a) {
        const double ur = wave_sum(u[a]);
        fm[a] = stiffness * (marker_vel[m * 3 + a] - ur);
    }
    if (lane == 0)
        for (int a = 0; a < 3; ++a) marker_force[m * 3 + a] = fm[a];
    if (ai >= 0) {
        const double s = w * marker_scale[m];
        for (int a = 0; a < D; ++a) atomicAdd(&aux_force[(int64_t)ai * 3 + a], fm[a] * s);
    }
}

// ----------------------------------------------------------------------------------------------
// host plans
// ----------------------------------------------------------------------------------------------
// one axis of FindNodesInPeriodicRegionCorner (grid/sfbitset_aux.cpp:380-557, 2493-2878):
// coordinates of the 2L region entries whose local index L-1 is `start`; entries past a
// non-periodic domain edge are -1; at a periodic edge the walk continues from the opposite
// domain plane, which duplicates the edge plane.  Returns the axis' bound on the valid half length.
inline int region_axis(int64_t start, int L, int64_t lo, int64_t hi, bool per_lo, bool per_hi, int64_t *seq) {
    int index_min = L;
    for (int k = 0; k < 2 * L; ++k) seq[k] = -1;
    int64_t v = start;
    for (int ix = 0; ix < L; ++ix) {
        seq[L - ix - 1] = v;
        if (v == lo) {
            if (per_lo) v = hi + 1;
            else { index_min = std::min(index_min, ix + 1); break; }
        }
        v = v > 0 ? v - 1 : v;
    }
    v = start;
    bool go = true;
    if (v == hi) {
        if (per_hi) v = lo > 0 ? lo - 1 : lo;
        else { index_min = 0; go = false; }
    }
    if (go) {
        for (int ix = 0; ix < L; ++ix) {
            v = v + 1;
            seq[L + ix] = v;
            if (v == hi) {
                if (per_hi) v = lo > 0 ? lo - 1 : lo;
                else { index_min = std::min(index_min, ix + 1); break; }
            }
        }
    }
    return index_min;
}

// CalculateLagrangianInterpCoeff (grid/grid_interpolation.cpp:87-112)
inline void lagrange_coeff(int half, double *c0, double *c1) {
    const int num = 2 * half, offset = half - 1;
    for (int iy = 0; iy < num; ++iy) {
        c1[iy] = 1.;
        c0[iy] = 0.;
        for (int ix = 0; ix < num; ++ix)
            if (ix != iy) c1[iy] *= (0.5 + offset - (double)ix) / ((double)iy - (double)ix);
    }
    if (num > 0) c0[offset] = 1.;
}

inline uint64_t encode_xyz(int dims, const int64_t *c) {
    uint64_t k = 0;
    for (int a = 0; a < dims; ++a) k |= spread(dims, (uint32_t)c[a], a);
    return k;
}

// stencils of TransferInfoFromCoarseGrid / TransferInfoToCoarseGrid (lbm_grid_interface.cpp:408-564)
inline zlh_status plan_transfer(zlh_ctx *ctx, int level) {
    LevelState &F = ctx->lv[level], &C = ctx->lv[level - 1];
    TransferPlan &P = F.xplan;
    P = TransferPlan();
    const int dims = ctx->d.dims;
    const int lc = level - 1;
    const int Lmax = 2;   // max_interp_length_ (grid/grid_info_interface.h:317-318)
    bool per_lo[3], per_hi[3];
    int64_t lo[3], hi[3];
    for (int a = 0; a < dims; ++a) {
        per_lo[a] = ctx->d.bc_scheme[1 + 2 * a] == ZLH_BC_PERIODIC;
        per_hi[a] = ctx->d.bc_scheme[2 + 2 * a] == ZLH_BC_PERIODIC;
        lo[a] = (int64_t)ctx->d.dom_min[a] << lc;
        hi[a] = (int64_t)ctx->d.dom_max[a] << lc;
    }
    std::vector<uint8_t> is_target(F.n, 0);
    P.begin.push_back(0);
    for (int layer = 1; layer < 3; ++layer) {
        for (int kind : {ZLH_INNER_F2C, ZLH_OUTER_F2C}) {
            auto it = F.layers.find({kind, layer});
            if (it == F.layers.end()) continue;
            for (uint64_t key : it->second) {
                const int64_t ti = find_key(F.keys, key);
                P.tgt.push_back(ti);
                is_target[ti] = 1;
                const uint64_t parent = key >> dims;
                int64_t seq[3][4];
                int valid = Lmax;
                for (int a = 0; a < dims; ++a)
                    valid = std::min(valid, region_axis(axis_coord(dims, parent, a), Lmax, lo[a], hi[a], per_lo[a],
                                                        per_hi[a], seq[a]));
                double c0[4], c1[4], cf[3][4];
                lagrange_coeff(valid, c0, c1);
                const int num = 2 * valid, off = Lmax - valid;
                for (int a = 0; a < dims; ++a)
                    for (int k = 0; k < num; ++k) cf[a][k] = (axis_coord(dims, key, a) & 1u) ? c1[k] : c0[k];
                const int nz = dims == 3 ? num : 1;
                bool broken = false;
                for (int iz = 0; iz < nz && !broken; ++iz)
                    for (int iy = 0; iy < num && !broken; ++iy)
                        for (int ix = 0; ix < num && !broken; ++ix) {
                            const double coeff = dims == 3 ? cf[0][ix] * cf[1][iy] * cf[2][iz] : cf[0][ix] * cf[1][iy];
                            if (!(std::fabs(coeff) > 1e-6)) continue;
                            int64_t c[3] = {seq[0][ix + off], seq[1][iy + off], dims == 3 ? seq[2][iz + off] : 0};
                            const uint64_t ck = encode_xyz(dims, c);
                            const int64_t j = find_key(C.keys, ck);
                            if (j >= 0 && !(C.flags[j] & (1 << 2))) {   // kNodeStatusCoarse2FineGhost_
                                P.contrib.push_back(j);
                                C.aux_mark[j] = 1;
                            } else {
                                const int64_t k = find_key(F.keys, ck << dims);
                                if (k < 0) { broken = true; ++P.incomplete; break; }
                                P.contrib.push_back(-k - 1);
                                F.aux_mark[k] = 1;
                            }
                            P.coeff.push_back(coeff);
                        }
                P.begin.push_back((int64_t)P.contrib.size());
                F.aux_mark[ti] = 1;
            }
        }
    }
    // contributors are read before any target is written: a fallback fine node that is itself
    // an interpolation target would make the result depend on the processing order
    for (int64_t c : P.contrib)
        if (c < 0 && is_target[-c - 1])
            return ctx->fail(ZLH_EINVAL, "level %d: interpolation fallback node is also a ghost target", level);
    for (int kind : {ZLH_INNER_C2F, ZLH_OUTER_C2F}) {
        auto it = C.layers.find({kind, 1});
        if (it == C.layers.end()) continue;
        for (uint64_t key : it->second) {
            const int64_t j = find_key(C.keys, key);
            const int64_t i = find_key(F.keys, key << dims);
            if (i < 0) continue;
            P.f2c_coarse.push_back(j);
            P.f2c_fine.push_back(i);
            C.aux_mark[j] = 1;
            F.aux_mark[i] = 1;
        }
    }
    return ZLH_OK;
}

inline double stencil_one(double d) { return std::fabs(d) < 1. ? 1. - std::fabs(d) : 0.; }
inline double stencil_two(double d) {
    if (d < 2.) {
        const double a = std::fabs(d), sq = d * d;
        if (a < 1.) return 0.125 * (3. - 2. * a + std::sqrt(1. + 4. * a - 4. * sq));
        return 0.125 * (5. - 2. * a - std::sqrt(-7. + 12. * a - 4. * sq));
    }
    return 0.;
}

// marker stencils of DirectForcingScheme{2D,3D} (immersed_boundary.cpp:108-364)
inline zlh_status plan_ib(zlh_ctx *ctx, int level) {
    LevelState &L = ctx->lv[level];
    IbPlan &P = L.ibplan;
    P = IbPlan();
    const IbMarkers &M = L.ib;
    if (M.n == 0) return ZLH_OK;
    const int dims = ctx->d.dims;
    for (int d = 1; d <= 2 * dims; ++d)
        if (ctx->d.bc_scheme[d] == ZLH_BC_PERIODIC)
            return ctx->fail(ZLH_EINVAL, "immersed boundary in a periodic domain is not supported");
    const int Ls = M.stencil_dis;
    P.support = 1;
    for (int a = 0; a < dims; ++a) P.support *= 2 * Ls;
    P.node.assign(M.n * P.support, -1);
    P.weight.assign(M.n * P.support, 0.);
    P.scale.resize(M.n);
    const double dx = 1. / (double)(1u << level), dt = dx;
    for (int64_t m = 0; m < M.n; ++m) {
        double dis[3];
        int64_t idx[3], seq[3][8];
        int valid = Ls;
        for (int a = 0; a < dims; ++a) {
            const double iref = M.xyz[m * 3 + a] / dx;
            idx[a] = (int64_t)(iref + 1e-6);
            dis[a] = iref - (double)idx[a];
            const int64_t lo = (int64_t)ctx->d.dom_min[a] << level, hi = (int64_t)ctx->d.dom_max[a] << level;
            valid = std::min(valid, region_axis(idx[a], Ls, lo, hi, false, false, seq[a]));
        }
        double (*stencil)(double) = valid == 2 ? stencil_two : stencil_one;
        const int n1 = 2 * valid, nz = dims == 3 ? n1 : 1;
        int e = 0;
        for (int iz = 0; iz < nz; ++iz) {
            const int oz = iz - valid;
            const double wz = dims == 3 ? stencil(std::fabs(dis[2] - oz - 1)) : 1.;
            for (int iy = 0; iy < n1; ++iy) {
                const int oy = iy - valid;
                const double wy = stencil(std::fabs(dis[1] - oy - 1));
                for (int ix = 0; ix < n1; ++ix, ++e) {
                    const int ox = ix - valid;
                    const double wx = stencil(std::fabs(dis[0] - ox - 1));
                    int64_t c[3] = {seq[0][Ls + ox], seq[1][Ls + oy], dims == 3 ? seq[2][Ls + oz] : 0};
                    if (c[0] < 0 || c[1] < 0 || c[2] < 0)
                        return ctx->fail(ZLH_EINVAL, "marker %lld stencil leaves the domain", (long long)m);
                    const int64_t j = find_key(L.keys, encode_xyz(dims, c));
                    if (j < 0) return ctx->fail(ZLH_EINVAL, "marker %lld stencil node absent at level %d", (long long)m, level);
                    P.node[m * P.support + e] = j;
                    P.weight[m * P.support + e] = dims == 3 ? wx * wy * wz : wx * wy;
                    L.aux_mark[j] = 1;
                }
            }
        }
        P.scale[m] = dims == 3 ? M.area[m] / dt / dx / dx : M.area[m] / dt / dx;
    }
    return ZLH_OK;
}

// ----------------------------------------------------------------------------------------------
// device tables
// ----------------------------------------------------------------------------------------------
inline zlh_status upload_transfer(zlh_ctx *ctx, int level) {
    LevelState &F = ctx->lv[level], &C = ctx->lv[level - 1];
    TransferPlan &P = F.xplan;
    TransferTables &X = F.xfer;
    zlh_status s;
    std::vector<int32_t> tgt, begin, src, src_ai, f2c_c, f2c_f;
    std::map<int64_t, int32_t> rows;
    for (int64_t t : P.tgt) tgt.push_back(F.node_aux_h[t]);
    for (int64_t b : P.begin) begin.push_back((int32_t)b);
    for (int64_t c : P.contrib) {
        if (c >= 0) {
            auto it = rows.find(c);
            if (it == rows.end()) {
                it = rows.emplace(c, (int32_t)src_ai.size()).first;
                src_ai.push_back(C.node_aux_h[c]);
            }
            src.push_back(it->second);
        } else {
            src.push_back(-F.node_aux_h[-c - 1] - 1);
        }
    }
    for (size_t k = 0; k < P.f2c_coarse.size(); ++k) {
        f2c_c.push_back(C.node_aux_h[P.f2c_coarse[k]]);
        f2c_f.push_back(F.node_aux_h[P.f2c_fine[k]]);
    }
    for (int32_t v : tgt) if (v < 0) return ctx->fail(ZLH_EDEVICE, "transfer target not deferred");
    for (int32_t v : src_ai) if (v < 0) return ctx->fail(ZLH_EDEVICE, "transfer source not deferred");
    X.n_c2f = (int64_t)tgt.size();
    X.n_src = (int64_t)src_ai.size();
    X.n_f2c = (int64_t)f2c_c.size();
    if ((s = upload(ctx, X.c2f_target, tgt))) return s;
    if ((s = upload(ctx, X.c2f_begin, begin))) return s;
    // gather groups: consecutive ghost nodes, their source rows staged in LDS; the group is
    // halved until the largest row set fits the LDS budget
    const int Qn = ctx->q;
    const size_t lds_budget = 48 * 1024;
    std::vector<int2> packed(src.size());
    std::vector<int32_t> group_begin, group_rows;
    int group = 256 / ((Qn + kGatherDirs - 1) / kGatherDirs);
    for (;; group /= 2) {
        group_begin.assign(1, 0);
        group_rows.clear();
        int max_rows = 0;
        int64_t max_contrib = 0;
        const int64_t nt = (int64_t)tgt.size();
        for (int64_t g0 = 0; g0 < nt; g0 += group) {
            std::map<int32_t, int32_t> local;
            for (int64_t t = g0; t < std::min<int64_t>(nt, g0 + group); ++t)
                for (int32_t k = begin[t]; k < begin[t + 1]; ++k) {
                    const float cf = (float)P.coeff[k];
                    if ((double)cf != P.coeff[k]) return ctx->fail(ZLH_EDEVICE, "interpolation weight not exact in FP32");
                    int32_t bits;
                    std::memcpy(&bits, &cf, 4);
                    int32_t sx = src[k];
                    if (sx >= 0) {
                        auto it = local.find(sx);
                        if (it == local.end()) {
                            it = local.emplace(sx, (int32_t)local.size()).first;
                            group_rows.push_back(sx);
                        }
                        sx = it->second;
                    }
                    packed[k] = make_int2(sx, bits);
                }
            max_rows = std::max<int>(max_rows, (int)local.size());
            max_contrib = std::max<int64_t>(max_contrib, begin[std::min<int64_t>(nt, g0 + group)] - begin[g0]);
            group_begin.push_back((int32_t)group_rows.size());
        }
        X.max_rows = max_rows;
        X.lds_bytes = (((size_t)max_rows * Qn * ctx->tsize() + 15) & ~(size_t)15) + (size_t)max_contrib * 8;
        if (X.lds_bytes <= lds_budget) break;
        if (group == 1) return ctx->fail(ZLH_EDEVICE, "interpolation stencil too large for LDS");
    }
    X.group = group;
    X.n_groups = (int64_t)group_begin.size() - 1;
    if ((s = upload(ctx, X.group_rows_begin, group_begin))) return s;
    if ((s = upload(ctx, X.group_rows, group_rows))) return s;
    if ((s = upload(ctx, X.c2f_contrib, packed))) return s;
    if ((s = upload(ctx, X.src_coarse_ai, src_ai))) return s;
    if ((s = dalloc(ctx, X.src_f, ctx->tsize() * ctx->q * std::max<int64_t>(X.n_src, 1)))) return s;
    if ((s = upload(ctx, X.f2c_coarse_ai, f2c_c))) return s;
    if ((s = upload(ctx, X.f2c_fine_ai, f2c_f))) return s;
    return ZLH_OK;
}

inline zlh_status upload_ib(zlh_ctx *ctx, int level) {
    LevelState &L = ctx->lv[level];
    IbTables &T = L.ibt;
    IbPlan &P = L.ibplan;
    T.n_markers = L.ib.n;
    if (T.n_markers == 0) return ZLH_OK;
    zlh_status s;
    T.support = P.support;
    T.stiffness = L.ib.stiffness;
    std::vector<int32_t> ai(P.node.size(), -1), reset;
    std::vector<uint8_t> seen(L.naux, 0);
    for (size_t k = 0; k < P.node.size(); ++k) {
        if (P.node[k] < 0) continue;
        ai[k] = L.node_aux_h[P.node[k]];
        if (ai[k] < 0) return ctx->fail(ZLH_EDEVICE, "immersed boundary node not deferred");
        if (!seen[ai[k]]) { seen[ai[k]] = 1; reset.push_back(ai[k]); }
    }
    T.n_reset = (int64_t)reset.size();
    if ((s = upload(ctx, T.node_ai, ai))) return s;
    if ((s = upload(ctx, T.weight, P.weight))) return s;
    if ((s = upload(ctx, T.reset_ai, reset))) return s;
    if ((s = upload(ctx, T.marker_vel, L.ib.vel))) return s;
    if ((s = upload(ctx, T.marker_scale, P.scale))) return s;
    if ((s = dalloc(ctx, T.marker_force, sizeof(double) * 3 * T.n_markers))) return s;
    if ((s = dalloc(ctx, T.node_u, sizeof(double) * 3 * std::max<int64_t>(L.naux, 1)))) return s;
    return ZLH_OK;
}

// ----------------------------------------------------------------------------------------------
// launches
// ----------------------------------------------------------------------------------------------
template <int D, typename T, bool SHIFT>
zlh_status launch_transfers_t(zlh_ctx *ctx, int level) {
    LevelState &F = ctx->lv[level], &C = ctx->lv[level - 1];
    TransferTables &X = F.xfer;
    const int comp = ctx->d.compressible;
    // tau ratios of PostStreamCoarse2Fine (coarse level) and PostStreamFine2Coarse (fine level),
    // CalRelaxationTimeRatio (lbm_collision_models.cpp:142-151)
    const double nu = ctx->d.nu0;
    const double dtc = 1. / (double)(1u << (level - 1)), dtf = 1. / (double)(1u << level);
    const double c2f = 0.5 * (nu * 3. / dtc * 2. + 0.5) / (nu * 3. / dtc + 0.5);
    const double f2c = 2. * (nu * 3. / dtf / 2 + 0.5) / (nu * 3. / dtf + 0.5);
    if (X.n_c2f > 0) {
        if (X.n_src > 0)
            hipLaunchKernelGGL((k_c2f_sources<D, T, SHIFT>), dim3(blocks(X.n_src)), dim3(256), 0, ctx->stream, X.n_src,
                               X.src_coarse_ai.as<int32_t>(), C.aux_f_d.as<T>(), C.naux, c2f, comp, X.src_f.as<T>());
        hipLaunchKernelGGL((k_c2f_gather<D, T, SHIFT>), dim3((unsigned)X.n_groups), dim3(256), X.lds_bytes,
                           ctx->stream, X.n_c2f, X.group, X.max_rows,
                           X.c2f_target.as<int32_t>(), X.c2f_begin.as<int32_t>(), X.c2f_contrib.as<int2>(),
                           X.group_rows_begin.as<int32_t>(), X.group_rows.as<int32_t>(), X.src_f.as<T>(),
                           F.aux_f_d.as<T>(), F.naux);
    }
    if (X.n_f2c > 0)
        hipLaunchKernelGGL((k_f2c<D, T, SHIFT>), dim3(blocks(X.n_f2c)), dim3(256), 0, ctx->stream, X.n_f2c,
                           X.f2c_coarse_ai.as<int32_t>(), X.f2c_fine_ai.as<int32_t>(), F.aux_f_d.as<T>(), F.naux,
                           F.aux_macro_d.as<double>(), f2c, comp, C.aux_f_d.as<T>(), C.naux);
    ZLH_TRY(hipGetLastError());
    return ZLH_OK;
}

inline zlh_status launch_transfers(zlh_ctx *ctx, int level) {
    if (ctx->d.dims == 2)
        return ctx->fp64() ? launch_transfers_t<2, double, false>(ctx, level) : launch_transfers_t<2, float, true>(ctx, level);
    return ctx->fp64() ? launch_transfers_t<3, double, false>(ctx, level) : launch_transfers_t<3, float, true>(ctx, level);
}

template <int D, typename T, typename Cc, bool SHIFT>
zlh_status launch_ib_t(zlh_ctx *ctx, int level) {
    LevelState &L = ctx->lv[level];
    IbTables &B = L.ibt;
    hipLaunchKernelGGL((k_ib_reset<D, T, Cc, SHIFT>), dim3(blocks(B.n_reset)), dim3(256), 0, ctx->stream, B.n_reset,
                       B.reset_ai.as<int32_t>(), L.aux_f_d.as<T>(), L.naux, L.aux_force_d.as<double>(),
                       B.node_u.as<double>(), L.sp);
    hipLaunchKernelGGL((k_ib_markers<D>), dim3(blocks(B.n_markers * 64)), dim3(256), 0, ctx->stream, B.n_markers,
                       B.support, B.node_ai.as<int32_t>(), B.weight.as<double>(), B.node_u.as<double>(),
                       B.marker_vel.as<double>(), B.marker_scale.as<double>(), B.stiffness,
                       B.marker_force.as<double>(), L.aux_force_d.as<double>());
    ZLH_TRY(hipGetLastError());
    return ZLH_OK;
}

inline zlh_status launch_ib(zlh_ctx *ctx, int level) {
    if (ctx->lv[level].ibt.n_markers == 0) return ZLH_OK;
    if (ctx->d.dims == 2)
        return ctx->fp64() ? launch_ib_t<2, double, double, false>(ctx, level) : launch_ib_t<2, float, float, true>(ctx, level);
    return ctx->fp64() ? launch_ib_t<3, double, double, false>(ctx, level) : launch_ib_t<3, float, float, true>(ctx, level);
}

}  // namespace zlh
