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
// (lbm_collision_models.cpp:244-251) for every coarse node that feeds an interpolation stencil,
// and a plain copy of every fine node that stands in for a missing coarse one.  The rows are
// stored node-major, [n_src + n_fb][Q], in the populations' storage form (FP64, or FP32 shifted by
// w), so that each stencil contributor is one contiguous row; both row sets are sorted by aux
// index, so consecutive threads read neighbouring entries of each direction's plane.  A
// workgroup's 256 rows go out through LDS as one contiguous, coalesced run.
// MID (time interpolation after an odd fine substep): each coarse row is the mean of the row
// computed now and the row `prev` of the last synchronised transfer, i.e. of the coarse state
// before and after the coarse level's last step; fallback rows stay current copies.
template <int D, typename T, bool SHIFT, bool MID>
__global__ __launch_bounds__(256) void k_c2f_sources(int64_t r_begin, int64_t n_src, int64_t n_fb, const int32_t *src_ai,
                                                     const T *aux_c, int64_t naux_c, const int32_t *fb_ai,
                                                     const T *aux_f, int64_t naux_f, double tau_ratio,
                                                     int compressible, const T *__restrict__ prev, T *src_f) {
    constexpr int Q = Lattice<D>::Q;
    __shared__ T tile[256 * Q];   // row-major; the odd stride Q keeps the row writes conflict-free
    const int64_t r0 = r_begin + (int64_t)blockIdx.x * 256;
    const int64_t r = r0 + threadIdx.x;
    const int64_t n = n_src + n_fb;
    if (r < n_src) {
        const int64_t j = src_ai[r];
        if constexpr (SHIFT) {
            // FP32 / BF16 storage: in FP32 on the stored deviations f - w, whose equilibrium
            // feq - w needs no cancellation
            float f[Q], rho, drho, u[3];
#pragma unroll
            for (int q = 0; q < Q; ++q) f[q] = float(aux_c[q * naux_c + j]);
            macro<D, true, float>(f, nullptr, 0.f, 0, compressible, rho, drho, u);
            const float usq = u[0] * u[0] + u[1] * u[1] + (D == 3 ? u[2] * u[2] : 0.f);
            const float tr = (float)tau_ratio;
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const float feq = feq_stored<D, true, float>(q, rho, drho, u[0], u[1], u[2], usq, compressible);
                tile[threadIdx.x * Q + q] = T(__builtin_fmaf(tr, f[q] - feq, feq));
            }
        } else {
            double f[Q], feq[Q], rho, u[3];
#pragma unroll
            for (int q = 0; q < Q; ++q) f[q] = double(aux_c[q * naux_c + j]);
            macro_feq_double<D>(f, compressible, rho, u, feq);
#pragma unroll
            for (int q = 0; q < Q; ++q) tile[threadIdx.x * Q + q] = T(feq[q] + tau_ratio * (f[q] - feq[q]));
        }
    } else if (r < n) {
        const int64_t fi = fb_ai[r - n_src];
        T v[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) v[q] = aux_f[q * naux_f + fi];
#pragma unroll
        for (int q = 0; q < Q; ++q) tile[threadIdx.x * Q + q] = v[q];
    }
    __syncthreads();
    const int rows = (int)(n - r0 < 256 ? n - r0 : 256);
    if constexpr (MID) {
        // coarse rows [0, n_src) of this workgroup: elements below (n_src - r0) Q
        const int64_t nc = (n_src - r0) * Q;
        for (int e = threadIdx.x; e < rows * Q; e += 256)
            src_f[r0 * Q + e] = e < nc ? T(0.5 * (double(prev[r0 * Q + e]) + double(tile[e]))) : tile[e];
    } else {
        for (int e = threadIdx.x; e < rows * Q; e += 256) src_f[r0 * Q + e] = tile[e];
    }
}

// InterpolationLagrangian{2D,3D} (grid/grid_interpolation.cpp:224-304, 433-525): each ghost node
// is the weighted sum of its contributors, accumulated in stencil order.  Ghost nodes are listed
// by contributor count (the parity of the node's coordinates fixes it: 4 Lagrange weights per odd
// coordinate, 1 per even one), stably, so the threads of a wavefront loop equally often; inside
// one count they keep Morton order, so the nodes of a workgroup are neighbours.  A
// workgroup takes a group of at most 256 / PARTS ghost nodes, as many as keep its staged data
// within the LDS budget: it first copies the union of their rows of src_f and their contributor
// words into LDS, then each ghost node is summed from LDS by PARTS threads of 5 directions (D2Q9:
// 5 and 4).  A contributor is one 32-bit word {LDS row offset, FP16 weight}: the tensor-product
// Lagrange weights are products of -1/16, 9/16, 1/2 and 1, dyadic rationals exact in FP16 (the
// host checks every one).
constexpr int kGatherDirs = 5;
constexpr size_t kGatherLdsBudget = 20 * 1024;
constexpr int kGatherMaxRows = 768;   // staged rows of one group (row offsets fit the word's 16 bits)
constexpr int64_t kGatherChunk = 2048;   // ghost nodes sorted by contributor count together
template <int D, typename T, bool SHIFT>
__device__ __forceinline__ void gather_group(int64_t gi, unsigned char *lds_raw, const GatherGroup *__restrict__ groups,
                                             const int32_t *__restrict__ tgt_ai, const int32_t *__restrict__ begin,
                                             const uint32_t *__restrict__ contrib, const int32_t *__restrict__ rows,
                                             const T *__restrict__ src_f, T *__restrict__ aux_f, int64_t naux) {
    constexpr int Q = Lattice<D>::Q;
    constexpr int NQ = kGatherDirs;
    constexpr int PARTS = (Q + NQ - 1) / NQ;
    constexpr int QP = PARTS * NQ;   // LDS row stride: every thread reads NQ entries, pads included
    const GatherGroup G = groups[gi];
    const int32_t nr = G.r1 - G.r0, nc = G.c1 - G.c0;
    T *rowbuf = reinterpret_cast<T *>(lds_raw);                                  // [nr][QP]
    uint32_t *cbuf = reinterpret_cast<uint32_t *>(lds_raw + (((size_t)nr * QP * sizeof(T) + 15) & ~(size_t)15));
    const int local = threadIdx.x / PARTS;
    const int q0 = (threadIdx.x - local * PARTS) * NQ;
    const int32_t t = G.t0 + local;
    const bool active = t < G.t1;
    // every load below is issued before the first LDS write that waits on it; none depends on
    // an earlier one except a row's data on its entry
    int32_t kb = 0, ke = 0, ai = 0;
    if (active) {
        kb = begin[t] - G.c0;
        ke = begin[t + 1] - G.c0;
        ai = tgt_ai[t];
    }
    for (int cb = 0; cb < nc; cb += 4 * 256) {
        uint32_t w[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int k = cb + j * 256 + (int)threadIdx.x;
            w[j] = k < nc ? contrib[G.c0 + k] : 0u;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int k = cb + j * 256 + (int)threadIdx.x;
            if (k < nc) cbuf[k] = w[j];
        }
    }
    // the rows as one flat run of nr * Q elements, consecutive threads on consecutive elements,
    // so a wavefront's load covers a few cache lines of src_f instead of 64
    constexpr int kBatch = 8;
    const int ne = nr * Q;
    for (int eb = 0; eb < ne; eb += kBatch * 256) {
        int32_t g[kBatch];
        T v[kBatch];
#pragma unroll
        for (int j = 0; j < kBatch; ++j) {
            const int e = eb + j * 256 + (int)threadIdx.x;
            g[j] = rows[G.r0 + (e < ne ? e : ne - 1) / Q];
        }
#pragma unroll
        for (int j = 0; j < kBatch; ++j) {
            const int e = eb + j * 256 + (int)threadIdx.x;
            const int ec = e < ne ? e : ne - 1;
            v[j] = src_f[(int64_t)g[j] * Q + (ec - (ec / Q) * Q)];
        }
#pragma unroll
        for (int j = 0; j < kBatch; ++j) {
            const int e = eb + j * 256 + (int)threadIdx.x;
            if (e < ne) {
                const int r = e / Q;
                rowbuf[r * QP + (e - r * Q)] = v[j];
            }
        }
    }
    __syncthreads();
    if (!active) return;
    // a contributor word is {row * QP (low 16 bits), FP16 weight (high 16 bits)}; the pad entry
    // of the last part is summed like the others and never stored
    auto weight = [](uint32_t cw) { return (float)__builtin_bit_cast(_Float16, (uint16_t)(cw >> 16)); };
    double acc[NQ];
#pragma unroll
    for (int i = 0; i < NQ; ++i) acc[i] = 0.;
    if constexpr (SHIFT) {
        // FP32 / BF16 storage holds f - w.  sum c (f~ + w) - w = sum c f~ + w (S - 1) with S =
        // sum c: the weights are multiples of 2^-12 of magnitude <= 1, so S is exact in FP32.
        // Runs of four products are summed in FP32 (rounding error a few 2^-24 of the run), the
        // runs in FP64.
        float S = 0.f;
        int32_t k = kb;
        for (; k + 4 <= ke; k += 4) {
            float c[4];
            const T *row[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint32_t cw = cbuf[k + j];
                c[j] = weight(cw);
                row[j] = rowbuf + (cw & 0xffffu) + q0;
            }
            S += (c[0] + c[1]) + (c[2] + c[3]);
#pragma unroll
            for (int i = 0; i < NQ; ++i) {
                float run = c[0] * float(row[0][i]);
                run = __builtin_fmaf(c[1], float(row[1][i]), run);
                run = __builtin_fmaf(c[2], float(row[2][i]), run);
                run = __builtin_fmaf(c[3], float(row[3][i]), run);
                acc[i] += (double)run;
            }
        }
        for (; k < ke; ++k) {
            const uint32_t cw = cbuf[k];
            const float c = weight(cw);
            const T *row = rowbuf + (cw & 0xffffu) + q0;
            S += c;
#pragma unroll
            for (int i = 0; i < NQ; ++i) acc[i] += (double)(c * float(row[i]));
        }
#pragma unroll
        for (int i = 0; i < NQ; ++i)
            if (q0 + i < Q) aux_f[(q0 + i) * naux + ai] = T(acc[i] + Lattice<D>::w[q0 + i] * ((double)S - 1.));
    } else {
        // FP64 storage: every product and sum in FP64, in stencil order
#pragma unroll 4
        for (int32_t k = kb; k < ke; ++k) {
            const uint32_t cw = cbuf[k];
            const double c = (double)weight(cw);
            const T *row = rowbuf + (cw & 0xffffu) + q0;
#pragma unroll
            for (int i = 0; i < NQ; ++i) acc[i] = acc[i] + c * double(row[i]);
        }
#pragma unroll
        for (int i = 0; i < NQ; ++i)
            if (q0 + i < Q) aux_f[(q0 + i) * naux + ai] = T(acc[i]);
    }
}

template <int D, typename T, bool SHIFT>
__global__ __launch_bounds__(256) void k_c2f_gather(int64_t n_groups, const GatherGroup *__restrict__ groups,
                                                    const int32_t *__restrict__ tgt_ai,
                                                    const int32_t *__restrict__ begin,
                                                    const uint32_t *__restrict__ contrib,
                                                    const int32_t *__restrict__ rows, const T *__restrict__ src_f,
                                                    T *__restrict__ aux_f, int64_t naux) {
    extern __shared__ unsigned char lds_raw[];
    const int64_t gi = xcd_group(n_groups);   // groups are in spatial order: an XCD takes a range
    if (gi < 0) return;
    gather_group<D, T, SHIFT>(gi, lds_raw, groups, tgt_ai, begin, contrib, rows, src_f, aux_f, naux);
}

// NodeInfoFine2Coarse (lbm_grid_interface.cpp:594-610) + PostStreamFine2Coarse
// (lbm_collision_models.cpp:259-267): rho, u of the fine node's last collision
template <int D, typename T, bool SHIFT>
__device__ __forceinline__ void f2c_node(int64_t p, const int32_t *coarse_ai, const int32_t *fine_ai, const T *fine_f,
                                         int64_t naux_f, const double *fine_macro, double tau_ratio, int compressible,
                                         T *coarse_f, int64_t naux_c) {
    using L = Lattice<D>;
    constexpr int Q = L::Q;
    const int64_t i = fine_ai[p], j = coarse_ai[p];
    if constexpr (SHIFT) {
        // FP32 / BF16 storage: in FP32 on the stored deviations f - w
        const float rho = (float)fine_macro[i * 4], ux = (float)fine_macro[i * 4 + 1];
        const float uy = (float)fine_macro[i * 4 + 2], uz = D == 3 ? (float)fine_macro[i * 4 + 3] : 0.f;
        const float drho = (float)(fine_macro[i * 4] - 1.);
        const float usq = ux * ux + uy * uy + (D == 3 ? uz * uz : 0.f);
        const float tr = (float)tau_ratio;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const float feq = feq_stored<D, true, float>(q, rho, drho, ux, uy, uz, usq, compressible);
            coarse_f[q * naux_c + j] = T(__builtin_fmaf(tr, float(fine_f[q * naux_f + i]) - feq, feq));
        }
        return;
    }
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

template <int D, typename T, bool SHIFT>
__global__ void k_f2c(int64_t n, const int32_t *coarse_ai, const int32_t *fine_ai, const T *fine_f, int64_t naux_f,
                      const double *fine_macro, double tau_ratio, int compressible, T *coarse_f, int64_t naux_c) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    f2c_node<D, T, SHIFT>(p, coarse_ai, fine_ai, fine_f, naux_f, fine_macro, tau_ratio, compressible, coarse_f, naux_c);
}

// The gather and the fine-to-coarse copy of one exchange in one launch, where the host found
// them independent (no fine node is both a gather target and a fine-to-coarse source; the
// coarse-to-fine source rows are taken before either): blocks below nb_gather (a multiple of 8)
// are the gather's, in its XCD-aware order; the blocks after them take 256 fine-to-coarse nodes
// each.
template <int D, typename T, bool SHIFT>
__global__ __launch_bounds__(256) void k_c2f_gather_f2c(int64_t nb_gather, int64_t n_groups,
                                                        const GatherGroup *__restrict__ groups,
                                                        const int32_t *__restrict__ tgt_ai,
                                                        const int32_t *__restrict__ begin,
                                                        const uint32_t *__restrict__ contrib,
                                                        const int32_t *__restrict__ rows, const T *__restrict__ src_f,
                                                        T *__restrict__ aux_f, int64_t naux, int64_t n_f2c,
                                                        const int32_t *coarse_ai, const int32_t *fine_ai,
                                                        const double *fine_macro, double tau_ratio, int compressible,
                                                        T *coarse_f, int64_t naux_c) {
    extern __shared__ unsigned char lds_raw[];
    const int64_t bid = blockIdx.x;
    if (bid < nb_gather) {
        const int64_t per = nb_gather >> 3;
        const int64_t gi = (bid & 7) * per + (bid >> 3);
        if (gi >= n_groups) return;
        gather_group<D, T, SHIFT>(gi, lds_raw, groups, tgt_ai, begin, contrib, rows, src_f, aux_f, naux);
    } else {
        const int64_t p = (bid - nb_gather) * 256 + threadIdx.x;
        if (p < n_f2c)
            f2c_node<D, T, SHIFT>(p, coarse_ai, fine_ai, aux_f, naux, fine_macro, tau_ratio, compressible, coarse_f,
                                  naux_c);
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

// immersed boundary, step 2: one wavefront per marker; lanes hold the stencil nodes.  Velocity
// interpolation by a wavefront reduction and the Lagrangian force F = k (U - u_ref)
// (DirectForcingScheme{2D,3D}, immersed_boundary.cpp:108-364)
template <int D>
__global__ __launch_bounds__(256) void k_ib_markers(int64_t n_markers, int support, const int32_t *node_ai,
                                                    const double *weight, const double *node_u,
                                                    const double *marker_vel, double stiffness,
                                                    double *marker_force) {
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
}

// immersed boundary, step 3: the spread F delta area / (dt dx^(d-1)) of every marker onto its
// stencil nodes (immersed_boundary.cpp:108-364), as a gather: one thread per supported node sums
// its markers' terms in FP64, in ascending marker order, and stores the node force.  Neighbouring
// markers share most of their stencil nodes, so a per-marker scatter would serialise on atomics.
template <int D>
__global__ __launch_bounds__(256) void k_ib_spread(int64_t n_reset, const int32_t *__restrict__ reset_ai,
                                                   const int32_t *__restrict__ begin,
                                                   const int32_t *__restrict__ marker,
                                                   const double *__restrict__ coef,
                                                   const double *__restrict__ marker_force,
                                                   double *__restrict__ aux_force) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_reset) return;
    double acc[3] = {0., 0., 0.};
#pragma unroll 4
    for (int32_t e = begin[r]; e < begin[r + 1]; ++e) {
        const int64_t m = marker[e];
        const double s = coef[e];
        for (int a = 0; a < D; ++a) acc[a] += marker_force[m * 3 + a] * s;
    }
    const int64_t ai = reset_ai[r];
    for (int a = 0; a < D; ++a) aux_force[ai * 3 + a] = acc[a];
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

// one axis of FindNodesInPeriodicRegionCornerOverlap (grid/sfbitset_aux.cpp:581-850,
// 2893-3825), the immersed boundary's region walk in a periodic domain
// (immersed_boundary.cpp:250-256): a walk reaching a periodic domain plane continues past its
// duplicate on the opposite side (lo -> hi - 1, hi -> lo + 1), and the duplicate is returned in
// twin[] at the plane node's entry (-1 elsewhere); it receives that entry's spread force too
// (immersed_boundary.cpp:341-363).  Every duplicate takes its own entry's weight; the reference's
// other weights at a start node on a maximum plane and at the region's first and last rows are
// not reproduced.
inline int region_axis_twin(int64_t start, int L, int64_t lo, int64_t hi, bool per_lo, bool per_hi, int64_t *seq,
                            int64_t *twin) {
    int index_min = L;
    for (int k = 0; k < 2 * L; ++k) seq[k] = twin[k] = -1;
    int64_t v = start;
    for (int ix = 0; ix < L; ++ix) {
        seq[L - ix - 1] = v;
        if (v == lo) {
            if (per_lo) { twin[L - ix - 1] = hi; v = hi; }
            else { index_min = std::min(index_min, ix + 1); break; }
        }
        v = v - 1;
    }
    v = start;
    if (v == hi) {
        if (per_hi) { twin[L - 1] = lo; v = lo; }
        else return 0;
    }
    for (int ix = 0; ix < L; ++ix) {
        v = v + 1;
        seq[L + ix] = v;
        if (v == hi) {
            if (per_hi) { twin[L + ix] = lo; v = lo; }
            else { index_min = std::min(index_min, ix + 1); break; }
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
    // InterpolationLagrangian{2D,3D} (grid/grid_interpolation.cpp:224-304, 433-525), or
    // InterpolationLinear{2D,3D} (:126-208, 319-417): the Lagrangian form with max_interp_length_ 1
    // (:23) and half-length-1 weights whatever the region's valid length
    const bool linear = ctx->interp_method[level] == ZLH_INTERP_LINEAR;
    const int Lmax = linear ? 1 : 2;   // max_interp_length_ (grid/grid_info_interface.h:317-318)
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
                const int half = linear ? 1 : valid;
                lagrange_coeff(half, c0, c1);
                const int num = 2 * half, off = Lmax - half;
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
                            if (c[0] < 0 || c[1] < 0 || c[2] < 0) { broken = true; ++P.incomplete; break; }
                            const uint64_t ck = encode_xyz(dims, c);
                            const int64_t j = find_key(C.keys, ck);
                            if (j >= 0 && !(C.flags[j] & (1 << 2))) {   // kNodeStatusCoarse2FineGhost_
                                P.contrib.push_back(j);
                                C.aux_mark[j] |= kAuxRecorded;   // read by the transfer
                            } else {
                                const int64_t k = find_key(F.keys, ck << dims);
                                if (k < 0) { broken = true; ++P.incomplete; break; }
                                P.contrib.push_back(-k - 1);
                                F.aux_mark[k] |= kAuxRecorded;
                            }
                            P.coeff.push_back(coeff);
                        }
                P.begin.push_back((int64_t)P.contrib.size());
                F.aux_mark[ti] |= kAuxRecorded | kAuxDeferOnTransfer;   // written by the transfer
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
            C.aux_mark[j] |= kAuxRecorded | kAuxDeferAlways;   // written after every coarse substep
            F.aux_mark[i] |= kAuxRecorded | kAuxDeferAlways;   // its collision's rho, u are read
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
    // DomainInfo::bool_periodic_domain_: any periodic face selects the overlap walk
    bool periodic = false;
    for (int d = 1; d <= 2 * dims; ++d) periodic = periodic || ctx->d.bc_scheme[d] == ZLH_BC_PERIODIC;
    const int Ls = M.stencil_dis;
    P.support = 1;
    for (int a = 0; a < dims; ++a) P.support *= 2 * Ls;
    P.node.assign(M.n * P.support, -1);
    P.weight.assign(M.n * P.support, 0.);
    P.scale.resize(M.n);
    const double dx = 1. / (double)(1u << level), dt = dx;
    for (int64_t m = 0; m < M.n; ++m) {
        double dis[3];
        int64_t idx[3], seq[3][8], twin[3][8];
        int valid = Ls;
        for (int a = 0; a < 3; ++a)
            for (int k = 0; k < 8; ++k) { seq[a][k] = 0; twin[a][k] = -1; }
        for (int a = 0; a < dims; ++a) {
            const double iref = M.xyz[m * 3 + a] / dx;
            idx[a] = (int64_t)(iref + 1e-6);
            dis[a] = iref - (double)idx[a];
            const int64_t lo = (int64_t)ctx->d.dom_min[a] << level, hi = (int64_t)ctx->d.dom_max[a] << level;
            const bool pl = ctx->d.bc_scheme[1 + 2 * a] == ZLH_BC_PERIODIC;
            const bool ph = ctx->d.bc_scheme[2 + 2 * a] == ZLH_BC_PERIODIC;
            valid = std::min(valid, periodic ? region_axis_twin(idx[a], Ls, lo, hi, pl, ph, seq[a], twin[a])
                                             : region_axis(idx[a], Ls, lo, hi, false, false, seq[a]));
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
                    L.aux_mark[j] |= kAuxRecorded | kAuxDeferAlways;   // forced before its collision
                }
            }
        }
        if (periodic) {
            // duplicates on periodic planes: every combination of an entry's coordinate or its
            // twin per axis other than the entry itself, with the entry's weight
            e = 0;
            for (int iz = 0; iz < nz; ++iz)
                for (int iy = 0; iy < n1; ++iy)
                    for (int ix = 0; ix < n1; ++ix, ++e) {
                        const int o[3] = {ix - valid, iy - valid, dims == 3 ? iz - valid : 0};
                        int64_t opt[3][2];
                        int nopt[3];
                        for (int a = 0; a < 3; ++a) {
                            opt[a][0] = a < dims ? seq[a][Ls + o[a]] : 0;
                            nopt[a] = 1;
                            if (a < dims && twin[a][Ls + o[a]] >= 0) opt[a][nopt[a]++] = twin[a][Ls + o[a]];
                        }
                        for (int cz = 0; cz < nopt[2]; ++cz)
                            for (int cy = 0; cy < nopt[1]; ++cy)
                                for (int cx = 0; cx < nopt[0]; ++cx) {
                                    if (cx == 0 && cy == 0 && cz == 0) continue;
                                    const int64_t c[3] = {opt[0][cx], opt[1][cy], opt[2][cz]};
                                    const int64_t j = find_key(L.keys, encode_xyz(dims, c));
                                    if (j < 0)
                                        return ctx->fail(ZLH_EINVAL, "marker %lld duplicate node absent at level %d",
                                                         (long long)m, level);
                                    P.dup_node.push_back(j);
                                    P.dup_entry.push_back(m * P.support + e);
                                    L.aux_mark[j] |= kAuxRecorded | kAuxDeferAlways;
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
// IEEE binary16 bits of w if w is exactly representable, else false
inline bool exact_half_bits(double w, uint16_t *bits) {
    const uint16_t sign = std::signbit(w) ? 0x8000u : 0u;
    const double a = std::fabs(w);
    if (a == 0.) { *bits = sign; return true; }
    int e;
    (void)std::frexp(a, &e);                      // a = f 2^e, f in [0.5, 1)
    const int E = e - 1;                          // a = 1.m 2^E
    if (E > 15) return false;
    if (E < -14) {                                // subnormal: a = m 2^-24
        const double m = std::ldexp(a, 24);
        if (m != std::floor(m) || m >= 1024.) return false;
        *bits = (uint16_t)(sign | (uint16_t)m);
        return true;
    }
    const double m = (std::ldexp(a, -E) - 1.) * 1024.;
    if (m != std::floor(m)) return false;
    *bits = (uint16_t)(sign | (uint16_t)((E + 15) << 10) | (uint16_t)m);
    return true;
}

inline zlh_status upload_transfer(zlh_ctx *ctx, int level) {
    LevelState &F = ctx->lv[level], &C = ctx->lv[level - 1];
    TransferPlan &P = F.xplan;
    TransferTables &X = F.xfer;
    zlh_status s;
    std::vector<int32_t> tgt, begin, src, src_ai, fb_ai, f2c_c, f2c_f;
    for (int64_t t : P.tgt) tgt.push_back(F.node_aux_h[t]);
    for (int64_t b : P.begin) begin.push_back((int32_t)b);
    // rows of src_f: the coarse sources by coarse aux index, then the fine fallback nodes by fine
    // aux index
    for (int64_t c : P.contrib) {
        if (c >= 0) src_ai.push_back(C.node_aux_h[c]);
        else fb_ai.push_back(F.node_aux_h[-c - 1]);
    }
    for (int32_t v : src_ai) if (v < 0) return ctx->fail(ZLH_EDEVICE, "transfer source not deferred");
    for (int32_t v : fb_ai) if (v < 0) return ctx->fail(ZLH_EDEVICE, "transfer fallback not deferred");
    for (auto *v : {&src_ai, &fb_ai}) {
        std::sort(v->begin(), v->end());
        v->erase(std::unique(v->begin(), v->end()), v->end());
    }
    X.n_fallback = 0;
    for (int64_t c : P.contrib) {
        if (c >= 0) {
            src.push_back((int32_t)(std::lower_bound(src_ai.begin(), src_ai.end(), C.node_aux_h[c]) - src_ai.begin()));
        } else {
            src.push_back((int32_t)(src_ai.size() + (std::lower_bound(fb_ai.begin(), fb_ai.end(), F.node_aux_h[-c - 1]) - fb_ai.begin())));
            ++X.n_fallback;
        }
    }
    for (size_t k = 0; k < P.f2c_coarse.size(); ++k) {
        f2c_c.push_back(C.node_aux_h[P.f2c_coarse[k]]);
        f2c_f.push_back(F.node_aux_h[P.f2c_fine[k]]);
    }
    for (int32_t v : tgt) if (v < 0) return ctx->fail(ZLH_EDEVICE, "transfer target not deferred");
    X.n_c2f = (int64_t)tgt.size();
    X.n_src = (int64_t)src_ai.size();
    X.n_fb = (int64_t)fb_ai.size();
    X.n_f2c = (int64_t)f2c_c.size();
    // gather order: ghost nodes by aux index (brick-major Morton order), cut into chunks of
    // kGatherChunk; inside a chunk by descending contributor count, then aux index.  A group
    // closes at 256 / PARTS nodes or when its rows and contributor words would pass the budget
    const int Qn = ctx->q;
    const int parts = (Qn + kGatherDirs - 1) / kGatherDirs, row_stride = parts * kGatherDirs;
    const size_t lds_budget = kGatherLdsBudget, row_bytes = (size_t)row_stride * ctx->tsize();
    const int group_max = 256 / parts;
    const int64_t nt = (int64_t)tgt.size();
    std::vector<int64_t> order(nt), chunk(nt);
    for (int64_t t = 0; t < nt; ++t) order[t] = t;
    std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return tgt[a] < tgt[b]; });
    for (int64_t k = 0; k < nt; ++k) chunk[order[k]] = k / kGatherChunk;
    std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
        if (chunk[a] != chunk[b]) return chunk[a] < chunk[b];
        const int32_t na = begin[a + 1] - begin[a], nb = begin[b + 1] - begin[b];
        return na != nb ? na > nb : tgt[a] < tgt[b];
    });
    auto lds_of = [&](size_t nrows, int64_t ncontrib) { return ((nrows * row_bytes + 15) & ~(size_t)15) + (size_t)ncontrib * 4; };
    std::vector<int32_t> tgt2, begin2(1, 0), group_tgt(1, 0), group_begin(1, 0), group_rows;
    std::vector<uint32_t> packed;
    packed.reserve(src.size());
    std::unordered_map<int32_t, int32_t> local;
    std::vector<int32_t> distinct;
    int64_t g_nc = 0;
    int g_nt = 0;
    size_t lds_max = 0;
    auto close_group = [&]() {
        lds_max = std::max(lds_max, lds_of(local.size(), g_nc));
        group_tgt.push_back((int32_t)tgt2.size());
        group_begin.push_back((int32_t)group_rows.size());
        local.clear();
        g_nc = 0;
        g_nt = 0;
    };
    for (int64_t t : order) {
        const int32_t k0 = begin[t], k1 = begin[t + 1];
        distinct.assign(src.begin() + k0, src.begin() + k1);
        std::sort(distinct.begin(), distinct.end());
        distinct.erase(std::unique(distinct.begin(), distinct.end()), distinct.end());
        size_t n_new = 0;
        for (int32_t r : distinct) n_new += local.count(r) ? 0 : 1;
        if (g_nt > 0 && (g_nt == group_max || local.size() + n_new > (size_t)kGatherMaxRows ||
                         lds_of(local.size() + n_new, g_nc + (k1 - k0)) > lds_budget))
            close_group();
        if (lds_of(distinct.size(), k1 - k0) > lds_budget || distinct.size() > (size_t)kGatherMaxRows)
            return ctx->fail(ZLH_EDEVICE, "interpolation stencil too large for LDS");
        for (int32_t r : distinct)
            if (local.emplace(r, (int32_t)local.size()).second) group_rows.push_back(r);
        for (int32_t k = k0; k < k1; ++k) {
            uint16_t bits;
            if (!exact_half_bits(P.coeff[k], &bits)) return ctx->fail(ZLH_EDEVICE, "interpolation weight not exact in FP16");
            packed.push_back((uint32_t)(local[src[k]] * row_stride) | ((uint32_t)bits << 16));
        }
        tgt2.push_back(tgt[t]);
        begin2.push_back((int32_t)packed.size());
        g_nc += k1 - k0;
        ++g_nt;
    }
    if (g_nt > 0) close_group();
    X.n_groups = (int64_t)group_tgt.size() - 1;
    X.lds_bytes = lds_max;
    X.staged_rows = (int64_t)group_rows.size();
    X.n_contrib = (int64_t)packed.size();
    std::vector<GatherGroup> groups(X.n_groups);
    for (int64_t g = 0; g < X.n_groups; ++g) {
        GatherGroup &G = groups[g];
        G.t0 = group_tgt[g];
        G.t1 = group_tgt[g + 1];
        G.c0 = begin2[G.t0];
        G.c1 = begin2[G.t1];
        G.r0 = group_begin[g];
        G.r1 = group_begin[g + 1];
        G.pad0 = G.pad1 = 0;
    }
    if ((s = upload(ctx, X.c2f_target, tgt2))) return s;
    if ((s = upload(ctx, X.c2f_begin, begin2))) return s;
    if ((s = upload(ctx, X.groups, groups))) return s;
    if ((s = upload(ctx, X.group_rows, group_rows))) return s;
    if ((s = upload(ctx, X.c2f_contrib, packed))) return s;
    if ((s = upload(ctx, X.src_coarse_ai, src_ai))) return s;
    if ((s = upload(ctx, X.fb_fine_ai, fb_ai))) return s;
    if ((s = dalloc(ctx, X.src_f, ctx->tsize() * ctx->q * std::max<int64_t>(X.n_src + X.n_fb, 1)))) return s;
    if ((s = dalloc(ctx, X.src_mid, ctx->tsize() * ctx->q * std::max<int64_t>(X.n_src + X.n_fb, 1)))) return s;
    if ((s = upload(ctx, X.f2c_coarse_ai, f2c_c))) return s;
    if ((s = upload(ctx, X.f2c_fine_ai, f2c_f))) return s;
    {
        std::vector<int32_t> a(tgt2.begin(), tgt2.end()), b(f2c_f.begin(), f2c_f.end());
        std::sort(a.begin(), a.end());
        std::sort(b.begin(), b.end());
        std::vector<int32_t> common;
        std::set_intersection(a.begin(), a.end(), b.begin(), b.end(), std::back_inserter(common));
        X.f2c_apart = common.empty();
    }
    F.xfer_writes = tgt;
    F.xfer_reads = fb_ai;
    F.xfer_reads.insert(F.xfer_reads.end(), f2c_f.begin(), f2c_f.end());
    if (level - 1 >= 1) {
        // the coarse level's own transfer (from level - 2) against this one's coarse side: this
        // one reads src_ai and writes f2c_c there
        std::vector<uint8_t> mark(std::max<int64_t>(C.naux, 1), 0);   // 1 read, 2 written
        for (int32_t i : C.xfer_reads) mark[i] |= 1;
        for (int32_t i : C.xfer_writes) mark[i] |= 2;
        C.xfer_overlap = false;
        for (int32_t i : src_ai) C.xfer_overlap |= (mark[i] & 2) != 0;
        for (int32_t i : f2c_c) C.xfer_overlap |= mark[i] != 0;
    }
    std::vector<uint8_t> macro_flag(std::max<int64_t>(F.naux, 1), 0);
    for (int32_t i : f2c_f) macro_flag[i] = 1;
    if ((s = upload(ctx, F.aux_macro_flag_d, macro_flag))) return s;
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
    std::vector<int32_t> ai(P.node.size(), -1), dup_ai(P.dup_node.size(), -1), reset;
    std::vector<uint8_t> seen(L.naux, 0);
    for (size_t k = 0; k < P.node.size(); ++k) {
        if (P.node[k] < 0) continue;
        ai[k] = L.node_aux_h[P.node[k]];
        if (ai[k] < 0) return ctx->fail(ZLH_EDEVICE, "immersed boundary node not deferred");
        if (!seen[ai[k]]) { seen[ai[k]] = 1; reset.push_back(ai[k]); }
    }
    // duplicates on periodic planes are reset and receive spread terms, but interpolate nothing
    for (size_t k = 0; k < P.dup_node.size(); ++k) {
        dup_ai[k] = L.node_aux_h[P.dup_node[k]];
        if (dup_ai[k] < 0) return ctx->fail(ZLH_EDEVICE, "immersed boundary node not deferred");
        if (!seen[dup_ai[k]]) { seen[dup_ai[k]] = 1; reset.push_back(dup_ai[k]); }
    }
    // supported nodes in aux order, so the reset's and the spread's row accesses coalesce
    std::sort(reset.begin(), reset.end());
    T.n_reset = (int64_t)reset.size();
    // node-major view of the marker table for the spread: the entries naming each supported
    // node, in ascending entry (marker) order
    if (ai.size() >= ((size_t)1 << 31)) return ctx->fail(ZLH_ENOMEM, "immersed boundary table too large");
    // each term's marker and coefficient weight x area / (dt dx^(dims-1)) (the same double
    // product the marker-major form forms), in ascending marker order per node
    std::vector<int32_t> pos(L.naux, -1), begin(reset.size() + 1, 0), marker;
    std::vector<double> coef;
    for (size_t r = 0; r < reset.size(); ++r) pos[reset[r]] = (int32_t)r;
    // terms: (node, marker, table entry) of every stencil entry and every duplicate, ordered by
    // marker per node (a node takes at most one term per marker)
    struct Term { int32_t a; int64_t m, k; };
    std::vector<Term> terms;
    for (size_t k = 0; k < ai.size(); ++k)
        if (ai[k] >= 0) terms.push_back({ai[k], (int64_t)(k / (size_t)P.support), (int64_t)k});
    for (size_t k = 0; k < dup_ai.size(); ++k)
        terms.push_back({dup_ai[k], P.dup_entry[k] / P.support, P.dup_entry[k]});
    std::stable_sort(terms.begin(), terms.end(), [](const Term &x, const Term &y) { return x.m < y.m; });
    for (const Term &t : terms) ++begin[pos[t.a] + 1];
    for (size_t r = 0; r < reset.size(); ++r) begin[r + 1] += begin[r];
    marker.resize(begin.back());
    coef.resize(begin.back());
    std::vector<int32_t> fill(begin.begin(), begin.end() - 1);
    for (const Term &t : terms) {
        const int32_t e = fill[pos[t.a]]++;
        marker[e] = (int32_t)t.m;
        coef[e] = P.weight[t.k] * P.scale[t.m];
    }
    if ((s = upload(ctx, T.spread_begin, begin))) return s;
    if ((s = upload(ctx, T.spread_marker, marker))) return s;
    if ((s = upload(ctx, T.spread_weight, coef))) return s;
    if ((s = upload(ctx, T.node_ai, ai))) return s;
    if ((s = upload(ctx, T.weight, P.weight))) return s;
    if ((s = upload(ctx, T.reset_ai, reset))) return s;
    if ((s = upload(ctx, T.marker_vel, L.ib.vel))) return s;
    if ((s = dalloc(ctx, T.marker_force, sizeof(double) * 3 * T.n_markers))) return s;
    if ((s = dalloc(ctx, T.node_u, sizeof(double) * 3 * std::max<int64_t>(L.naux, 1)))) return s;
    return ZLH_OK;
}

// ----------------------------------------------------------------------------------------------
// launches
// ----------------------------------------------------------------------------------------------
// parts: kTransferC2F (TransferInfoFromCoarseGrid) and / or kTransferF2C (TransferInfoToCoarseGrid),
// in that order (InformationFromGridOfDifferentLevel, lbm_solver_interface.cpp:224-238);
// kTransferMid makes the coarse-to-fine part the time-interpolated refill after an odd substep
// (rows into src_mid, src_f kept as the synchronised rows), kTransferRows only computes the
// synchronised rows (the earlier state of the next time-interpolated refill).
// kTransferCoarseRows forms only the rescaled coarse rows (they depend on the coarse level's
// state alone, so the coarse level's stream forms them right after its step), and
// kTransferRowsReady makes a coarse-to-fine part that finds them in place and forms only the
// fallback rows before the gather (not with kTransferMid)
constexpr int kTransferC2F = 1, kTransferF2C = 2, kTransferMid = 4, kTransferRows = 8, kTransferCoarseRows = 16,
              kTransferRowsReady = 32;
template <int D, typename T, bool SHIFT>
zlh_status launch_transfers_t(zlh_ctx *ctx, int level, int parts) {
    LevelState &F = ctx->lv[level], &C = ctx->lv[level - 1];
    TransferTables &X = F.xfer;
    const int comp = ctx->d.compressible;
    // tau ratios of PostStreamCoarse2Fine (coarse level) and PostStreamFine2Coarse (fine level),
    // CalRelaxationTimeRatio (lbm_collision_models.cpp:142-151)
    const double nu = ctx->d.nu0;
    const double dtc = 1. / (double)(1u << (level - 1)), dtf = 1. / (double)(1u << level);
    const double c2f = 0.5 * (nu * 3. / dtc * 2. + 0.5) / (nu * 3. / dtc + 0.5);
    const double f2c = 2. * (nu * 3. / dtf / 2 + 0.5) / (nu * 3. / dtf + 0.5);
    if ((parts & kTransferCoarseRows) && X.n_c2f > 0) {
        // the coarse rows alone, rows [0, n_src)
        if (X.n_src > 0)
            hipLaunchKernelGGL((k_c2f_sources<D, T, SHIFT, false>), dim3(blocks(X.n_src)), dim3(256), 0, ctx->stream,
                               int64_t(0), X.n_src, int64_t(0), X.src_coarse_ai.as<int32_t>(), C.aux_f_d.as<T>(),
                               C.naux, X.fb_fine_ai.as<int32_t>(), F.aux_f_d.as<T>(), F.naux, c2f, comp, nullptr,
                               X.src_f.as<T>());
        ZLH_TRY(hipGetLastError());
        return ZLH_OK;
    }
    if ((parts & (kTransferC2F | kTransferRows)) && X.n_c2f > 0) {
        const bool mid = (parts & kTransferMid) != 0;
        T *rows = mid ? X.src_mid.as<T>() : X.src_f.as<T>();
        // with kTransferRowsReady the coarse rows are in place: only the fallback rows remain
        const int64_t r_begin = (parts & kTransferRowsReady) ? X.n_src : 0;
        if (X.n_src + X.n_fb > r_begin) {
            if (mid)
                hipLaunchKernelGGL((k_c2f_sources<D, T, SHIFT, true>), dim3(blocks(X.n_src + X.n_fb - r_begin)),
                                   dim3(256), 0, ctx->stream, r_begin, X.n_src, X.n_fb, X.src_coarse_ai.as<int32_t>(),
                                   C.aux_f_d.as<T>(), C.naux, X.fb_fine_ai.as<int32_t>(), F.aux_f_d.as<T>(), F.naux,
                                   c2f, comp, X.src_f.as<T>(), rows);
            else
                hipLaunchKernelGGL((k_c2f_sources<D, T, SHIFT, false>), dim3(blocks(X.n_src + X.n_fb - r_begin)),
                                   dim3(256), 0, ctx->stream, r_begin, X.n_src, X.n_fb, X.src_coarse_ai.as<int32_t>(),
                                   C.aux_f_d.as<T>(), C.naux, X.fb_fine_ai.as<int32_t>(), F.aux_f_d.as<T>(), F.naux,
                                   c2f, comp, nullptr, rows);
        }
        const int64_t nb_gather = (X.n_groups + 7) / 8 * 8;
        if ((parts & kTransferC2F) && (parts & kTransferF2C) && X.n_f2c > 0 && X.f2c_apart) {
            // one launch for both halves (independent: see k_c2f_gather_f2c)
            hipLaunchKernelGGL((k_c2f_gather_f2c<D, T, SHIFT>), dim3((unsigned)(nb_gather + blocks(X.n_f2c))), dim3(256),
                               X.lds_bytes, ctx->stream, nb_gather, X.n_groups, X.groups.as<GatherGroup>(),
                               X.c2f_target.as<int32_t>(), X.c2f_begin.as<int32_t>(), X.c2f_contrib.as<uint32_t>(),
                               X.group_rows.as<int32_t>(), rows, F.aux_f_d.as<T>(), F.naux, X.n_f2c,
                               X.f2c_coarse_ai.as<int32_t>(), X.f2c_fine_ai.as<int32_t>(), F.aux_macro_d.as<double>(),
                               f2c, comp, C.aux_f_d.as<T>(), C.naux);
            ZLH_TRY(hipGetLastError());
            return ZLH_OK;
        }
        if (parts & kTransferC2F)
            hipLaunchKernelGGL((k_c2f_gather<D, T, SHIFT>), dim3((unsigned)nb_gather), dim3(256),
                               X.lds_bytes, ctx->stream, X.n_groups, X.groups.as<GatherGroup>(),
                               X.c2f_target.as<int32_t>(), X.c2f_begin.as<int32_t>(), X.c2f_contrib.as<uint32_t>(),
                               X.group_rows.as<int32_t>(), rows, F.aux_f_d.as<T>(), F.naux);
    }
    if ((parts & kTransferF2C) && X.n_f2c > 0)
        hipLaunchKernelGGL((k_f2c<D, T, SHIFT>), dim3(blocks(X.n_f2c)), dim3(256), 0, ctx->stream, X.n_f2c,
                           X.f2c_coarse_ai.as<int32_t>(), X.f2c_fine_ai.as<int32_t>(), F.aux_f_d.as<T>(), F.naux,
                           F.aux_macro_d.as<double>(), f2c, comp, C.aux_f_d.as<T>(), C.naux);
    ZLH_TRY(hipGetLastError());
    return ZLH_OK;
}

inline zlh_status launch_transfers(zlh_ctx *ctx, int level, int parts = kTransferC2F | kTransferF2C) {
    if (ctx->d.dims == 2)
        return ctx->fp64() ? launch_transfers_t<2, double, false>(ctx, level, parts)
             : ctx->bf16() ? launch_transfers_t<2, bf16_t, true>(ctx, level, parts)
                           : launch_transfers_t<2, float, true>(ctx, level, parts);
    return ctx->fp64() ? launch_transfers_t<3, double, false>(ctx, level, parts)
         : ctx->bf16() ? launch_transfers_t<3, bf16_t, true>(ctx, level, parts)
                       : launch_transfers_t<3, float, true>(ctx, level, parts);
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
                       B.marker_vel.as<double>(), B.stiffness, B.marker_force.as<double>());
    hipLaunchKernelGGL((k_ib_spread<D>), dim3(blocks(B.n_reset)), dim3(256), 0, ctx->stream, B.n_reset,
                       B.reset_ai.as<int32_t>(), B.spread_begin.as<int32_t>(), B.spread_marker.as<int32_t>(),
                       B.spread_weight.as<double>(), B.marker_force.as<double>(), L.aux_force_d.as<double>());
    ZLH_TRY(hipGetLastError());
    return ZLH_OK;
}

inline zlh_status launch_ib(zlh_ctx *ctx, int level) {
    if (ctx->lv[level].ibt.n_markers == 0) return ZLH_OK;
    if (ctx->d.dims == 2)
        return ctx->fp64() ? launch_ib_t<2, double, double, false>(ctx, level)
             : ctx->bf16() ? launch_ib_t<2, bf16_t, float, true>(ctx, level)
                           : launch_ib_t<2, float, float, true>(ctx, level);
    return ctx->fp64() ? launch_ib_t<3, double, double, false>(ctx, level)
         : ctx->bf16() ? launch_ib_t<3, bf16_t, float, true>(ctx, level)
                       : launch_ib_t<3, float, float, true>(ctx, level);
}

}  // namespace zlh
