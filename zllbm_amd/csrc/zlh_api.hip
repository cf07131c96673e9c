// zlh_api.hip -- host side of the C ABI declared in include/zllbm_hip.h.
//
// The context keeps, per level, the sorted Morton keys on the host (for table construction) and
// the brick-layout device state described in kernels.hpp.  Device tables are built by kernels
// (slot lookup by binary search over sorted brick keys, neighbour bricks from Morton neighbour
// arithmetic); the host only classifies which nodes take the deferred path and lays out the
// boundary / transfer / forcing tables that address them.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <iterator>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>
#include "zllbm_hip.h"
#include <functional>
#include "kernels.hpp"
#include "state.hpp"

namespace zlh {

// ----------------------------------------------------------------------------------------------
// Morton helpers (host and device): same bit layout as SFBitsetAux2D/3D (grid/morton_aux.cpp)
// ----------------------------------------------------------------------------------------------
__host__ __device__ inline uint64_t axis_mask(int dims, int axis) {
    uint64_t m = 0;
    const int bits = 64 / dims;
    for (int i = 0; i < bits; ++i) m |= 1ull << (i * dims + axis);
    return m;
}

__host__ __device__ inline uint32_t axis_coord(int dims, uint64_t key, int axis) {
    uint32_t v = 0;
    const int bits = 64 / dims;
    for (int i = 0; i < bits && i < 32; ++i) v |= (uint32_t)((key >> (i * dims + axis)) & 1ull) << i;
    return v;
}

__host__ __device__ inline uint64_t spread(int dims, uint32_t v, int axis) {
    uint64_t k = 0;
    const int bits = 64 / dims;
    for (int i = 0; i < bits && i < 32; ++i) k |= (uint64_t)((v >> i) & 1u) << (i * dims + axis);
    return k;
}

// neighbour along one axis without wrap-around: returns false past either end of the code space
__host__ __device__ inline bool step_axis(int dims, uint64_t key, int axis, int dir, uint64_t *out) {
    const uint64_t cur = axis_mask(dims, axis);
    const uint64_t a = key & cur;
    if (dir < 0) {
        if (a == 0) return false;
        *out = ((a - 1) & cur) | (key & ~cur);       // FindXNeg
        return true;
    }
    if (a == cur) return false;
    *out = (((key | ~cur) + 1) & cur) | (key & ~cur);  // FindXPos
    return true;
}

__host__ __device__ inline int64_t bsearch_u64(const uint64_t *a, int64_t n, uint64_t v) {
    int64_t lo = 0, hi = n - 1;
    while (lo <= hi) {
        const int64_t mid = (lo + hi) >> 1;
        const uint64_t x = a[mid];
        if (x == v) return mid;
        if (x < v) lo = mid + 1; else hi = mid - 1;
    }
    return -1;
}

// ----------------------------------------------------------------------------------------------
// K0 / K1: slot lookup and neighbour tables
// ----------------------------------------------------------------------------------------------
__global__ void k_node_slots(const uint64_t *keys, int64_t n, const uint64_t *bkeys, int64_t nbrick,
                             int32_t *node_slot, unsigned long long *occ, int32_t *slot_node) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t k = keys[i];
    const int64_t b = bsearch_u64(bkeys, nbrick, k >> 6);
    if (b < 0) { node_slot[i] = -1; return; }
    const int lane = (int)(k & 63ull);
    const int64_t slot = b * 64 + lane;
    node_slot[i] = (int32_t)slot;
    slot_node[slot] = (int32_t)i;
    atomicOr(&occ[b], 1ull << lane);
}

__global__ void k_brick_nbrs(int dims, const uint64_t *bkeys, int64_t nbrick, int32_t *nbr) {
    const int nn = dims == 3 ? 27 : 9;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nbrick * nn) return;
    const int64_t b = t / nn;
    const int k = (int)(t % nn);
    int o[3] = {k % 3 - 1, (k / 3) % 3 - 1, dims == 3 ? k / 9 - 1 : 0};
    uint64_t key = bkeys[b];
    bool ok = true;
    for (int a = 0; a < dims && ok; ++a)
        if (o[a] != 0) ok = step_axis(dims, key, a, o[a], &key);
    nbr[t] = ok ? (int32_t)bsearch_u64(bkeys, nbrick, key) : -1;
}

template <int D>
__global__ void k_missing(const int32_t *node_slot, int64_t n, const unsigned long long *occ,
                          const int32_t *nbr, uint32_t *missing) {
    using L = Lattice<D>;
    using B = Brick<D>;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t slot = node_slot[i];
    if (slot < 0) { missing[i] = 0xffffffffu; return; }
    const int64_t b = slot >> 6;
    const int lane = (int)(slot & 63);
    const int lx = B::lane_coord(lane, 0), ly = B::lane_coord(lane, 1), lz = D == 3 ? B::lane_coord(lane, 2) : 0;
    uint32_t m = 0;
    for (int q = 1; q < L::Q; ++q) {
        const int sx = lx - L::cx[q], sy = ly - L::cy[q], sz = lz - L::cz[q];
        const int ox = sx < 0 ? -1 : (sx >= B::edge ? 1 : 0);
        const int oy = sy < 0 ? -1 : (sy >= B::edge ? 1 : 0);
        const int oz = D == 3 ? (sz < 0 ? -1 : (sz >= B::edge ? 1 : 0)) : 0;
        const int k = D == 3 ? (oz + 1) * 9 + (oy + 1) * 3 + (ox + 1) : (oy + 1) * 3 + (ox + 1);
        const int sb = nbr[b * B::nnbr + k];
        const int sl = B::lane_of(sx & (B::edge - 1), sy & (B::edge - 1), sz & (B::edge - 1));
        if (sb < 0 || !((occ[sb] >> sl) & 1ull)) m |= 1u << q;
    }
    missing[i] = m;
}

template <int D>
__global__ void k_source_table(const int32_t *node_slot, int64_t n, const unsigned long long *occ,
                               const int32_t *nbr, const int32_t *slot_node, int32_t *src) {
    using L = Lattice<D>;
    using B = Brick<D>;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t slot = node_slot[i];
    const int64_t b = slot >> 6;
    const int lane = (int)(slot & 63);
    const int lx = B::lane_coord(lane, 0), ly = B::lane_coord(lane, 1), lz = D == 3 ? B::lane_coord(lane, 2) : 0;
    src[i * L::Q] = (int32_t)i;
    for (int q = 1; q < L::Q; ++q) {
        const int sx = lx - L::cx[q], sy = ly - L::cy[q], sz = lz - L::cz[q];
        const int ox = sx < 0 ? -1 : (sx >= B::edge ? 1 : 0);
        const int oy = sy < 0 ? -1 : (sy >= B::edge ? 1 : 0);
        const int oz = D == 3 ? (sz < 0 ? -1 : (sz >= B::edge ? 1 : 0)) : 0;
        const int k = D == 3 ? (oz + 1) * 9 + (oy + 1) * 3 + (ox + 1) : (oy + 1) * 3 + (ox + 1);
        const int sb = nbr[b * B::nnbr + k];
        const int sl = B::lane_of(sx & (B::edge - 1), sy & (B::edge - 1), sz & (B::edge - 1));
        src[i * L::Q + q] = (sb < 0 || !((occ[sb] >> sl) & 1ull)) ? -1 : slot_node[(int64_t)sb * 64 + sl];
    }
}

// ----------------------------------------------------------------------------------------------
// state upload / materialisation
// ----------------------------------------------------------------------------------------------
// f_in is [n][Q] double in key order (unshifted).  Aux nodes store it; bulk nodes store it in
// the "previous" buffer (for materialisation before the first step) and their first collision
// into the buffer the first fused kernel reads.
template <int D, typename T, typename C, bool SHIFT, int COLL>
__global__ void k_upload(BrickLevel<T> lv, const double *f_in, int64_t n, const int32_t *node_slot,
                         const int32_t *node_aux, T *fa, T *fb, T *aux_f, StepParams p) {
    constexpr int Q = Lattice<D>::Q;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    // lanes without a bulk node run the collision on the rest state (the FP32 learned closure
    // is wavefront-wide) and store nothing
    const bool active = i < n;
    if (!mlp_on_mfma<C, -1>() && !active) return;
    C f[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q)
        f[q] = active ? C(f_in[i * Q + q] - (SHIFT ? Lattice<D>::w[q] : 0.)) : C(SHIFT ? 0. : Lattice<D>::w[q]);
    const int32_t ai = active ? node_aux[i] : -1;
    if (ai >= 0) {
#pragma unroll
        for (int q = 0; q < Q; ++q) aux_f[q * lv.naux + ai] = T(f[q]);
    }
    const int64_t slot = active ? node_slot[i] : 0;
    if (active && ai < 0) {
#pragma unroll
        for (int q = 0; q < Q; ++q) fb[pop_at<Q>(q, slot)] = T(f[q]);
    }
    C F[3] = {C(p.const_force[0]), C(p.const_force[1]), C(p.const_force[2])};
    C rho, u[3];
    collide<D, SHIFT, COLL, C>(f, F, p, rho, u, p.mlp ? &p.mlp->ops : nullptr);
    if (!active || ai >= 0) return;
#pragma unroll
    for (int q = 0; q < Q; ++q) fa[pop_at<Q>(q, slot)] = T(f[q]);
}

// post-stream population q of one node (brick slot, aux index or -1), in double and unshifted:
// deferred nodes hold it in the aux array; a bulk node's is the pull of the post-collision value
// at x - c_q from the buffer the last fused kernel read (before the first step, the uploaded
// value itself)
template <int D, typename T, bool SHIFT>
__device__ __forceinline__ double post_stream_value(const BrickLevel<T> &lv, int64_t slot, int32_t ai, int q,
                                                    const T *last, const T *aux_f, int fresh) {
    using L = Lattice<D>;
    using B = Brick<D>;
    constexpr int Q = L::Q;
    double v;
    if (ai >= 0) {
        v = double(aux_f[q * lv.naux + ai]);
    } else if (fresh || q == 0) {
        v = double(last[pop_at<Q>(q, slot)]);
    } else {
        const int64_t b = slot >> 6;
        const int lane = (int)(slot & 63);
        const int lx = B::lane_coord(lane, 0), ly = B::lane_coord(lane, 1), lz = D == 3 ? B::lane_coord(lane, 2) : 0;
        const int sx = lx - L::cx[q], sy = ly - L::cy[q], sz = lz - L::cz[q];
        const int ox = sx < 0 ? -1 : (sx >= B::edge ? 1 : 0);
        const int oy = sy < 0 ? -1 : (sy >= B::edge ? 1 : 0);
        const int oz = D == 3 ? (sz < 0 ? -1 : (sz >= B::edge ? 1 : 0)) : 0;
        const int k = D == 3 ? (oz + 1) * 9 + (oy + 1) * 3 + (ox + 1) : (oy + 1) * 3 + (ox + 1);
        const int64_t sb = lv.nbr[b * B::nnbr + k];
        const int sl = B::lane_of(sx & (B::edge - 1), sy & (B::edge - 1), sz & (B::edge - 1));
        v = double(last[pop_at<Q>(q, sb * 64 + sl)]);
    }
    return v + (SHIFT ? L::w[q] : 0.);
}

// post-stream populations of every node, [Q][n] double in key order (unshifted)
template <int D, typename T, bool SHIFT>
__global__ void k_materialize(BrickLevel<T> lv, const int32_t *node_slot, const int32_t *node_aux, int64_t n,
                              const T *last, const T *aux_f, int fresh, double *out) {
    constexpr int Q = Lattice<D>::Q;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t ai = node_aux[i];
    const int64_t slot = node_slot[i];
    for (int q = 0; q < Q; ++q) out[q * n + i] = post_stream_value<D, T, SHIFT>(lv, slot, ai, q, last, aux_f, fresh);
}

__global__ void k_sum4(const double *fq, int64_t n, int Q, const int *cx, const int *cy, const int *cz,
                       double *acc) {
    __shared__ double sh[4][256];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double m = 0, jx = 0, jy = 0, jz = 0;
    if (i < n) {
        for (int q = 0; q < Q; ++q) {
            const double f = fq[q * n + i];
            m += f; jx += cx[q] * f; jy += cy[q] * f; jz += cz[q] * f;
        }
    }
    sh[0][threadIdx.x] = m; sh[1][threadIdx.x] = jx; sh[2][threadIdx.x] = jy; sh[3][threadIdx.x] = jz;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int c = 0; c < 4; ++c) sh[c][threadIdx.x] += sh[c][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int c = 0; c < 4; ++c) atomicAdd(&acc[c], sh[c][0]);
}

// ----------------------------------------------------------------------------------------------
// Domain boundary conditions on aux nodes, one launch per direction in the reference order
// (ComputeDomainBoundaryCondition, lbm_grid_interface.cpp:225-314)
// ----------------------------------------------------------------------------------------------
struct BcIdx {
    int n;
    int ind[8];
    int inv[8];
};

// BoundaryBounceBack{2D,3D} (lbm_bounce_back.cpp:29-91): f[in] = f_collide[out] - 6 rho0 w (c.u_w)
template <int D, typename T>
__global__ void k_bc_bounce_back(int64_t nb, const int32_t *node_ai, const int32_t *node_slot, const T *fa,
                                 int64_t nslot, T *aux_f, int64_t naux, BcIdx idx, double rho0, double ux,
                                 double uy, double uz) {
    using L = Lattice<D>;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nb) return;
    const int64_t ai = node_ai[t], slot = node_slot[t];
    for (int k = 0; k < idx.n; ++k) {
        const int o = idx.inv[k];
        const double cu = L::cx[o] * ux + L::cy[o] * uy + (D == 3 ? L::cz[o] * uz : 0.);
        aux_f[idx.ind[k] * naux + ai] = T(double(fa[pop_at<L::Q>(o, slot)]) - 6. * rho0 * L::w[o] * cu);
    }
}

// BoundaryPeriodic{2D,3D} (lbm_periodic.cpp:22-133): incoming directions copied from the twin
template <typename T>
__global__ void k_bc_periodic(int64_t nb, const int32_t *node_ai, const int32_t *twin_ai, T *aux_f,
                              int64_t naux, BcIdx idx) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nb) return;
    const int64_t ai = node_ai[t], tj = twin_ai[t];
    if (tj < 0) return;
    for (int k = 0; k < idx.n; ++k) aux_f[idx.ind[k] * naux + ai] = aux_f[idx.ind[k] * naux + tj];
}

// BoundaryNonEqExtrapolation{2D,3D} (lbm_noneq_extrapolation.cpp:29-144):
// f = feq(rho0, u_b) + f(inner) - feq(rho, u of inner)
template <int D, typename T, typename C, bool SHIFT>
__global__ void k_bc_nee(int64_t nb, const int32_t *node_ai, const int32_t *inner_ai, T *aux_f, int64_t naux,
                         const double *aux_force, StepParams p, double rho0, double ux, double uy, double uz,
                         const double *ub_node) {
    constexpr int Q = Lattice<D>::Q;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nb) return;
    const int64_t ai = node_ai[t], j = inner_ai[t];
    if (j < 0) return;
    C f[Q];
    for (int q = 0; q < Q; ++q) f[q] = C(aux_f[q * naux + j]);
    C F[3] = {C(p.const_force[0]), C(p.const_force[1]), C(p.const_force[2])};
    if (p.has_force && aux_force) {
        F[0] += C(aux_force[j * 3]); F[1] += C(aux_force[j * 3 + 1]); F[2] += C(aux_force[j * 3 + 2]);
    }
    C rho, drho, u[3];
    macro<D, SHIFT, C>(f, F, C(p.dt), p.has_force, p.compressible, rho, drho, u);
    const C usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    const C rb = C(rho0), drb = C(rho0 - 1.);
    // a profiled face carries one velocity per boundary node
    const C ubx = C(ub_node ? ub_node[t * 3] : ux), uby = C(ub_node ? ub_node[t * 3 + 1] : uy);
    const C ubz = C(D == 3 ? (ub_node ? ub_node[t * 3 + 2] : uz) : 0.);
    const C ubsq = ubx * ubx + uby * uby + ubz * ubz;
    for (int q = 0; q < Q; ++q) {
        const C fb = feq_stored<D, SHIFT, C>(q, rb, drb, ubx, uby, ubz, ubsq, p.compressible);
        const C fi = feq_stored<D, SHIFT, C>(q, rho, drho, u[0], u[1], u[2], usq, p.compressible);
        aux_f[q * naux + ai] = T(fb + f[q] - fi);
    }
}


// closure inputs, learned nu_t and fallback flag of each node's post-stream populations
// ([Q][n] double, unshifted), without forcing terms
template <int D>
__global__ void k_closure_features(const double *fq, int64_t n, double tau0, double cs18, const MlpDev *m,
                                   float *x_out, float *nu_out, int32_t *fb_out) {
    using L = Lattice<D>;
    constexpr int Q = L::Q;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double f[Q], feq[Q], rho, u[3];
    for (int q = 0; q < Q; ++q) f[q] = fq[q * n + i];
    double r = f[0], jx = 0, jy = 0, jz = 0;
    for (int q = 1; q < Q; ++q) { r += f[q]; jx += L::cx[q] * f[q]; jy += L::cy[q] * f[q]; jz += L::cz[q] * f[q]; }
    rho = r; u[0] = jx / r; u[1] = jy / r; u[2] = D == 3 ? jz / r : 0.;
    const double usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    for (int q = 0; q < Q; ++q) {
        const double cu = u[0] * L::cx[q] + u[1] * L::cy[q] + u[2] * L::cz[q];
        feq[q] = L::w[q] * rho * (1. + 3. * cu + 4.5 * cu * cu - 1.5 * usq);
    }
    double pi[6] = {0, 0, 0, 0, 0, 0};
    for (int q = 1; q < Q; ++q) {
        const double d = f[q] - feq[q];
        pi[0] += L::cx[q] * L::cx[q] * d; pi[1] += L::cy[q] * L::cy[q] * d; pi[2] += L::cz[q] * L::cz[q] * d;
        pi[3] += L::cx[q] * L::cy[q] * d; pi[4] += L::cx[q] * L::cz[q] * d; pi[5] += L::cy[q] * L::cz[q] * d;
    }
    float x[kMlpF];
    closure_features<double>(pi, rho, rho - 1., u, tau0, x);
    for (int k = 0; k < kMlpF; ++k) x_out[i * kMlpF + k] = x[k];
    bool fb = false;
    const double tt = mlp_tau_t<double>(*m, pi, rho, rho - 1., u, tau0, cs18, &fb);
    nu_out[i] = float(tt / 3.);
    fb_out[i] = fb ? 1 : 0;
}

}  // namespace zlh

// ==============================================================================================
// static device array of zlh_mlp_gradient_noalloc: the parameters, the samples and targets,
// and the gradient and loss accumulators, at fixed float offsets (each a multiple of 64 floats,
// so every region starts 256-byte aligned).  One copy per GPU, at its own address, for the life
// of the process.  Host code never uses the array's name as an address: the per-device address
// comes from hipGetSymbolAddress (scratch_setup).
// ==============================================================================================
namespace {
constexpr int64_t kNoAllocMaxSamples = ZLH_MLP_NOALLOC_MAX_SAMPLES;
constexpr int64_t kScratchParams = 0;                                     // [kMlpParams]
constexpr int64_t kScratchGrad = 1536;                                    // [kMlpParams]
constexpr int64_t kScratchLoss = 3072;                                    // [1]
constexpr int64_t kScratchT = 3136;                                       // [max samples]
constexpr int64_t kScratchX = kScratchT + kNoAllocMaxSamples;             // [max samples][kMlpF]
constexpr int64_t kScratchFloats = kScratchX + kNoAllocMaxSamples * zlh::kMlpF;
static_assert(zlh::kMlpParams <= kScratchGrad - kScratchParams && zlh::kMlpParams <= kScratchLoss - kScratchGrad,
              "network parameters fit their regions");
static_assert(kScratchT % 64 == 0 && kScratchX % 64 == 0, "regions are 256-byte aligned");
}  // namespace

static __device__ __align__(256) float scratch[kScratchFloats];

// clears the accumulator regions (a kernel, not a memset, so that the call stays capturable)
__global__ void k_scratch_clear(float *p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

// ==============================================================================================
// host context
// ==============================================================================================
using namespace zlh;

namespace {

#define ZLH_TRY(expr)                                                                         \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) return ctx->fail(ZLH_EDEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

}  // namespace

struct zlh_ctx {
    zlh_solver_desc d{};
    int q = 0;
    hipStream_t stream = nullptr;
    std::vector<LevelState> lv;
    bool built = false, initialised = false;
    std::string err;
    bool timing = false;
    struct TimedLaunch { hipEvent_t start, stop; int level; };
    std::vector<TimedLaunch> events;
    double timing_ms[16] = {};       // per level
    int64_t timing_launches[16] = {};
    DevBuf lattice_d;
    DevBuf mlp_d;           // MlpDev of the learned closure
    bool mlp_set = false;
    int mlp_train_form = ZLH_MLP_FP32;   // kernel of zlh_mlp_gradient / zlh_mlp_train
    int n_cu = 0;           // compute units of the context's device
    int device = 0;         // the device the context was created on
    float *scratch_addr = nullptr;   // that device's address of the static no-allocation array
    // two background steps captured as one graph, per combination of the levels' buffer
    // indices (two steps flip every level's buffers an even number of times, so replays chain)
    bool use_graphs = true;
    std::map<uint64_t, hipGraphExec_t> graph_exec;
    // time marching runs each level on a stream of its own; events order a transfer after the
    // coarse step it pairs with, and the coarse level's next step after the transfer
    int bc_profile[7] = {};   // per face: ZLH_PROFILE_*
    std::map<std::pair<int, int>, std::vector<double>> bc_node_u;   // (level, face) -> [n][3] per-node velocities
    int interp_method[16] = {};   // per fine level: ZLH_INTERP_LAGRANGIAN or ZLH_INTERP_LINEAR
    bool time_interp = false;     // refill ghost layers after odd substeps too (time-interpolated)
    int mrt_transform = ZLH_TRANSFORM_VALU;   // FP32 learned-closure kernels: ZLH_TRANSFORM_*
    bool level_streams = true;
    std::vector<hipStream_t> lstream;
    std::vector<hipEvent_t> ev_step, ev_xfer, ev_join, ev_rows;
    hipEvent_t ev_fork = nullptr;
    void drop_graphs() {
        for (auto &g : graph_exec) (void)hipGraphExecDestroy(g.second);
        graph_exec.clear();
    }

    zlh_status fail(zlh_status code, const char *fmt, ...) {
        char buf[1024];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        err = buf;
        return code;
    }
    bool fp64() const { return d.precision == ZLH_FP64; }
    bool bf16() const { return d.precision == ZLH_BF16_SHIFTED; }
    size_t tsize() const { return fp64() ? 8 : (bf16() ? 2 : 4); }
};

namespace {

// device address of the static no-allocation array, looked up once per device with that device
// current (each GPU holds its own copy of the array at its own address)
constexpr int kScratchMaxDevices = 64;
std::mutex g_scratch_mutex;
float *g_scratch_addr[kScratchMaxDevices] = {};

zlh_status scratch_setup(zlh_ctx *ctx) {
    if (ctx->device < 0 || ctx->device >= kScratchMaxDevices)
        return ctx->fail(ZLH_EDEVICE, "device %d: no-allocation array supports devices 0..%d", ctx->device,
                         kScratchMaxDevices - 1);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_addr[ctx->device]) {
        int cur = -1;
        ZLH_TRY(hipGetDevice(&cur));
        if (cur != ctx->device) ZLH_TRY(hipSetDevice(ctx->device));
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != ctx->device) ZLH_TRY(hipSetDevice(cur));
        if (e != hipSuccess || !p)
            return ctx->fail(ZLH_EDEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_scratch_addr[ctx->device] = static_cast<float *>(p);
    }
    ctx->scratch_addr = g_scratch_addr[ctx->device];
    return ZLH_OK;
}

zlh_status dalloc(zlh_ctx *ctx, DevBuf &b, size_t bytes, bool zero = true) {
    if (b.p) { (void)hipFree(b.p); b.p = nullptr; }
    b.bytes = bytes;
    if (bytes == 0) bytes = 16;
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) { b.p = nullptr; return ctx->fail(ZLH_ENOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e)); }
    if (zero) ZLH_TRY(hipMemsetAsync(b.p, 0, bytes, ctx->stream));
    return ZLH_OK;
}

template <typename X>
zlh_status upload(zlh_ctx *ctx, DevBuf &b, const std::vector<X> &v) {
    zlh_status s = dalloc(ctx, b, v.size() * sizeof(X), false);
    if (s) return s;
    if (!v.empty()) ZLH_TRY(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(X), hipMemcpyHostToDevice, ctx->stream));
    ZLH_TRY(hipStreamSynchronize(ctx->stream));
    return ZLH_OK;
}

inline unsigned blocks(int64_t n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

// FP32 -> BF16 bit pattern, round to nearest even (NaN stays NaN)
// FP32 -> FP16 bit pattern, round to nearest even, with subnormals, overflow to infinity
inline uint16_t f16_bits(float v) {
    uint32_t x;
    std::memcpy(&x, &v, 4);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const int32_t ef = (int32_t)((x >> 23) & 0xffu);
    uint32_t m = x & 0x7fffffu;
    if (ef == 0xff) return (uint16_t)(sign | 0x7c00u | (m ? 0x200u : 0u));
    const int32_t e = ef - 127 + 15;
    if (e >= 31) return (uint16_t)(sign | 0x7c00u);
    if (e <= 0) {
        if (e < -10) return (uint16_t)sign;
        m |= 0x800000u;
        const int shift = 14 - e;
        uint32_t hm = m >> shift;
        const uint32_t rem = m & ((1u << shift) - 1u), half = 1u << (shift - 1);
        if (rem > half || (rem == half && (hm & 1u))) ++hm;
        return (uint16_t)(sign | hm);
    }
    uint32_t out = sign | ((uint32_t)e << 10) | (m >> 13);
    const uint32_t rem = m & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (out & 1u))) ++out;     // a carry rounds into the exponent
    return (uint16_t)out;
}

inline uint16_t bf16_bits(float v) {
    uint32_t u;
    std::memcpy(&u, &v, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

StepParams level_params(const zlh_ctx *ctx, int level) {
    StepParams p{};
    const double nu = ctx->d.nu0;
    p.dt = 1. / (double)(1u << level);
    p.tau0 = nu * 3. / p.dt + 0.5;   // CalRelaxationTime (lbm_collision_models.cpp:114-118)
    p.cs18 = 18. * ctx->d.cs * ctx->d.cs;
    for (int a = 0; a < 3; ++a) p.const_force[a] = ctx->d.has_force ? ctx->d.const_force[a] : 0.;
    p.les = ctx->d.les == ZLH_LES_SMAGORINSKY ? 1 : (ctx->d.les == ZLH_LES_MLP ? 2 : 0);
    p.mlp_operands = ctx->d.mlp_operands;
    p.mlp = ctx->mlp_d.as<MlpDev>();
    p.les_tau_mode = ctx->d.les_tau_mode;
    p.has_force = ctx->d.has_force;
    p.compressible = ctx->d.compressible;
    return p;
}

template <typename T>
BrickLevel<T> view(const LevelState &L) {
    BrickLevel<T> v;
    v.nbrick = L.nbrick;
    v.nslot = L.nbrick * 64;
    v.naux = L.naux;
    v.occ = L.occ_d.as<uint64_t>();
    v.nbr = L.nbr_d.as<int32_t>();
    v.auxmask = L.auxmask_d.as<uint64_t>();
    v.defmask = L.defmask_d[0].as<uint64_t>();
    v.auxbase = L.auxbase_d.as<int32_t>();
    v.aux_missing = L.aux_missing_d.as<uint32_t>();
    v.aux_slot = L.aux_slot_d.as<int32_t>();
    return v;
}

// incoming / inverse direction lists of a face (GetBoundaryNInverseIndices,
// lbm_boundary_conditions.cpp:102-156)
BcIdx bc_indices(int dims, int dir) {
    BcIdx r{};
    const int axis = (dir - 1) / 2;
    const int sign_in = (dir % 2 == 1) ? 1 : -1;
    const int Q = dims == 2 ? 9 : 19;
    for (int q = 1; q < Q; ++q) {
        const int c = dims == 2 ? (axis == 0 ? Lattice<2>::cx[q] : Lattice<2>::cy[q])
                                : (axis == 0 ? Lattice<3>::cx[q] : axis == 1 ? Lattice<3>::cy[q] : Lattice<3>::cz[q]);
        if (c == sign_in) {
            r.ind[r.n] = q;
            r.inv[r.n] = dims == 2 ? opposite<2>(q) : opposite<3>(q);
            ++r.n;
        }
    }
    return r;
}

int64_t find_key(const std::vector<uint64_t> &keys, uint64_t k) {
    auto it = std::lower_bound(keys.begin(), keys.end(), k);
    return (it != keys.end() && *it == k) ? (int64_t)(it - keys.begin()) : -1;
}

}  // namespace

#include "amr_ib.hpp"

// ==============================================================================================
// templated dispatch helpers
// ==============================================================================================
namespace {

// learned-closure kernel with R rounds of four bricks per workgroup
template <int D, typename T, typename C, bool SHIFT, int COLL, int V, int R>
void launch_mlp_rounds(zlh_ctx *ctx, const BrickLevel<T> &v, T *fa, T *fb, T *aux_f, const StepParams &sp) {
    const int64_t ng = (v.nbrick + 4 * R - 1) / (4 * R);
    hipLaunchKernelGGL((k_stream_collide_mlp<D, T, C, SHIFT, COLL, V, R>), dim3((unsigned)((ng + 7) / 8 * 8)), dim3(256),
                       0, ctx->stream, v.nbrick, v.naux, v.occ, v.auxmask, v.auxbase, v.nbr, v.aux_missing, fa, fb,
                       aux_f, sp);
}

// one round per workgroup while the grid is at most two waves of resident workgroups (the
// latency of a short grid is one brick's), two on larger levels (each wave's second pull then
// overlaps its first brick's network; more rounds measured slower on a 3M-node level)
template <int D, typename T, typename C, bool SHIFT, int COLL, int V>
void launch_mlp(zlh_ctx *ctx, const BrickLevel<T> &v, T *fa, T *fb, T *aux_f, const StepParams &sp) {
    static int per_cu = [] {
        int n = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_stream_collide_mlp<D, T, C, SHIFT, COLL, V, 1>, 256, 0) !=
                hipSuccess || n < 1)
            n = 1;
        return n;
    }();
    const int64_t resident = (int64_t)per_cu * ctx->n_cu;
    if ((v.nbrick + 3) / 4 <= 2 * resident) launch_mlp_rounds<D, T, C, SHIFT, COLL, V, 1>(ctx, v, fa, fb, aux_f, sp);
    else launch_mlp_rounds<D, T, C, SHIFT, COLL, V, 2>(ctx, v, fa, fb, aux_f, sp);
}

// bulk kernel variant: collision operator x (sub-grid model, forcing)
template <int D, typename T, typename C, bool SHIFT, int COLL>
void launch_stream_collide_v(zlh_ctx *ctx, int var, unsigned grid, const BrickLevel<T> &v, T *fa, T *fb, T *aux_f,
                             const StepParams &sp) {
#define ZLH_SC(V)                                                                                             \
    case V:                                                                                                   \
        if constexpr (mlp_on_mfma<C, V>() && sizeof(T) == 4) {                                                \
            launch_mlp<D, T, C, SHIFT, COLL, V>(ctx, v, fa, fb, aux_f, sp);                                   \
        } else if constexpr (mlp_on_mfma<C, V>() || V >= kVarMrtMfma) {                                       \
            /* BF16 storage with the learned closure: refused by zlh_create; kVarMrtMfma is only */           \
            /* selected for the FP32 learned closure */                                                       \
        } else {                                                                                              \
            hipLaunchKernelGGL((k_stream_collide<D, T, C, SHIFT, COLL, V>), dim3(grid), dim3(256), 0, ctx->stream, \
                               v, fa, fb, aux_f, sp);                                                         \
        }                                                                                                     \
        break;
    switch (var) {
        ZLH_SC(0) ZLH_SC(1) ZLH_SC(2) ZLH_SC(3) ZLH_SC(4) ZLH_SC(5) ZLH_SC(6) ZLH_SC(7) ZLH_SC(8) ZLH_SC(9)
        ZLH_SC(20) ZLH_SC(21) ZLH_SC(24) ZLH_SC(25)
        default: break;
    }
#undef ZLH_SC
}

template <int D, typename T, typename C, bool SHIFT, int COLL>
void launch_aux_collide_v(zlh_ctx *ctx, int var, LevelState &L, const BrickLevel<T> &v, const int32_t *list,
                          int64_t nlist, T *aux_f, const double *aux_force, T *fa) {
#define ZLH_AC(V)                                                                                             \
    case V:                                                                                                   \
        hipLaunchKernelGGL((k_aux_collide<D, T, C, SHIFT, COLL, V>), dim3(blocks(nlist)), dim3(256), 0,        \
                           ctx->stream, v, list, nlist, aux_f, aux_force, fa, L.aux_macro_d.as<double>(),     \
                           L.aux_macro_flag_d.as<uint8_t>(), L.sp);                                           \
        break;
    switch (var) {
        ZLH_AC(0) ZLH_AC(1) ZLH_AC(2) ZLH_AC(3) ZLH_AC(4) ZLH_AC(5) ZLH_AC(6) ZLH_AC(7) ZLH_AC(8) ZLH_AC(9)
        default:
            if constexpr (std::is_same<C, float>::value && sizeof(T) == 4) {
                switch (var) { ZLH_AC(20) ZLH_AC(21) ZLH_AC(24) ZLH_AC(25) default: break; }
            }
            break;
    }
#undef ZLH_AC
}

template <int D, typename T, typename C, bool SHIFT>
void launch_stream_collide(zlh_ctx *ctx, int coll, int var, unsigned grid, const BrickLevel<T> &v, T *fa, T *fb,
                           T *aux_f, const StepParams &sp) {
    if (coll == ZLH_SRT) launch_stream_collide_v<D, T, C, SHIFT, 0>(ctx, var, grid, v, fa, fb, aux_f, sp);
    else launch_stream_collide_v<D, T, C, SHIFT, 1>(ctx, var, grid, v, fa, fb, aux_f, sp);
}

template <int D, typename T, typename C, bool SHIFT>
zlh_status launch_step_kernels(zlh_ctx *ctx, int level, int coll) {
    LevelState &L = ctx->lv[level];
    BrickLevel<T> v = view<T>(L);
    T *fa = L.f_d[L.cur].as<T>(), *fb = L.f_d[L.cur ^ 1].as<T>();
    T *aux_f = L.aux_f_d.as<T>();
    const double *aux_force = L.aux_force_d.as<double>();
    int var = 2 * les_code(L.sp.les, L.sp.les_tau_mode, L.sp.mlp_operands) + (L.sp.has_force ? 1 : 0);
    if (std::is_same<C, float>::value && sizeof(T) == 4 && (var >> 1 == 2 || var >> 1 == 4) &&
        ctx->mrt_transform == ZLH_TRANSFORM_MFMA)
        var += kVarMrtMfma;
    // the aux collision takes the nodes deferred by the last substep (every recorded node after
    // an upload)
    const int32_t *list = nullptr;
    int64_t nlist = L.naux;
    if (L.pending != kPendingAll) {
        const int pk = L.pending == kPendingTransfer ? 1 : 0;
        list = L.deflist_d[pk].as<int32_t>();
        nlist = L.ndef[pk];
    }
    if (nlist > 0) {
        if (coll == ZLH_SRT) launch_aux_collide_v<D, T, C, SHIFT, 0>(ctx, var, L, v, list, nlist, aux_f, aux_force, fa);
        else launch_aux_collide_v<D, T, C, SHIFT, 1>(ctx, var, L, v, list, nlist, aux_f, aux_force, fa);
    }
    // this substep defers the ghost targets too when the transfer into the level follows it
    // (after every substep with time interpolation)
    const int kind = (level > 0 && (L.last_tsl % 2 == 0 || ctx->time_interp)) ? 1 : 0;
    v.defmask = L.defmask_d[kind].as<uint64_t>();
    const int64_t groups = (L.nbrick + 3) / 4;
    const unsigned grid = (unsigned)(((groups + 7) / 8) * 8);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (ctx->timing) {
        ZLH_TRY(hipEventCreate(&e0));
        ZLH_TRY(hipEventCreate(&e1));
        ZLH_TRY(hipEventRecord(e0, ctx->stream));
    }
    launch_stream_collide<D, T, C, SHIFT>(ctx, coll, var, grid, v, fa, fb, aux_f, L.sp);
    if (ctx->timing) {
        ZLH_TRY(hipEventRecord(e1, ctx->stream));
        ctx->events.push_back({e0, e1, level});
    }
    ZLH_TRY(hipGetLastError());
    // domain boundaries in the reference order: every minimum face, then every maximum face
    const int dims = D;
    int order[6];
    for (int a = 0; a < dims; ++a) { order[a] = 1 + 2 * a; order[dims + a] = 2 + 2 * a; }
    for (int k = 0; k < 2 * dims; ++k) {
        const int dir = order[k];
        if (L.bc_n[dir] == 0) continue;
        const BcIdx idx = bc_indices(dims, dir);
        const double *bv = ctx->d.bc_value[dir];
        switch (ctx->d.bc_scheme[dir]) {
            case ZLH_BC_BOUNCE_BACK:
                hipLaunchKernelGGL((k_bc_bounce_back<D, T>), dim3(blocks(L.bc_n[dir])), dim3(256), 0, ctx->stream,
                                   L.bc_n[dir], L.bc_ai_d[dir].as<int32_t>(), L.bc_slot_d[dir].as<int32_t>(), fa,
                                   L.nbrick * 64, aux_f, L.naux, idx, ctx->d.rho0, bv[0], bv[1], bv[2]);
                break;
            case ZLH_BC_PERIODIC:
                hipLaunchKernelGGL((k_bc_periodic<T>), dim3(blocks(L.bc_n[dir])), dim3(256), 0, ctx->stream,
                                   L.bc_n[dir], L.bc_ai_d[dir].as<int32_t>(), L.bc_partner_d[dir].as<int32_t>(),
                                   aux_f, L.naux, idx);
                break;
            case ZLH_BC_NEE:
                hipLaunchKernelGGL((k_bc_nee<D, T, C, SHIFT>), dim3(blocks(L.bc_n[dir])), dim3(256), 0, ctx->stream,
                                   L.bc_n[dir], L.bc_ai_d[dir].as<int32_t>(), L.bc_partner_d[dir].as<int32_t>(),
                                   aux_f, L.naux, aux_force, L.sp, ctx->d.rho0, bv[0], bv[1], bv[2],
                                   L.bc_u_d[dir].p ? L.bc_u_d[dir].as<double>() : nullptr);
                break;
            default:
                break;
        }
    }
    ZLH_TRY(hipGetLastError());
    L.cur ^= 1;
    L.fresh = false;
    L.pending = kind ? kPendingTransfer : kPendingPlain;
    return ZLH_OK;
}

template <int D, typename T, typename C, bool SHIFT>
zlh_status dispatch_upload(zlh_ctx *ctx, int level, const double *f_dev) {
    LevelState &L = ctx->lv[level];
    BrickLevel<T> v = view<T>(L);
    T *fa = L.f_d[L.cur].as<T>(), *fb = L.f_d[L.cur ^ 1].as<T>();
    const int coll = ctx->d.collision[level];
    if (coll == ZLH_SRT)
        hipLaunchKernelGGL((k_upload<D, T, C, SHIFT, 0>), dim3(blocks(L.n)), dim3(256), 0, ctx->stream, v, f_dev, L.n,
                           L.node_slot_d.as<int32_t>(), L.node_aux_d.as<int32_t>(), fa, fb, L.aux_f_d.as<T>(), L.sp);
    else
        hipLaunchKernelGGL((k_upload<D, T, C, SHIFT, 1>), dim3(blocks(L.n)), dim3(256), 0, ctx->stream, v, f_dev, L.n,
                           L.node_slot_d.as<int32_t>(), L.node_aux_d.as<int32_t>(), fa, fb, L.aux_f_d.as<T>(), L.sp);
    ZLH_TRY(hipGetLastError());
    L.fresh = true;
    L.pending = kPendingAll;
    L.last_tsl = 0;
    return ZLH_OK;
}

template <int D, typename T, bool SHIFT>
zlh_status dispatch_materialize(zlh_ctx *ctx, int level, double *out_dev) {
    LevelState &L = ctx->lv[level];
    BrickLevel<T> v = view<T>(L);
    // the buffer read by the last fused kernel (or, before any step, the uploaded populations)
    const T *last = L.f_d[L.cur ^ 1].as<T>();
    hipLaunchKernelGGL((k_materialize<D, T, SHIFT>), dim3(blocks(L.n)), dim3(256), 0, ctx->stream, v,
                       L.node_slot_d.as<int32_t>(), L.node_aux_d.as<int32_t>(), L.n, last, L.aux_f_d.as<T>(),
                       L.fresh ? 1 : 0, out_dev);
    ZLH_TRY(hipGetLastError());
    return ZLH_OK;
}

zlh_status do_step_kernels(zlh_ctx *ctx, int level) {
    const int coll = ctx->d.collision[level];
    if (ctx->d.dims == 2) {
        return ctx->fp64() ? launch_step_kernels<2, double, double, false>(ctx, level, coll)
             : ctx->bf16() ? launch_step_kernels<2, bf16_t, float, true>(ctx, level, coll)
                           : launch_step_kernels<2, float, float, true>(ctx, level, coll);
    }
    return ctx->fp64() ? launch_step_kernels<3, double, double, false>(ctx, level, coll)
         : ctx->bf16() ? launch_step_kernels<3, bf16_t, float, true>(ctx, level, coll)
                       : launch_step_kernels<3, float, float, true>(ctx, level, coll);
}

zlh_status do_upload(zlh_ctx *ctx, int level, const double *f_dev) {
    if (ctx->d.dims == 2)
        return ctx->fp64() ? dispatch_upload<2, double, double, false>(ctx, level, f_dev)
             : ctx->bf16() ? dispatch_upload<2, bf16_t, float, true>(ctx, level, f_dev)
                           : dispatch_upload<2, float, float, true>(ctx, level, f_dev);
    return ctx->fp64() ? dispatch_upload<3, double, double, false>(ctx, level, f_dev)
         : ctx->bf16() ? dispatch_upload<3, bf16_t, float, true>(ctx, level, f_dev)
                       : dispatch_upload<3, float, float, true>(ctx, level, f_dev);
}

zlh_status do_materialize(zlh_ctx *ctx, int level, double *out_dev) {
    if (ctx->d.dims == 2)
        return ctx->fp64() ? dispatch_materialize<2, double, false>(ctx, level, out_dev)
             : ctx->bf16() ? dispatch_materialize<2, bf16_t, true>(ctx, level, out_dev)
                           : dispatch_materialize<2, float, true>(ctx, level, out_dev);
    return ctx->fp64() ? dispatch_materialize<3, double, false>(ctx, level, out_dev)
         : ctx->bf16() ? dispatch_materialize<3, bf16_t, true>(ctx, level, out_dev)
                       : dispatch_materialize<3, float, true>(ctx, level, out_dev);
}

// feq of given macros (double, unshifted) [n][Q]
void host_feq(int dims, double rho, const double *u, int compressible, double *feq) {
    const int Q = dims == 2 ? 9 : 19;
    const double ux = u[0], uy = u[1], uz = dims == 3 ? u[2] : 0.;
    const double usq = ux * ux + uy * uy + uz * uz;
    for (int q = 0; q < Q; ++q) {
        const double cx = dims == 2 ? Lattice<2>::cx[q] : Lattice<3>::cx[q];
        const double cy = dims == 2 ? Lattice<2>::cy[q] : Lattice<3>::cy[q];
        const double cz = dims == 2 ? 0. : Lattice<3>::cz[q];
        const double w = dims == 2 ? Lattice<2>::w[q] : Lattice<3>::w[q];
        const double cu = ux * cx + uy * cy + uz * cz;
        feq[q] = compressible ? w * rho * (1. + 3. * cu + 4.5 * cu * cu - 1.5 * usq)
                              : w * (rho + 3. * cu + 4.5 * cu * cu - 1.5 * usq);
    }
}

}  // namespace

// ==============================================================================================
// build: device tables + host classification
// ==============================================================================================
namespace {

// phase 1: device slot / neighbour tables and the absent-source masks
zlh_status build_level_slots(zlh_ctx *ctx, int level) {
    LevelState &L = ctx->lv[level];
    const int dims = ctx->d.dims;
    const int NN = dims == 3 ? 27 : 9;
    L.n = (int64_t)L.keys.size();
    // brick keys: keys are sorted and the brick key is a prefix of the Morton code, so the
    // brick keys of sorted nodes are non-decreasing
    L.bkeys.clear();
    for (uint64_t k : L.keys)
        if (L.bkeys.empty() || L.bkeys.back() != (k >> 6)) L.bkeys.push_back(k >> 6);
    L.nbrick = (int64_t)L.bkeys.size();
    const int64_t nslot = L.nbrick * 64;
    int64_t max_bricks = 0;
    (void)zlh_max_level_bricks(&ctx->d, &max_bricks);
    if (L.nbrick > max_bricks)
        return ctx->fail(ZLH_EINVAL, "level %d too large: %lld bricks of 64 nodes, at most %lld with these settings",
                         level, (long long)L.nbrick, (long long)max_bricks);
    zlh_status s;
    if ((s = upload(ctx, L.keys_d, L.keys))) return s;
    if ((s = upload(ctx, L.bkeys_d, L.bkeys))) return s;
    if ((s = dalloc(ctx, L.node_slot_d, sizeof(int32_t) * L.n))) return s;
    if ((s = dalloc(ctx, L.slot_node_d, sizeof(int32_t) * nslot))) return s;
    ZLH_TRY(hipMemsetAsync(L.slot_node_d.p, 0xff, sizeof(int32_t) * nslot, ctx->stream));
    if ((s = dalloc(ctx, L.occ_d, sizeof(uint64_t) * L.nbrick))) return s;
    if ((s = dalloc(ctx, L.nbr_d, sizeof(int32_t) * L.nbrick * NN))) return s;
    hipLaunchKernelGGL(k_node_slots, dim3(blocks(L.n)), dim3(256), 0, ctx->stream, L.keys_d.as<uint64_t>(), L.n,
                       L.bkeys_d.as<uint64_t>(), L.nbrick, L.node_slot_d.as<int32_t>(),
                       L.occ_d.as<unsigned long long>(), L.slot_node_d.as<int32_t>());
    hipLaunchKernelGGL(k_brick_nbrs, dim3(blocks(L.nbrick * NN)), dim3(256), 0, ctx->stream, dims,
                       L.bkeys_d.as<uint64_t>(), L.nbrick, L.nbr_d.as<int32_t>());
    ZLH_TRY(hipGetLastError());
    DevBuf missing_d;
    if ((s = dalloc(ctx, missing_d, sizeof(uint32_t) * L.n))) return s;
    if (dims == 2)
        hipLaunchKernelGGL(k_missing<2>, dim3(blocks(L.n)), dim3(256), 0, ctx->stream, L.node_slot_d.as<int32_t>(),
                           L.n, L.occ_d.as<unsigned long long>(), L.nbr_d.as<int32_t>(), missing_d.as<uint32_t>());
    else
        hipLaunchKernelGGL(k_missing<3>, dim3(blocks(L.n)), dim3(256), 0, ctx->stream, L.node_slot_d.as<int32_t>(),
                           L.n, L.occ_d.as<unsigned long long>(), L.nbr_d.as<int32_t>(), missing_d.as<uint32_t>());
    ZLH_TRY(hipGetLastError());
    L.missing_h.resize(L.n);
    L.node_slot_h.resize(L.n);
    ZLH_TRY(hipMemcpyAsync(L.missing_h.data(), missing_d.p, sizeof(uint32_t) * L.n, hipMemcpyDeviceToHost, ctx->stream));
    ZLH_TRY(hipMemcpyAsync(L.node_slot_h.data(), L.node_slot_d.p, sizeof(int32_t) * L.n, hipMemcpyDeviceToHost,
                           ctx->stream));
    ZLH_TRY(hipStreamSynchronize(ctx->stream));
    for (int64_t i = 0; i < L.n; ++i)
        if (L.node_slot_h[i] < 0) return ctx->fail(ZLH_EDEVICE, "slot lookup failed for node %lld", (long long)i);
    // nodes with an absent pull source keep that population from the previous substep
    L.aux_mark.assign(L.n, 0);
    for (int64_t i = 0; i < L.n; ++i)
        if (L.missing_h[i]) L.aux_mark[i] = kAuxRecorded | kAuxDeferAlways;
    // domain boundaries (CheckIfNodeOutsideCubicDomain, grid/grid_info.cpp:84-142) and the
    // nodes their conditions read
    for (int d = 0; d < 7; ++d) { L.bnode[d].clear(); L.bpart[d].clear(); }
    for (int64_t i = 0; i < L.n; ++i) {
        for (int a = 0; a < dims; ++a) {
            const uint32_t c = axis_coord(dims, L.keys[i], a);
            const uint32_t lo = ctx->d.dom_min[a] << level, hi = ctx->d.dom_max[a] << level;
            for (int side = 0; side < 2; ++side) {
                if (c != (side ? hi : lo)) continue;
                const int dir = 1 + 2 * a + side;
                int64_t partner = -1;
                const int scheme = ctx->d.bc_scheme[dir];
                if (scheme == ZLH_BC_PERIODIC) {
                    // (key & others) | bound of the opposite face (lbm_periodic.cpp:77-133)
                    const uint64_t tw = spread(dims, side ? lo : hi, a);
                    partner = find_key(L.keys, (L.keys[i] & ~axis_mask(dims, a)) | tw);
                } else if (scheme == ZLH_BC_NEE) {
                    uint64_t ik;
                    if (step_axis(dims, L.keys[i], a, side ? -1 : 1, &ik)) partner = find_key(L.keys, ik);
                }
                L.bnode[dir].push_back(i);
                L.bpart[dir].push_back(partner);
                L.aux_mark[i] |= kAuxRecorded | kAuxDeferAlways;
                if (partner >= 0) L.aux_mark[partner] |= kAuxRecorded;   // read by the condition
            }
        }
    }
    return ZLH_OK;
}

// phase 2: deferred-node arrays, boundary tables and population storage
zlh_status build_level_aux(zlh_ctx *ctx, int level) {
    LevelState &L = ctx->lv[level];
    const int dims = ctx->d.dims;
    const int Q = ctx->q;
    const int64_t nslot = L.nbrick * 64;
    zlh_status s;
    std::vector<int64_t> aux_nodes;
    for (int64_t i = 0; i < L.n; ++i)
        if (L.aux_mark[i]) aux_nodes.push_back(i);
    std::sort(aux_nodes.begin(), aux_nodes.end(),
              [&](int64_t a, int64_t b) { return L.node_slot_h[a] < L.node_slot_h[b]; });
    L.naux = (int64_t)aux_nodes.size();
    std::vector<uint64_t> auxmask(L.nbrick, 0);
    std::vector<int32_t> auxbase(L.nbrick, 0), aux_slot(L.naux);
    std::vector<uint32_t> aux_missing(L.naux);
    L.node_aux_h.assign(L.n, -1);
    for (int64_t k = 0; k < L.naux; ++k) {
        const int64_t i = aux_nodes[k];
        const int32_t slot = L.node_slot_h[i];
        auxmask[slot >> 6] |= 1ull << (slot & 63);
        aux_slot[k] = slot;
        aux_missing[k] = L.missing_h[i];
        L.node_aux_h[i] = (int32_t)k;
    }
    int32_t acc = 0;
    for (int64_t b = 0; b < L.nbrick; ++b) {
        auxbase[b] = acc;
        acc += __builtin_popcountll(auxmask[b]);
    }
    // deferred nodes per kind of substep: always-deferred roles on both, ghost targets only on the
    // substeps that the transfer into this level follows
    std::vector<uint64_t> defmask[2] = {std::vector<uint64_t>(L.nbrick, 0), std::vector<uint64_t>(L.nbrick, 0)};
    std::vector<int32_t> deflist[2];
    for (int64_t k = 0; k < L.naux; ++k) {
        const uint8_t m = L.aux_mark[aux_nodes[k]];
        const int32_t slot = L.node_slot_h[aux_nodes[k]];
        for (int kind = 0; kind < 2; ++kind) {
            const uint8_t deferred = kind ? (kAuxDeferAlways | kAuxDeferOnTransfer) : kAuxDeferAlways;
            if (m & deferred) {
                defmask[kind][slot >> 6] |= 1ull << (slot & 63);
                deflist[kind].push_back((int32_t)k);
            }
        }
    }
    for (int kind = 0; kind < 2; ++kind) {
        L.ndef[kind] = (int64_t)deflist[kind].size();
        if ((s = upload(ctx, L.defmask_d[kind], defmask[kind]))) return s;
        if ((s = upload(ctx, L.deflist_d[kind], deflist[kind]))) return s;
    }
    L.pending = kPendingAll;
    if ((s = upload(ctx, L.auxmask_d, auxmask))) return s;
    if ((s = upload(ctx, L.auxbase_d, auxbase))) return s;
    if ((s = upload(ctx, L.aux_slot_d, aux_slot))) return s;
    if ((s = upload(ctx, L.aux_missing_d, aux_missing))) return s;
    if ((s = upload(ctx, L.node_aux_d, L.node_aux_h))) return s;
    for (int dir = 1; dir <= 2 * dims; ++dir) {
        std::vector<int32_t> ai, part, slot;
        for (size_t k = 0; k < L.bnode[dir].size(); ++k) {
            ai.push_back(L.node_aux_h[L.bnode[dir][k]]);
            part.push_back(L.bpart[dir][k] >= 0 ? L.node_aux_h[L.bpart[dir][k]] : -1);
            slot.push_back(L.node_slot_h[L.bnode[dir][k]]);
        }
        L.bc_n[dir] = (int64_t)ai.size();
        const auto given = ctx->bc_node_u.find({level, dir});
        if (given != ctx->bc_node_u.end()) {
            if ((int64_t)given->second.size() != 3 * L.bc_n[dir])
                return ctx->fail(ZLH_EINVAL, "level %d face %d: %lld boundary velocities for %lld boundary nodes", level,
                                 dir, (long long)(given->second.size() / 3), (long long)L.bc_n[dir]);
            if ((s = upload(ctx, L.bc_u_d[dir], given->second))) return s;
        } else if (ctx->bc_profile[dir] == ZLH_PROFILE_PARABOLIC && ctx->d.bc_scheme[dir] == ZLH_BC_NEE) {
            // the face velocity times prod over the transverse axes of 4 t (1 - t)
            const int axis = (dir - 1) / 2;
            std::vector<double> ub;
            for (int64_t i : L.bnode[dir]) {
                double fac = 1.0;
                for (int a = 0; a < dims; ++a) {
                    if (a == axis) continue;
                    const double lo = (double)((int64_t)ctx->d.dom_min[a] << level);
                    const double hi = (double)((int64_t)ctx->d.dom_max[a] << level);
                    const double t = ((double)axis_coord(dims, L.keys[i], a) - lo) / (hi - lo);
                    fac *= 4.0 * t * (1.0 - t);
                }
                for (int c = 0; c < 3; ++c) ub.push_back(ctx->d.bc_value[dir][c] * fac);
            }
            if ((s = upload(ctx, L.bc_u_d[dir], ub))) return s;
        }
        if ((s = upload(ctx, L.bc_ai_d[dir], ai))) return s;
        if ((s = upload(ctx, L.bc_partner_d[dir], part))) return s;
        if ((s = upload(ctx, L.bc_slot_d[dir], slot))) return s;
    }
    const size_t ts = ctx->tsize();
    if ((s = dalloc(ctx, L.f_d[0], ts * Q * nslot))) return s;
    if ((s = dalloc(ctx, L.f_d[1], ts * Q * nslot))) return s;
    // one spare element past [Q][naux]: the learned-closure kernel parks a deferred node's
    // missing directions there instead of branching around their stores
    if ((int64_t)Q * L.naux + 1 >= ((int64_t)1 << 31)) return ctx->fail(ZLH_ENOMEM, "level %d: aux array too large", level);
    if ((s = dalloc(ctx, L.aux_f_d, ts * (Q * std::max<int64_t>(L.naux, 1) + 1)))) return s;
    if ((s = dalloc(ctx, L.aux_macro_d, sizeof(double) * 4 * std::max<int64_t>(L.naux, 1)))) return s;
    if ((s = dalloc(ctx, L.aux_macro_flag_d, std::max<int64_t>(L.naux, 1)))) return s;
    if ((s = dalloc(ctx, L.aux_force_d, sizeof(double) * 3 * std::max<int64_t>(L.naux, 1)))) return s;
    L.sp = level_params(ctx, level);
    ZLH_TRY(hipStreamSynchronize(ctx->stream));
    return ZLH_OK;
}

}  // namespace

// ==============================================================================================
// C ABI
// ==============================================================================================
namespace {
zlh_status refresh_transfer_rows(zlh_ctx *ctx, int coarse);
}  // namespace

extern "C" {

zlh_status zlh_create(const zlh_solver_desc *desc, zlh_ctx **out) {
    if (!desc || !out) return ZLH_EINVAL;
    *out = nullptr;
    if ((desc->dims != 2 && desc->dims != 3) || desc->max_level < 0 || desc->max_level > 15) return ZLH_EINVAL;
    if (desc->precision != ZLH_FP64 && desc->precision != ZLH_FP32_SHIFTED && desc->precision != ZLH_BF16_SHIFTED)
        return ZLH_EINVAL;
    // the learned closure's fused kernel stages 4-byte populations with direct-to-LDS loads
    if (desc->precision == ZLH_BF16_SHIFTED && desc->les == ZLH_LES_MLP) return ZLH_EINVAL;
    if (!(desc->nu0 > 0.)) return ZLH_EINVAL;
    for (int l = 0; l <= desc->max_level; ++l)
        if (desc->collision[l] != ZLH_SRT && desc->collision[l] != ZLH_MRT) return ZLH_EINVAL;
    if (desc->les != ZLH_LES_NONE && desc->les != ZLH_LES_SMAGORINSKY && desc->les != ZLH_LES_MLP) return ZLH_EINVAL;
    if (desc->mlp_operands != ZLH_MLP_OPERANDS_BF16 && desc->mlp_operands != ZLH_MLP_OPERANDS_FP16) return ZLH_EINVAL;
    for (int dir = 1; dir <= 2 * desc->dims; ++dir)
        if (desc->bc_scheme[dir] < ZLH_BC_BOUNCE_BACK || desc->bc_scheme[dir] > ZLH_BC_NEE) return ZLH_EINVAL;
    for (int a = 0; a < desc->dims; ++a)
        if (desc->dom_min[a] < 1 || desc->dom_max[a] <= desc->dom_min[a]) return ZLH_EINVAL;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return ZLH_EDEVICE;
    std::unique_ptr<zlh_ctx> ctx(new zlh_ctx());
    ctx->d = *desc;
    ctx->q = desc->dims == 2 ? 9 : 19;
    ctx->lv.resize(desc->max_level + 1);
    if (hipStreamCreate(&ctx->stream) != hipSuccess) return ZLH_EDEVICE;
    const int nl = desc->max_level + 1;
    ctx->lstream.assign(nl, nullptr);
    ctx->ev_step.assign(nl, nullptr);
    ctx->ev_xfer.assign(nl, nullptr);
    ctx->ev_join.assign(nl, nullptr);
    ctx->ev_rows.assign(nl, nullptr);
    if (hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming) != hipSuccess) return ZLH_EDEVICE;
    for (int l = 0; l < nl; ++l)
        if (hipStreamCreate(&ctx->lstream[l]) != hipSuccess ||
            hipEventCreateWithFlags(&ctx->ev_step[l], hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&ctx->ev_xfer[l], hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&ctx->ev_join[l], hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&ctx->ev_rows[l], hipEventDisableTiming) != hipSuccess)
            return ZLH_EDEVICE;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ctx->n_cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ctx->n_cu < 1)
        return ZLH_EDEVICE;
    ctx->device = dev;
    if (scratch_setup(ctx.get())) return ZLH_EDEVICE;
    if (dalloc(ctx.get(), ctx->mlp_d, sizeof(MlpDev))) return ZLH_ENOMEM;
    *out = ctx.release();
    return ZLH_OK;
}

void zlh_destroy(zlh_ctx *ctx) {
    if (!ctx) return;
    // every stream drains before the device buffers are freed
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (hipStream_t s : ctx->lstream)
        if (s) (void)hipStreamSynchronize(s);
    for (auto &e : ctx->events) { (void)hipEventDestroy(e.start); (void)hipEventDestroy(e.stop); }
    ctx->drop_graphs();
    ctx->lv.clear();
    for (hipStream_t s : ctx->lstream)
        if (s) (void)hipStreamDestroy(s);
    for (auto *v : {&ctx->ev_step, &ctx->ev_xfer, &ctx->ev_join, &ctx->ev_rows})
        for (hipEvent_t e : *v)
            if (e) (void)hipEventDestroy(e);
    if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *zlh_last_error(const zlh_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

zlh_status zlh_max_level_bricks(const zlh_solver_desc *desc, int64_t *max_bricks) {
    if (!desc || !max_bricks || (desc->dims != 2 && desc->dims != 3)) return ZLH_EINVAL;
    // slot and aux indices are 32-bit: 64 nbrick 19 < 2^31
    int64_t m = (((int64_t)1 << 31) / 19 - 1) / 64;
    // the FP32 learned closure's fused kernel pulls through 32-bit byte offsets
    // nbrick Q 64 4 (pull_to_lds in kernels.hpp) and stores deferred nodes through 32-bit byte
    // offsets up to Q naux 4 <= nbrick Q 256: nbrick Q 256 < 2^32
    if (desc->les == ZLH_LES_MLP && desc->precision == ZLH_FP32_SHIFTED) {
        const int64_t q = desc->dims == 2 ? 9 : 19;
        m = std::min<int64_t>(m, (((int64_t)1 << 32) - 1) / (q * 64 * 4));
    }
    *max_bricks = m;
    return ZLH_OK;
}

zlh_status zlh_set_level_nodes(zlh_ctx *ctx, int level, const uint64_t *keys, const int32_t *flags, int64_t n) {
    if (!ctx) return ZLH_EINVAL;
    if (level < 0 || level > ctx->d.max_level) return ctx->fail(ZLH_EINVAL, "level %d out of range", level);
    if (n < 1 || !keys) return ctx->fail(ZLH_EINVAL, "empty level %d", level);
    const int dims = ctx->d.dims;
    const uint32_t top = dims == 3 ? (1u << 21) - 1 : 0xffffffffu;
    for (int64_t i = 0; i < n; ++i) {
        if (i > 0 && keys[i] <= keys[i - 1])
            return ctx->fail(ZLH_EINVAL, "keys of level %d not sorted/unique at %lld", level, (long long)i);
        for (int a = 0; a < dims; ++a) {
            const uint32_t c = axis_coord(dims, keys[i], a);
            if (c == 0 || c == top) return ctx->fail(ZLH_EINVAL, "node %lld has index %u on axis %d", (long long)i, c, a);
        }
    }
    LevelState &L = ctx->lv[level];
    L.keys.assign(keys, keys + n);
    L.flags.assign(n, 0);
    if (flags) L.flags.assign(flags, flags + n);
    L.have = true;
    ctx->built = false;
    ctx->initialised = false;
    return ZLH_OK;
}

zlh_status zlh_set_interface_layer(zlh_ctx *ctx, int level, int kind, int layer, const uint64_t *keys, int64_t n) {
    if (!ctx) return ZLH_EINVAL;
    if (level < 0 || level > ctx->d.max_level || kind < 0 || kind > 3 || layer < 0 || layer > 2 || n < 0 || (n && !keys))
        return ctx->fail(ZLH_EINVAL, "bad interface layer arguments");
    LevelState &L = ctx->lv[level];
    if (!L.have) return ctx->fail(ZLH_ESTATE, "level %d nodes not set", level);
    std::vector<uint64_t> v(keys, keys + n);
    std::sort(v.begin(), v.end());
    for (uint64_t k : v)
        if (find_key(L.keys, k) < 0) return ctx->fail(ZLH_EINVAL, "layer node not in level %d", level);
    L.layers[{kind, layer}] = std::move(v);
    ctx->built = false;
    return ZLH_OK;
}

zlh_status zlh_ib_set_markers(zlh_ctx *ctx, int level, const double *xyz, const double *area, const double *vel,
                              int64_t n, int stencil_dis, double stiffness) {
    if (!ctx) return ZLH_EINVAL;
    if (level < 0 || level > ctx->d.max_level || n < 0 || stencil_dis < 1 || stencil_dis > 2 || (n && (!xyz || !area || !vel)))
        return ctx->fail(ZLH_EINVAL, "bad immersed boundary arguments");
    if (n && !ctx->d.has_force) return ctx->fail(ZLH_EINVAL, "immersed boundary needs has_force");
    IbMarkers &m = ctx->lv[level].ib;
    m.n = n;
    m.xyz.assign(xyz, xyz + 3 * n);
    m.vel.assign(vel, vel + 3 * n);
    m.area.assign(area, area + n);
    m.stencil_dis = stencil_dis;
    m.stiffness = stiffness;
    ctx->built = false;
    return ZLH_OK;
}

zlh_status zlh_build(zlh_ctx *ctx) {
    if (!ctx) return ZLH_EINVAL;
    ctx->drop_graphs();
    for (int l = 0; l <= ctx->d.max_level; ++l)
        if (!ctx->lv[l].have) return ctx->fail(ZLH_ESTATE, "level %d nodes not set", l);
    zlh_status s;
    for (int l = 0; l <= ctx->d.max_level; ++l)
        if ((s = build_level_slots(ctx, l))) return s;
    for (int l = 1; l <= ctx->d.max_level; ++l)
        if ((s = plan_transfer(ctx, l))) return s;
    for (int l = 0; l <= ctx->d.max_level; ++l)
        if ((s = plan_ib(ctx, l))) return s;
    for (int l = 0; l <= ctx->d.max_level; ++l)
        if ((s = build_level_aux(ctx, l))) return s;
    for (int l = 1; l <= ctx->d.max_level; ++l)
        if ((s = upload_transfer(ctx, l))) return s;
    for (int l = 0; l <= ctx->d.max_level; ++l)
        if ((s = upload_ib(ctx, l))) return s;
    ctx->built = true;
    return ZLH_OK;
}

zlh_status zlh_upload_populations(zlh_ctx *ctx, int level, const double *f) {
    if (!ctx) return ZLH_EINVAL;
    if (!ctx->built) return ctx->fail(ZLH_ESTATE, "build before upload");
    if (level < 0 || level > ctx->d.max_level || !f) return ctx->fail(ZLH_EINVAL, "bad upload arguments");
    LevelState &L = ctx->lv[level];
    DevBuf tmp;
    zlh_status s = dalloc(ctx, tmp, sizeof(double) * L.n * ctx->q, false);
    if (s) return s;
    ZLH_TRY(hipMemcpyAsync(tmp.p, f, sizeof(double) * L.n * ctx->q, hipMemcpyHostToDevice, ctx->stream));
    L.cur = 0;
    if ((s = do_upload(ctx, level, tmp.as<double>()))) return s;
    if ((s = refresh_transfer_rows(ctx, level))) return s;
    ZLH_TRY(hipStreamSynchronize(ctx->stream));
    return ZLH_OK;
}

zlh_status zlh_init_macros(zlh_ctx *ctx, int level, const double *rho, const double *u) {
    if (!ctx) return ZLH_EINVAL;
    if (!ctx->built) return ctx->fail(ZLH_ESTATE, "build before init");
    if (level < 0 || level > ctx->d.max_level || !rho || !u) return ctx->fail(ZLH_EINVAL, "bad init arguments");
    LevelState &L = ctx->lv[level];
    const int dims = ctx->d.dims, Q = ctx->q;
    std::vector<double> f((size_t)L.n * Q);
    for (int64_t i = 0; i < L.n; ++i) host_feq(dims, rho[i], u + i * dims, ctx->d.compressible, &f[i * Q]);
    return zlh_upload_populations(ctx, level, f.data());
}

zlh_status zlh_init_feq(zlh_ctx *ctx) {
    if (!ctx) return ZLH_EINVAL;
    if (!ctx->built) return ctx->fail(ZLH_ESTATE, "build before init");
    for (int l = 0; l <= ctx->d.max_level; ++l) {
        LevelState &L = ctx->lv[l];
        std::vector<double> rho(L.n, ctx->d.rho0), u((size_t)L.n * ctx->d.dims);
        for (int64_t i = 0; i < L.n; ++i)
            for (int a = 0; a < ctx->d.dims; ++a) u[i * ctx->d.dims + a] = ctx->d.u0[a];
        zlh_status s = zlh_init_macros(ctx, l, rho.data(), u.data());
        if (s) return s;
    }
    ctx->initialised = true;
    return ZLH_OK;
}

zlh_status zlh_step_level(zlh_ctx *ctx, int level, int time_step_level) {
    if (!ctx) return ZLH_EINVAL;
    if (!ctx->built) return ctx->fail(ZLH_ESTATE, "build before stepping");
    if (level < 0 || level > ctx->d.max_level || time_step_level < 1)
        return ctx->fail(ZLH_EINVAL, "bad level / substep");
    if (ctx->d.les == ZLH_LES_MLP && !ctx->mlp_set) return ctx->fail(ZLH_ESTATE, "learned closure not set (zlh_mlp_set)");
    zlh_status s;
    // immersed-boundary forcing (UpdateCriterion, lbm_grid_interface.cpp:61-95)
    if ((s = launch_ib(ctx, level))) return s;
    // collision, streaming and domain boundaries (RunSolverOnGivenGrid + ComputeDomainBoundaryCondition)
    ctx->lv[level].last_tsl = time_step_level;
    if ((s = do_step_kernels(ctx, level))) return s;
    return ZLH_OK;
}

}  // extern "C"

namespace {
zlh_status transfer_checked(zlh_ctx *ctx, int fine_level, int parts) {
    if (!ctx) return ZLH_EINVAL;
    if (!ctx->built) return ctx->fail(ZLH_ESTATE, "build before transfer");
    if (fine_level < 1 || fine_level > ctx->d.max_level) return ctx->fail(ZLH_EINVAL, "bad fine level %d", fine_level);
    // InformationFromGridOfDifferentLevel (lbm_solver_interface.cpp:224-238) exchanges after the
    // fine level's even substeps, whose fused kernel defers the ghost targets this writes.  With
    // time interpolation every substep defers them, and after an odd one (the coarse level a
    // step ahead) the coarse-to-fine part is the time-interpolated refill and the fine-to-coarse
    // part has nothing to give.
    if (ctx->lv[fine_level].last_tsl % 2 != 0) {
        if (!ctx->time_interp || !(parts & kTransferC2F))
            return ctx->fail(ZLH_ESTATE, "transfer %s level %d after its odd substep %d",
                             (parts & kTransferC2F) ? "into" : "from", fine_level, ctx->lv[fine_level].last_tsl);
        return launch_transfers(ctx, fine_level, kTransferC2F | kTransferMid);
    }
    return launch_transfers(ctx, fine_level, parts);
}

// the synchronised coarse rows of every refined level whose coarse level is `coarse` (-1: all),
// from the coarse state as it is now: the earlier state of the next time-interpolated refill
zlh_status refresh_transfer_rows(zlh_ctx *ctx, int coarse) {
    if (!ctx->time_interp) return ZLH_OK;
    for (int l = 1; l <= ctx->d.max_level; ++l) {
        if (coarse >= 0 && l - 1 != coarse) continue;
        zlh_status s = launch_transfers(ctx, l, kTransferRows);
        if (s) return s;
    }
    return ZLH_OK;
}
}  // namespace

extern "C" {

zlh_status zlh_transfer_levels(zlh_ctx *ctx, int fine_level) {
    return transfer_checked(ctx, fine_level, kTransferC2F | kTransferF2C);
}

zlh_status zlh_transfer_from_coarse(zlh_ctx *ctx, int fine_level) {
    return transfer_checked(ctx, fine_level, kTransferC2F);
}

zlh_status zlh_transfer_to_coarse(zlh_ctx *ctx, int fine_level) {
    return transfer_checked(ctx, fine_level, kTransferF2C);
}

}  // extern "C"

namespace {

zlh_status march_launches(zlh_ctx *ctx, int nsteps) {
    // level order: a level step is followed by two substeps of the next finer level
    std::vector<int> order;
    std::function<void(int)> rec = [&](int l) {
        order.push_back(l);
        if (l < ctx->d.max_level) { rec(l + 1); rec(l + 1); }
    };
    rec(0);
    const int nl = ctx->d.max_level + 1;
    // InformationFromGridOfDifferentLevel (lbm_solver_interface.cpp:224-238): refined levels
    // exchange with their parent when their substep count is even.  With level streams, level l
    // runs on lstream[l]; the transfer into level l (on lstream[l]) waits for the latest step of
    // level l - 1 (and, where the host found that they share nodes, for the latest transfers into
    // levels l + 1 and l - 1), and level l - 1's next step waits for it.  Nothing else is shared between levels, so a
    // coarse level's step overlaps the finer levels' substeps.  Timed runs keep one stream, so
    // that per-launch events time a kernel alone.
    const bool split = ctx->level_streams && !ctx->timing && nl > 1;
    // With level streams (and no time interpolation), level l - 1's stream forms the coarse rows
    // of the next transfer into l right after its step and the transfer into l - 1 (the coarse
    // state those rows read is then final: nothing else writes level l - 1 before that transfer
    // into l, whose own fine-to-coarse copy comes after them); the transfer into l waits for
    // them (ev_rows[l], which also follows ev_step[l - 1]) and forms only the fallback rows, so
    // the coarse rows leave the finer level's chain.  Level l - 1's step itself waits for the
    // previous transfer into l, whose gather reads the rows it overwrites.
    const bool early_rows = split && !ctx->time_interp;
    hipStream_t main = ctx->stream;
    if (split) {
        ZLH_TRY(hipEventRecord(ctx->ev_fork, main));
        for (int l = 0; l < nl; ++l) ZLH_TRY(hipStreamWaitEvent(ctx->lstream[l], ctx->ev_fork, 0));
    }
    std::vector<uint8_t> after_xfer(nl, 0);   // level l's next step must wait for ev_xfer[l + 1]
    std::vector<uint8_t> xfer_recorded(nl, 0);   // ev_xfer[l] recorded in this call
    zlh_status s = ZLH_OK;
    hipError_t e = hipSuccess;
    for (int t = 0; t < nsteps && !s && e == hipSuccess; ++t) {
        std::vector<int> tsl(nl, 0);
        for (int l : order) {
            ++tsl[l];
            if (split) {
                ctx->stream = ctx->lstream[l];
                if (after_xfer[l] && (e = hipStreamWaitEvent(ctx->stream, ctx->ev_xfer[l + 1], 0)) != hipSuccess) break;
                after_xfer[l] = 0;
            }
            if ((s = zlh_step_level(ctx, l, tsl[l]))) break;
            if (split && (e = hipEventRecord(ctx->ev_step[l], ctx->stream)) != hipSuccess) break;
            // (with time interpolation, after odd substeps too: the refill reads the coarse level's
            // latest state, which its next step overwrites, so the same ordering holds)
            if (l > 0 && (tsl[l] % 2 == 0 || ctx->time_interp)) {
                if (split && (e = hipStreamWaitEvent(ctx->stream, early_rows ? ctx->ev_rows[l] : ctx->ev_step[l - 1],
                                                     0)) != hipSuccess)
                    break;
                // where the transfers into l and l + 1 touch common nodes of level l (interfaces
                // that come within a stencil), both keep the serial order: the transfer into l
                // after the latest transfer into l + 1, and the transfer into l after the latest
                // transfer into l - 1 (which shares nodes of level l - 1 with it)
                if (split && l < nl - 1 && ctx->lv[l].xfer_overlap &&
                    (e = hipStreamWaitEvent(ctx->stream, ctx->ev_xfer[l + 1], 0)) != hipSuccess)
                    break;
                // (a transfer into l - 1 of an earlier call is ordered by the fork and join)
                if (split && l >= 2 && ctx->lv[l - 1].xfer_overlap && xfer_recorded[l - 1] &&
                    (e = hipStreamWaitEvent(ctx->stream, ctx->ev_xfer[l - 1], 0)) != hipSuccess)
                    break;
                if ((s = transfer_checked(ctx, l, kTransferC2F | kTransferF2C | (early_rows ? kTransferRowsReady : 0))))
                    break;
                if (split) {
                    if ((e = hipEventRecord(ctx->ev_xfer[l], ctx->stream)) != hipSuccess) break;
                    after_xfer[l - 1] = 1;
                    xfer_recorded[l] = 1;
                }
            }
            if (early_rows && l < nl - 1) {
                if ((s = launch_transfers(ctx, l + 1, kTransferCoarseRows))) break;
                if ((e = hipEventRecord(ctx->ev_rows[l + 1], ctx->stream)) != hipSuccess) break;
            }
        }
    }
    if (split) {
        // join every level stream back into the caller's stream (also on failure, so that a
        // capture can end)
        ctx->stream = main;
        for (int l = 0; l < nl; ++l) {
            const hipError_t ej = hipEventRecord(ctx->ev_join[l], ctx->lstream[l]);
            const hipError_t ew = ej == hipSuccess ? hipStreamWaitEvent(main, ctx->ev_join[l], 0) : ej;
            if (e == hipSuccess) e = ew;
        }
    }
    if (s) return s;
    if (e != hipSuccess) return ctx->fail(ZLH_EDEVICE, "level streams: %s", hipGetErrorString(e));
    return ZLH_OK;
}

// capture two background steps into an executable graph (the host-side state advances as if
// they ran; the caller launches the graph at once)
zlh_status capture_steps(zlh_ctx *ctx, int nsteps, hipGraphExec_t *exec) {
    ZLH_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    const zlh_status s = march_launches(ctx, nsteps);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
    if (s) {
        if (graph) (void)hipGraphDestroy(graph);
        return s;
    }
    if (e != hipSuccess) return ctx->fail(ZLH_EDEVICE, "graph capture: %s", hipGetErrorString(e));
    const hipError_t ei = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) return ctx->fail(ZLH_EDEVICE, "graph instantiation: %s", hipGetErrorString(ei));
    return ZLH_OK;
}

}  // namespace

extern "C" {

zlh_status zlh_time_marching(zlh_ctx *ctx, int nsteps) {
    if (!ctx) return ZLH_EINVAL;
    if (!ctx->built) return ctx->fail(ZLH_ESTATE, "not built");
    if (nsteps < 0) return ctx->fail(ZLH_EINVAL, "negative step count");
    // launch-bound multi-level steps replay captured graphs: of two background steps while two
    // remain (level 0's buffer index comes back to where it was), of one step for an odd last
    // one.  Timing (per-launch events) runs the launches directly.
    if (!ctx->use_graphs || ctx->timing) return march_launches(ctx, nsteps);
    int t = 0;
    while (t < nsteps) {
        const int n = nsteps - t >= 2 ? 2 : 1;
        // steps per graph, and the buffer index and the aux collision's pending node set of every level
        uint64_t key = n == 1 ? 1ull << 63 : 0;
        for (int l = 0; l <= ctx->d.max_level; ++l)
            key |= (uint64_t)(ctx->lv[l].cur & 1) << l | (uint64_t)(ctx->lv[l].pending & 3) << (16 + 2 * l);
        auto it = ctx->graph_exec.find(key);
        if (it == ctx->graph_exec.end()) {
            hipGraphExec_t g = nullptr;
            zlh_status s = capture_steps(ctx, n, &g);   // advances the host state
            if (s) return s;
            ctx->graph_exec[key] = g;
            ZLH_TRY(hipGraphLaunch(g, ctx->stream));
        } else {
            ZLH_TRY(hipGraphLaunch(it->second, ctx->stream));
            // host state after the steps: post-stream data current, each level's last substep the
            // last of a background step (even on refined levels), and the buffer index flipped
            // once per substep: refined levels take an even number, level 0 one per step
            for (int l = 0; l <= ctx->d.max_level; ++l) {
                LevelState &L = ctx->lv[l];
                L.fresh = false;
                L.last_tsl = 1 << l;
                L.pending = l > 0 ? kPendingTransfer : kPendingPlain;
            }
            ctx->lv[0].cur ^= n & 1;
        }
        t += n;
    }
    return ZLH_OK;
}

zlh_status zlh_graphs_enable(zlh_ctx *ctx, int enable) {
    if (!ctx) return ZLH_EINVAL;
    ctx->use_graphs = enable != 0;
    if (!ctx->use_graphs) ctx->drop_graphs();
    return ZLH_OK;
}

zlh_status zlh_set_boundary_profile(zlh_ctx *ctx, int dir, int profile) {
    if (!ctx) return ZLH_EINVAL;
    if (ctx->built) return ctx->fail(ZLH_ESTATE, "boundary profile after build");
    if (dir < 1 || dir > 2 * ctx->d.dims) return ctx->fail(ZLH_EINVAL, "bad face %d", dir);
    if (profile != ZLH_PROFILE_UNIFORM && profile != ZLH_PROFILE_PARABOLIC) return ctx->fail(ZLH_EINVAL, "bad profile %d", profile);
    if (profile == ZLH_PROFILE_PARABOLIC && ctx->d.bc_scheme[dir] != ZLH_BC_NEE)
        return ctx->fail(ZLH_EINVAL, "face %d: profiles apply to non-equilibrium extrapolation faces", dir);
    ctx->bc_profile[dir] = profile;
    return ZLH_OK;
}

zlh_status zlh_set_boundary_velocities(zlh_ctx *ctx, int level, int dir, const double *u, int64_t n) {
    if (!ctx) return ZLH_EINVAL;
    if (ctx->built) return ctx->fail(ZLH_ESTATE, "boundary velocities after build");
    if (level < 0 || level > ctx->d.max_level) return ctx->fail(ZLH_EINVAL, "bad level %d", level);
    if (dir < 1 || dir > 2 * ctx->d.dims) return ctx->fail(ZLH_EINVAL, "bad face %d", dir);
    if (ctx->d.bc_scheme[dir] != ZLH_BC_NEE)
        return ctx->fail(ZLH_EINVAL, "face %d: node velocities apply to non-equilibrium extrapolation faces", dir);
    if (n < 0 || (n > 0 && !u)) return ctx->fail(ZLH_EINVAL, "bad boundary velocities");
    ctx->bc_node_u[{level, dir}] = std::vector<double>(u, u + 3 * n);
    return ZLH_OK;
}

zlh_status zlh_set_interpolation_method(zlh_ctx *ctx, int level, int method) {
    if (!ctx) return ZLH_EINVAL;
    if (ctx->built) return ctx->fail(ZLH_ESTATE, "interpolation method after build");
    if (level < 1 || level > ctx->d.max_level) return ctx->fail(ZLH_EINVAL, "bad fine level %d", level);
    if (method != ZLH_INTERP_LAGRANGIAN && method != ZLH_INTERP_LINEAR)
        return ctx->fail(ZLH_EINVAL, "bad interpolation method %d", method);
    ctx->interp_method[level] = method;
    return ZLH_OK;
}

zlh_status zlh_set_time_interpolation(zlh_ctx *ctx, int enable) {
    if (!ctx) return ZLH_EINVAL;
    if (ctx->time_interp == (enable != 0)) return ZLH_OK;
    ctx->time_interp = enable != 0;
    ctx->drop_graphs();   // captured graphs hold the other transfer schedule
    if (ctx->built) {
        // the synchronised rows from the state as it is now (after a whole background step, the
        // rows the last transfers computed)
        zlh_status s = refresh_transfer_rows(ctx, -1);
        if (s) return s;
        ZLH_TRY(hipStreamSynchronize(ctx->stream));
    }
    return ZLH_OK;
}

zlh_status zlh_set_moment_transform(zlh_ctx *ctx, int where) {
    if (!ctx) return ZLH_EINVAL;
    if (where != ZLH_TRANSFORM_VALU && where != ZLH_TRANSFORM_MFMA)
        return ctx->fail(ZLH_EINVAL, "bad moment transform %d", where);
    ctx->mrt_transform = where;
    ctx->drop_graphs();   // captured graphs hold the other kernels
    return ZLH_OK;
}

zlh_status zlh_level_streams_enable(zlh_ctx *ctx, int enable) {
    if (!ctx) return ZLH_EINVAL;
    ctx->level_streams = enable != 0;
    ctx->drop_graphs();   // captured graphs hold the other ordering
    return ZLH_OK;
}

zlh_status zlh_download_populations(zlh_ctx *ctx, int level, double *f) {
    if (!ctx) return ZLH_EINVAL;
    if (!ctx->built) return ctx->fail(ZLH_ESTATE, "not built");
    if (level < 0 || level > ctx->d.max_level || !f) return ctx->fail(ZLH_EINVAL, "bad download arguments");
    LevelState &L = ctx->lv[level];
    const int Q = ctx->q;
    DevBuf tmp;
    zlh_status s = dalloc(ctx, tmp, sizeof(double) * L.n * Q, false);
    if (s) return s;
    if ((s = do_materialize(ctx, level, tmp.as<double>()))) return s;
    std::vector<double> h((size_t)L.n * Q);
    ZLH_TRY(hipMemcpyAsync(h.data(), tmp.p, sizeof(double) * L.n * Q, hipMemcpyDeviceToHost, ctx->stream));
    ZLH_TRY(hipStreamSynchronize(ctx->stream));
    for (int64_t i = 0; i < L.n; ++i)
        for (int q = 0; q < Q; ++q) f[i * Q + q] = h[(size_t)q * L.n + i];
    return ZLH_OK;
}

zlh_status zlh_download_macros(zlh_ctx *ctx, int level, double *rho, double *u) {
    if (!ctx) return ZLH_EINVAL;
    if (level < 0 || level > ctx->d.max_level || !rho || !u) return ctx->fail(ZLH_EINVAL, "bad download arguments");
    LevelState &L = ctx->lv[level];
    const int Q = ctx->q, dims = ctx->d.dims;
    std::vector<double> f((size_t)L.n * Q);
    zlh_status s = zlh_download_populations(ctx, level, f.data());
    if (s) return s;
    for (int64_t i = 0; i < L.n; ++i) {
        double r = f[i * Q], j[3] = {0, 0, 0};
        for (int q = 1; q < Q; ++q) {
            const double v = f[i * Q + q];
            r += v;
            j[0] += (dims == 2 ? Lattice<2>::cx[q] : Lattice<3>::cx[q]) * v;
            j[1] += (dims == 2 ? Lattice<2>::cy[q] : Lattice<3>::cy[q]) * v;
            if (dims == 3) j[2] += Lattice<3>::cz[q] * v;
        }
        rho[i] = r;
        for (int a = 0; a < dims; ++a) u[i * dims + a] = ctx->d.compressible ? j[a] / r : j[a];
    }
    return ZLH_OK;
}

zlh_status zlh_download_source_table(zlh_ctx *ctx, int level, int32_t *src) {
    if (!ctx) return ZLH_EINVAL;
    if (!ctx->built) return ctx->fail(ZLH_ESTATE, "not built");
    if (level < 0 || level > ctx->d.max_level || !src) return ctx->fail(ZLH_EINVAL, "bad arguments");
    LevelState &L = ctx->lv[level];
    DevBuf tmp;
    zlh_status s = dalloc(ctx, tmp, sizeof(int32_t) * L.n * ctx->q, false);
    if (s) return s;
    if (ctx->d.dims == 2)
        hipLaunchKernelGGL(k_source_table<2>, dim3(blocks(L.n)), dim3(256), 0, ctx->stream, L.node_slot_d.as<int32_t>(),
                           L.n, L.occ_d.as<unsigned long long>(), L.nbr_d.as<int32_t>(), L.slot_node_d.as<int32_t>(),
                           tmp.as<int32_t>());
    else
        hipLaunchKernelGGL(k_source_table<3>, dim3(blocks(L.n)), dim3(256), 0, ctx->stream, L.node_slot_d.as<int32_t>(),
                           L.n, L.occ_d.as<unsigned long long>(), L.nbr_d.as<int32_t>(), L.slot_node_d.as<int32_t>(),
                           tmp.as<int32_t>());
    ZLH_TRY(hipGetLastError());
    ZLH_TRY(hipMemcpyAsync(src, tmp.p, sizeof(int32_t) * L.n * ctx->q, hipMemcpyDeviceToHost, ctx->stream));
    ZLH_TRY(hipStreamSynchronize(ctx->stream));
    return ZLH_OK;
}

zlh_status zlh_download_node_forces(zlh_ctx *ctx, int level, double *force) {
    if (!ctx) return ZLH_EINVAL;
    if (!ctx->built) return ctx->fail(ZLH_ESTATE, "not built");
    if (level < 0 || level > ctx->d.max_level || !force) return ctx->fail(ZLH_EINVAL, "bad arguments");
    LevelState &L = ctx->lv[level];
    std::vector<double> aux((size_t)3 * std::max<int64_t>(L.naux, 1));
    ZLH_TRY(hipMemcpyAsync(aux.data(), L.aux_force_d.p, sizeof(double) * aux.size(), hipMemcpyDeviceToHost,
                           ctx->stream));
    ZLH_TRY(hipStreamSynchronize(ctx->stream));
    for (int64_t i = 0; i < L.n; ++i) {
        const int32_t ai = L.node_aux_h[i];
        for (int a = 0; a < 3; ++a) force[i * 3 + a] = ai >= 0 ? aux[(size_t)ai * 3 + a] : 0.;
    }
    return ZLH_OK;
}

zlh_status zlh_ib_download_forces(zlh_ctx *ctx, int level, double *force) {
    if (!ctx) return ZLH_EINVAL;
    if (level < 0 || level > ctx->d.max_level || !force) return ctx->fail(ZLH_EINVAL, "bad arguments");
    LevelState &L = ctx->lv[level];
    if (L.ib.n == 0) return ZLH_OK;
    ZLH_TRY(hipMemcpyAsync(force, L.ibt.marker_force.p, sizeof(double) * 3 * L.ib.n, hipMemcpyDeviceToHost, ctx->stream));
    ZLH_TRY(hipStreamSynchronize(ctx->stream));
    return ZLH_OK;
}

zlh_status zlh_conservation(zlh_ctx *ctx, int level, double *out4) {
    if (!ctx) return ZLH_EINVAL;
    if (level < 0 || level > ctx->d.max_level || !out4) return ctx->fail(ZLH_EINVAL, "bad arguments");
    LevelState &L = ctx->lv[level];
    const int Q = ctx->q;
    zlh_status s;
    if (!ctx->lattice_d.p) {
        std::vector<int> c(3 * 19, 0);
        for (int q = 0; q < Q; ++q) {
            c[q] = ctx->d.dims == 2 ? Lattice<2>::cx[q] : Lattice<3>::cx[q];
            c[19 + q] = ctx->d.dims == 2 ? Lattice<2>::cy[q] : Lattice<3>::cy[q];
            c[38 + q] = ctx->d.dims == 2 ? 0 : Lattice<3>::cz[q];
        }
        if ((s = upload(ctx, ctx->lattice_d, c))) return s;
    }
    DevBuf tmp, acc;
    if ((s = dalloc(ctx, tmp, sizeof(double) * L.n * Q, false))) return s;
    if ((s = dalloc(ctx, acc, sizeof(double) * 4))) return s;
    if ((s = do_materialize(ctx, level, tmp.as<double>()))) return s;
    const int *c = ctx->lattice_d.as<int>();
    hipLaunchKernelGGL(k_sum4, dim3(blocks(L.n)), dim3(256), 0, ctx->stream, tmp.as<double>(), L.n, Q, c, c + 19,
                       c + 38, acc.as<double>());
    ZLH_TRY(hipGetLastError());
    ZLH_TRY(hipMemcpyAsync(out4, acc.p, sizeof(double) * 4, hipMemcpyDeviceToHost, ctx->stream));
    ZLH_TRY(hipStreamSynchronize(ctx->stream));
    return ZLH_OK;
}

zlh_status zlh_level_info(zlh_ctx *ctx, int level, int64_t *n_nodes, int64_t *n_aux, int64_t *n_bricks) {
    if (!ctx) return ZLH_EINVAL;
    if (level < 0 || level > ctx->d.max_level) return ctx->fail(ZLH_EINVAL, "bad level");
    const LevelState &L = ctx->lv[level];
    if (n_nodes) *n_nodes = (int64_t)L.keys.size();
    if (n_aux) *n_aux = L.naux;
    if (n_bricks) *n_bricks = L.nbrick;
    return ZLH_OK;
}

zlh_status zlh_transfer_info(zlh_ctx *ctx, int level, int64_t *out) {
    if (!ctx || !out) return ZLH_EINVAL;
    if (level < 0 || level > ctx->d.max_level) return ctx->fail(ZLH_EINVAL, "bad level");
    const TransferTables &X = ctx->lv[level].xfer;
    const int64_t v[ZLH_XFER_INFO_N] = {X.n_c2f, X.n_src, X.n_f2c, X.n_groups, X.staged_rows, X.n_contrib,
                                        X.n_fallback, (int64_t)X.lds_bytes, ctx->lv[level].xfer_overlap ? 1 : 0};
    for (int i = 0; i < ZLH_XFER_INFO_N; ++i) out[i] = level > 0 ? v[i] : 0;
    return ZLH_OK;
}

zlh_status zlh_synchronize(zlh_ctx *ctx) {
    if (!ctx) return ZLH_EINVAL;
    ZLH_TRY(hipStreamSynchronize(ctx->stream));
    return ZLH_OK;
}

zlh_status zlh_timing_enable(zlh_ctx *ctx, int enable) {
    if (!ctx) return ZLH_EINVAL;
    ZLH_TRY(hipStreamSynchronize(ctx->stream));
    for (auto &e : ctx->events) { (void)hipEventDestroy(e.start); (void)hipEventDestroy(e.stop); }
    ctx->events.clear();
    for (int l = 0; l < 16; ++l) { ctx->timing_ms[l] = 0.; ctx->timing_launches[l] = 0; }
    ctx->timing = enable != 0;
    return ZLH_OK;
}

zlh_status zlh_timing_read(zlh_ctx *ctx, int level, double *total_ms, int64_t *launches) {
    if (!ctx) return ZLH_EINVAL;
    if (level < -1 || level > ctx->d.max_level) return ctx->fail(ZLH_EINVAL, "bad level %d", level);
    ZLH_TRY(hipStreamSynchronize(ctx->stream));
    for (auto &e : ctx->events) {
        float ms = 0.f;
        ZLH_TRY(hipEventElapsedTime(&ms, e.start, e.stop));
        ctx->timing_ms[e.level] += ms;
        ++ctx->timing_launches[e.level];
        (void)hipEventDestroy(e.start);
        (void)hipEventDestroy(e.stop);
    }
    ctx->events.clear();
    double ms = 0.;
    int64_t n = 0;
    for (int l = 0; l <= ctx->d.max_level; ++l)
        if (level < 0 || l == level) { ms += ctx->timing_ms[l]; n += ctx->timing_launches[l]; }
    if (total_ms) *total_ms = ms;
    if (launches) *launches = n;
    return ZLH_OK;
}

zlh_status zlh_mlp_set(zlh_ctx *ctx, const float *params, const float *mean, const float *std_dev, const float *fmin,
                       const float *fmax, double nu_scale) {
    if (!ctx) return ZLH_EINVAL;
    if (!params || !mean || !std_dev || !fmin || !fmax || !(nu_scale > 0.)) return ctx->fail(ZLH_EINVAL, "bad closure arguments");
    MlpDev m;
    std::memcpy(m.p, params, sizeof(m.p));
    for (int i = 0; i < kMlpF; ++i) {
        if (!(std_dev[i] > 0.f)) return ctx->fail(ZLH_EINVAL, "feature %d has zero spread", i);
        m.mean[i] = mean[i];
        m.inv_std[i] = 1.f / std_dev[i];
        m.fmin[i] = fmin[i];
        m.fmax[i] = fmax[i];
    }
    m.nu_scale = (float)nu_scale;
    // operand images of the wavefront MFMA network (layout in mlp.hpp).  W1, b1 and b2 are scaled
    // by c = -log2(e) (in FP64, stored as FP32, then rounded to the operand format), so that both
    // hidden layers' pre-activations arrive as c z; the output weights carry 1/c (silu_scaled)
    const float *W1 = m.p, *b1 = W1 + kMlpH * kMlpF, *W2 = b1 + kMlpH, *b2 = W2 + kMlpH * kMlpH, *w3 = b2 + kMlpH;
    auto scaled = [](float v) { return (float)(kSiluScale * (double)v); };
    MlpOps &o = m.ops;
    for (int l = 0; l < 64; ++l) {
        const int i = l % 32, h = l / 32;
        for (int e = 0; e < 8; ++e) {
            const int k = 8 * h + e;
            const float a1 = scaled(k < kMlpF ? W1[i * kMlpF + k] : (k == kMlpF ? b1[i] : 0.f));
            o.w1a[l][e] = bf16_bits(a1);
            o.h1a[l][e] = f16_bits(a1);
            for (int st = 0; st < 2; ++st) {
                const float a2 = W2[i * kMlpH + mfma_row(h, 8 * st + e)];
                o.w2a[st][l][e] = bf16_bits(a2);
                o.h2a[st][l][e] = f16_bits(a2);
            }
        }
    }
    for (int h = 0; h < 2; ++h)
        for (int r = 0; r < 16; ++r) {
            o.w3r[h][r] = (float)((double)w3[mfma_row(h, r)] / kSiluScale);
            o.b2r[h][r] = scaled(b2[mfma_row(h, r)]);
        }
    // operand images of the moment transform on the matrix cores (M and M^T, integers)
    {
        auto fill = [&o](auto S) {
            constexpr int Q = decltype(S)::Q;
            for (int s = 0; s < kMrtSteps; ++s)
                for (int l = 0; l < 64; ++l) {
                    const int i = l % 32, k = 2 * s + l / 32;
                    o.mfwd[s][l] = (i < Q && k < Q) ? (float)S.m[i][k] : 0.f;
                    o.mbwd[s][l] = (i < Q && k < Q) ? (float)S.m[k][i] : 0.f;
                }
        };
        if (ctx->d.dims == 2) fill(MomentSpace<2>{});
        else fill(MomentSpace<3>{});
    }
    for (int i = 0; i < kMlpF; ++i) {
        o.mean[i] = m.mean[i];
        o.inv_std[i] = m.inv_std[i];
        o.fmin[i] = m.fmin[i];
        o.fmax[i] = m.fmax[i];
    }
    o.nu_scale = m.nu_scale;
    o.b3 = m.p[kMlpParams - 1];
    o.pad[0] = o.pad[1] = 0.f;
    ZLH_TRY(hipMemcpyAsync(ctx->mlp_d.p, &m, sizeof(m), hipMemcpyHostToDevice, ctx->stream));
    ZLH_TRY(hipStreamSynchronize(ctx->stream));
    ctx->mlp_set = true;
    return ZLH_OK;
}

// gradient kernel of the selected training form: one thread per sample on the vector units, or
// 64 samples per wavefront on the matrix cores (a grid of at most kGradMaxBlocks workgroups that
// loop over the batch)
static void launch_mlp_grad(zlh_ctx *ctx, const float *p, const float *x, const float *t, int64_t n, int nb,
                            uint32_t seed, int random, float *grad, float *loss) {
    if (ctx->mlp_train_form == ZLH_MLP_BF16_MFMA)
        hipLaunchKernelGGL(k_mlp_grad_mfma, dim3(std::min<unsigned>(blocks(nb), kGradMaxBlocks)), dim3(256), 0,
                           ctx->stream, p, x, t, n, nb, seed, random, grad, loss);
    else
        hipLaunchKernelGGL(k_mlp_grad, dim3(blocks(nb)), dim3(256), 0, ctx->stream, p, x, t, n, nb, seed, random, grad,
                           loss);
}

zlh_status zlh_mlp_training_form(zlh_ctx *ctx, int form) {
    if (!ctx) return ZLH_EINVAL;
    if (form != ZLH_MLP_FP32 && form != ZLH_MLP_BF16_MFMA)
        return ctx->fail(ZLH_EINVAL, "training form must be ZLH_MLP_FP32 or ZLH_MLP_BF16_MFMA (FP16 gradients underflow)");
    ctx->mlp_train_form = form;
    return ZLH_OK;
}

zlh_status zlh_mlp_gradient(zlh_ctx *ctx, const float *params, const float *x, const float *t, int64_t n, float *grad,
                            float *loss) {
    if (!ctx) return ZLH_EINVAL;
    if (!params || !x || !t || !grad || n < 1 || n > (1 << 26)) return ctx->fail(ZLH_EINVAL, "bad gradient arguments");
    DevBuf dp, dx, dt, dg, dl;
    zlh_status s;
    if ((s = dalloc(ctx, dp, sizeof(float) * kMlpParams, false))) return s;
    if ((s = dalloc(ctx, dx, sizeof(float) * kMlpF * n, false))) return s;
    if ((s = dalloc(ctx, dt, sizeof(float) * n, false))) return s;
    if ((s = dalloc(ctx, dg, sizeof(float) * kMlpParams))) return s;
    if ((s = dalloc(ctx, dl, sizeof(float)))) return s;
    ZLH_TRY(hipMemcpyAsync(dp.p, params, sizeof(float) * kMlpParams, hipMemcpyHostToDevice, ctx->stream));
    ZLH_TRY(hipMemcpyAsync(dx.p, x, sizeof(float) * kMlpF * n, hipMemcpyHostToDevice, ctx->stream));
    ZLH_TRY(hipMemcpyAsync(dt.p, t, sizeof(float) * n, hipMemcpyHostToDevice, ctx->stream));
    launch_mlp_grad(ctx, dp.as<float>(), dx.as<float>(), dt.as<float>(), n, (int)n, 0u, 0, dg.as<float>(),
                    dl.as<float>());
    ZLH_TRY(hipGetLastError());
    ZLH_TRY(hipMemcpyAsync(grad, dg.p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, ctx->stream));
    if (loss) ZLH_TRY(hipMemcpyAsync(loss, dl.p, sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    ZLH_TRY(hipStreamSynchronize(ctx->stream));
    return ZLH_OK;
}

zlh_status zlh_mlp_gradient_noalloc(zlh_ctx *ctx, const float *params, const float *x, const float *t, int64_t n,
                                    float *grad, float *loss) {
    if (!ctx) return ZLH_EINVAL;
    if (!params || !x || !t || !grad || n < 1) return ctx->fail(ZLH_EINVAL, "bad gradient arguments");
    if (n > kNoAllocMaxSamples)
        return ctx->fail(ZLH_EINVAL, "zlh_mlp_gradient_noalloc holds at most %lld samples, got %lld",
                         (long long)kNoAllocMaxSamples, (long long)n);
    if (!ctx->scratch_addr) return ctx->fail(ZLH_ESTATE, "no-allocation array not set up");
    float *dp = ctx->scratch_addr + kScratchParams, *dg = ctx->scratch_addr + kScratchGrad;
    float *dl = ctx->scratch_addr + kScratchLoss, *dt = ctx->scratch_addr + kScratchT, *dx = ctx->scratch_addr + kScratchX;
    // gradient and loss regions (and the gap between them) start at zero
    const int64_t nclear = kScratchT - kScratchGrad;
    hipLaunchKernelGGL(k_scratch_clear, dim3(blocks(nclear)), dim3(256), 0, ctx->stream, dg, nclear);
    ZLH_TRY(hipGetLastError());
    ZLH_TRY(hipMemcpyAsync(dp, params, sizeof(float) * kMlpParams, hipMemcpyHostToDevice, ctx->stream));
    ZLH_TRY(hipMemcpyAsync(dx, x, sizeof(float) * kMlpF * n, hipMemcpyHostToDevice, ctx->stream));
    ZLH_TRY(hipMemcpyAsync(dt, t, sizeof(float) * n, hipMemcpyHostToDevice, ctx->stream));
    launch_mlp_grad(ctx, dp, dx, dt, n, (int)n, 0u, 0, dg, dl);
    ZLH_TRY(hipGetLastError());
    ZLH_TRY(hipMemcpyAsync(grad, dg, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, ctx->stream));
    if (loss) ZLH_TRY(hipMemcpyAsync(loss, dl, sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    ZLH_TRY(hipStreamSynchronize(ctx->stream));
    return ZLH_OK;
}

zlh_status zlh_mlp_train(zlh_ctx *ctx, const float *x, const float *t, int64_t n, int steps, int batch, double lr,
                         uint64_t seed, float *params, float *loss_history) {
    if (!ctx) return ZLH_EINVAL;
    if (!params || !x || !t || n < 1 || n > (1 << 26) || steps < 1 || batch < 1 || batch > (1 << 20) || !(lr > 0.))
        return ctx->fail(ZLH_EINVAL, "bad training arguments");
    DevBuf dp, dx, dt, dg, dm, dv, dl;
    zlh_status s;
    if ((s = dalloc(ctx, dp, sizeof(float) * kMlpParams, false))) return s;
    if ((s = dalloc(ctx, dx, sizeof(float) * kMlpF * n, false))) return s;
    if ((s = dalloc(ctx, dt, sizeof(float) * n, false))) return s;
    if ((s = dalloc(ctx, dg, sizeof(float) * kMlpParams))) return s;
    if ((s = dalloc(ctx, dm, sizeof(float) * kMlpParams))) return s;
    if ((s = dalloc(ctx, dv, sizeof(float) * kMlpParams))) return s;
    if ((s = dalloc(ctx, dl, sizeof(float) * steps))) return s;
    ZLH_TRY(hipMemcpyAsync(dp.p, params, sizeof(float) * kMlpParams, hipMemcpyHostToDevice, ctx->stream));
    ZLH_TRY(hipMemcpyAsync(dx.p, x, sizeof(float) * kMlpF * n, hipMemcpyHostToDevice, ctx->stream));
    ZLH_TRY(hipMemcpyAsync(dt.p, t, sizeof(float) * n, hipMemcpyHostToDevice, ctx->stream));
    // a batch of at least n samples is the full set in order (plain full-batch Adam); smaller
    // batches are drawn with replacement from a per-step hash
    const int full = (int64_t)batch >= n;
    const int nb = full ? (int)n : batch;
    for (int it = 1; it <= steps; ++it) {
        ZLH_TRY(hipMemsetAsync(dg.p, 0, sizeof(float) * kMlpParams, ctx->stream));
        launch_mlp_grad(ctx, dp.as<float>(), dx.as<float>(), dt.as<float>(), n, nb,
                        (uint32_t)(seed * 1000003ull + (uint64_t)it), full ? 0 : 1, dg.as<float>(),
                        dl.as<float>() + (it - 1));
        hipLaunchKernelGGL(k_adam, dim3(blocks(kMlpParams)), dim3(256), 0, ctx->stream, dp.as<float>(), dg.as<float>(),
                           dm.as<float>(), dv.as<float>(), kMlpParams, (float)lr, 0.9f, 0.999f, 1e-8f, it);
    }
    ZLH_TRY(hipGetLastError());
    ZLH_TRY(hipMemcpyAsync(params, dp.p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, ctx->stream));
    if (loss_history) ZLH_TRY(hipMemcpyAsync(loss_history, dl.p, sizeof(float) * steps, hipMemcpyDeviceToHost, ctx->stream));
    ZLH_TRY(hipStreamSynchronize(ctx->stream));
    return ZLH_OK;
}

zlh_status zlh_mlp_eval(zlh_ctx *ctx, const float *x, int64_t n, int form, float *y) {
    if (!ctx) return ZLH_EINVAL;
    if (!ctx->mlp_set) return ctx->fail(ZLH_ESTATE, "learned closure not set");
    if (!x || !y || n < 1 || form < ZLH_MLP_FP32 || form > ZLH_MLP_FP16_MFMA)
        return ctx->fail(ZLH_EINVAL, "bad evaluation arguments");
    DevBuf dx, dy;
    zlh_status s;
    if ((s = dalloc(ctx, dx, sizeof(float) * kMlpF * n, false))) return s;
    if ((s = dalloc(ctx, dy, sizeof(float) * n, false))) return s;
    ZLH_TRY(hipMemcpyAsync(dx.p, x, sizeof(float) * kMlpF * n, hipMemcpyHostToDevice, ctx->stream));
    if (form == ZLH_MLP_BF16_MFMA)
        hipLaunchKernelGGL(k_mlp_eval_wave<OperandsBf16>, dim3(blocks(n)), dim3(256), 0, ctx->stream,
                           ctx->mlp_d.as<MlpDev>(), dx.as<float>(), n, dy.as<float>());
    else if (form == ZLH_MLP_FP16_MFMA)
        hipLaunchKernelGGL(k_mlp_eval_wave<OperandsF16>, dim3(blocks(n)), dim3(256), 0, ctx->stream,
                           ctx->mlp_d.as<MlpDev>(), dx.as<float>(), n, dy.as<float>());
    else
        hipLaunchKernelGGL(k_mlp_eval, dim3(blocks(n)), dim3(256), 0, ctx->stream, ctx->mlp_d.as<MlpDev>(),
                           dx.as<float>(), n, dy.as<float>());
    ZLH_TRY(hipGetLastError());
    ZLH_TRY(hipMemcpyAsync(y, dy.p, sizeof(float) * n, hipMemcpyDeviceToHost, ctx->stream));
    ZLH_TRY(hipStreamSynchronize(ctx->stream));
    return ZLH_OK;
}

zlh_status zlh_closure_features(zlh_ctx *ctx, int level, float *features, float *nu_t, int32_t *fallback) {
    if (!ctx) return ZLH_EINVAL;
    if (!ctx->built) return ctx->fail(ZLH_ESTATE, "not built");
    if (!ctx->mlp_set) return ctx->fail(ZLH_ESTATE, "learned closure not set");
    if (level < 0 || level > ctx->d.max_level || !features || !nu_t || !fallback) return ctx->fail(ZLH_EINVAL, "bad arguments");
    LevelState &L = ctx->lv[level];
    DevBuf fq, dx, dn, df;
    zlh_status s;
    if ((s = dalloc(ctx, fq, sizeof(double) * L.n * ctx->q, false))) return s;
    if ((s = dalloc(ctx, dx, sizeof(float) * kMlpF * L.n, false))) return s;
    if ((s = dalloc(ctx, dn, sizeof(float) * L.n, false))) return s;
    if ((s = dalloc(ctx, df, sizeof(int32_t) * L.n, false))) return s;
    if ((s = do_materialize(ctx, level, fq.as<double>()))) return s;
    if (ctx->d.dims == 2)
        hipLaunchKernelGGL(k_closure_features<2>, dim3(blocks(L.n)), dim3(256), 0, ctx->stream, fq.as<double>(), L.n,
                           L.sp.tau0, L.sp.cs18, ctx->mlp_d.as<MlpDev>(), dx.as<float>(), dn.as<float>(), df.as<int32_t>());
    else
        hipLaunchKernelGGL(k_closure_features<3>, dim3(blocks(L.n)), dim3(256), 0, ctx->stream, fq.as<double>(), L.n,
                           L.sp.tau0, L.sp.cs18, ctx->mlp_d.as<MlpDev>(), dx.as<float>(), dn.as<float>(), df.as<int32_t>());
    ZLH_TRY(hipGetLastError());
    ZLH_TRY(hipMemcpyAsync(features, dx.p, sizeof(float) * kMlpF * L.n, hipMemcpyDeviceToHost, ctx->stream));
    ZLH_TRY(hipMemcpyAsync(nu_t, dn.p, sizeof(float) * L.n, hipMemcpyDeviceToHost, ctx->stream));
    ZLH_TRY(hipMemcpyAsync(fallback, df.p, sizeof(int32_t) * L.n, hipMemcpyDeviceToHost, ctx->stream));
    ZLH_TRY(hipStreamSynchronize(ctx->stream));
    return ZLH_OK;
}

}  // extern "C"
