// kernels.hpp -- gfx950 kernels of the lattice Boltzmann time step.
//
// Data layout of a level in HBM:
//   * nodes are grouped into bricks of 64 (4x4x4 in 3D, 8x8 in 2D); the brick key is the Morton
//     code shifted right by 6, the lane inside the brick is the low 6 Morton bits, so one
//     wavefront owns one brick and a node's slot is brick * 64 + lane;
//   * populations are structure-of-arrays per direction: f[q][slot], two buffers (A/B);
//   * per brick: an occupancy mask, the 27 (9 in 2D) neighbouring brick indices, and a mask of
//     the nodes that take the deferred ("aux") path with the prefix count of aux nodes before it.
//
// Time-step structure (the reference's collide -> stream -> boundaries -> transfer per level
// substep, lbm_grid_interface.cpp:104-153):
//   * "bulk" nodes have every pull source, are not touched by boundaries, transfers or the
//     immersed boundary: the fused kernel pulls their post-collision neighbours (stream) and
//     immediately applies the next collision, writing post-collision values: 2Q accesses per
//     node, no per-node tables.
//   * "aux" nodes (domain boundary nodes and their partners, refinement-interface layers and
//     interpolation stencils, immersed-boundary supports, nodes with an absent pull source) keep
//     their post-stream populations in a compact side array, are modified there by the boundary,
//     transfer and forcing kernels, and are collided at the start of their level's next substep.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include "lattice.h"
#include "mlp.hpp"

namespace zlh {

struct StepParams {
    double tau0;           // level relaxation time (CalRelaxationTime)
    double dt;             // level time step 1/2^l
    double cs18;           // 18 * Cs^2 (k0Smagorinsky)
    double const_force[3];
    int les;               // 0 none, 1 Smagorinsky, 2 learned closure
    int les_tau_mode;      // 0 reference, 1 physical
    int has_force;
    int compressible;
    const MlpDev *mlp;     // learned closure (les == 2)
};

template <typename T>
struct BrickLevel {
    int64_t nbrick;
    int64_t nslot;            // nbrick * 64
    int64_t naux;
    const uint64_t *occ;      // [nbrick]
    const int32_t *nbr;       // [nbrick][NN]
    const uint64_t *auxmask;  // [nbrick]
    const int32_t *auxbase;   // [nbrick]
    const uint32_t *aux_missing;  // [naux]
    const int32_t *aux_slot;  // [naux]
};

template <int D>
struct MrtTable {
    static constexpr int Q = Lattice<D>::Q;
    int m[Q][Q];
    int n2[Q];
    constexpr MrtTable() : m(), n2() {
        for (int k = 0; k < Q; ++k) {
            int s = 0;
            for (int i = 0; i < Q; ++i) {
                m[k][i] = Lattice<D>::moment(k, i);
                s += m[k][i] * m[k][i];
            }
            n2[k] = s;
        }
    }
};

// shift applied to stored populations: FP32 stores f - w_i
template <int D, bool SHIFT, typename C>
__device__ __forceinline__ C wshift(int q) {
    return SHIFT ? C(Lattice<D>::w[q]) : C(0);
}

// equilibrium in storage form (feq - w_i when shifted), CalFeq{2D,3D}{Compressible,Incompressible}
// (lbm_solver_interface.cpp:374-435)
template <int D, bool SHIFT, typename C>
__device__ __forceinline__ C feq_stored(int q, C rho, C drho, C ux, C uy, C uz, C usq, int compressible) {
    using L = Lattice<D>;
    C cu = ux * C(L::cx[q]) + uy * C(L::cy[q]);
    if (D == 3) cu += uz * C(L::cz[q]);
    C poly = C(3) * cu + C(4.5) * cu * cu - C(1.5) * usq;
    C w = C(L::w[q]);
    if (SHIFT) {
        return compressible ? w * (drho + rho * poly) : w * (drho + poly);
    } else {
        return compressible ? w * rho * (C(1) + poly) : w * (rho + poly);
    }
}

// macroscopic variables of stored populations; F is the total force (zero when unforced)
// (CalMacro*/CalMacroForce*, lbm_solver_interface.cpp:477-592)
template <int D, bool SHIFT, typename C>
__device__ __forceinline__ void macro(const C (&f)[Lattice<D>::Q], const C *F, C dt, int has_force,
                                      int compressible, C &rho, C &drho, C (&u)[3]) {
    using L = Lattice<D>;
    constexpr int Q = L::Q;
    C s = f[0], jx = C(0), jy = C(0), jz = C(0);
#pragma unroll
    for (int q = 1; q < Q; ++q) {
        s += f[q];
        if (L::cx[q]) jx += C(L::cx[q]) * f[q];
        if (L::cy[q]) jy += C(L::cy[q]) * f[q];
        if (D == 3 && L::cz[q]) jz += C(L::cz[q]) * f[q];
    }
    if (SHIFT) { drho = s; rho = C(1) + s; } else { rho = s; drho = s - C(1); }
    if (has_force) {
        jx += C(0.5) * F[0] * dt;
        jy += C(0.5) * F[1] * dt;
        if (D == 3) jz += C(0.5) * F[2] * dt;
    }
    if (compressible) {
        C inv = C(1) / rho;
        u[0] = jx * inv; u[1] = jy * inv; u[2] = D == 3 ? jz * inv : C(0);
    } else {
        u[0] = jx; u[1] = jy; u[2] = D == 3 ? jz : C(0);
    }
}

// One node collision on stored populations (in place): func_collision_node_
// (lbm_solver_interface.cpp:266-300), Smagorinsky SGS time (les_models/lbm_les_smagorinsky.cpp:22-94),
// SRT (lbm_collision_model_srt.cpp:20-39) and MRT (lbm_collision_model_mrt.cpp:114-143).
// The MRT transform is applied as M^T diag(1/|m_k|^2) (S M neq - D M F dt) with the constant
// integer basis, which equals M^{-1} S M (f - feq) - M^{-1} D M F dt.
template <int D, bool SHIFT, int COLL, typename C>
__device__ __forceinline__ void collide(C (&f)[Lattice<D>::Q], const C *F, const StepParams &p,
                                        C &rho_out, C (&u)[3]) {
    using L = Lattice<D>;
    constexpr int Q = L::Q;
    const C dt = C(p.dt);
    C rho, drho;
    macro<D, SHIFT, C>(f, F, dt, p.has_force, p.compressible, rho, drho, u);
    rho_out = rho;
    C usq = u[0] * u[0] + u[1] * u[1] + (D == 3 ? u[2] * u[2] : C(0));
    C neq[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q)
        neq[q] = f[q] - feq_stored<D, SHIFT, C>(q, rho, drho, u[0], u[1], u[2], usq, p.compressible);
    C fi[Q];
    if (p.has_force) {
        C uF = u[0] * F[0] + u[1] * F[1] + (D == 3 ? u[2] * F[2] : C(0));
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            C cF = C(L::cx[q]) * F[0] + C(L::cy[q]) * F[1] + (D == 3 ? C(L::cz[q]) * F[2] : C(0));
            C cu = C(L::cx[q]) * u[0] + C(L::cy[q]) * u[1] + (D == 3 ? C(L::cz[q]) * u[2] : C(0));
            fi[q] = C(L::w[q]) * (C(3) * (cF - uF) + C(9) * cu * cF);
        }
    }
    C tau = C(p.tau0);
    if (p.les) {
        C pxx = 0, pxy = 0, pyy = 0, pzz = 0, pxz = 0, pyz = 0;
#pragma unroll
        for (int q = 1; q < Q; ++q) {
            const int ax = L::cx[q], ay = L::cy[q], az = L::cz[q];
            if (ax * ax) pxx += neq[q];
            if (ay * ay) pyy += neq[q];
            if (ax * ay) pxy += C(ax * ay) * neq[q];
            if (D == 3) {
                if (az * az) pzz += neq[q];
                if (ax * az) pxz += C(ax * az) * neq[q];
                if (ay * az) pyz += C(ay * az) * neq[q];
            }
        }
        if (p.has_force) {
            pxx += dt * F[0] * u[0];
            pyy += dt * F[1] * u[1];
            pxy += C(0.5) * dt * (F[0] * u[0] + F[1] * u[1]);
            if (D == 3) {
                pzz += dt * F[2] * u[2];
                pxz += C(0.5) * dt * (F[0] * u[0] + F[2] * u[2]);
                pyz += C(0.5) * dt * (F[1] * u[1] + F[2] * u[2]);
            }
        }
        const C t0 = C(p.tau0);
        if (p.les == 1) {
            C pi2 = pxx * pxx + pyy * pyy + pzz * pzz + C(2) * (pxy * pxy + pxz * pxz + pyz * pyz);
            C tau_t = C(0.5) * (sqrt(t0 * t0 + C(p.cs18) * sqrt(C(2) * pi2)) - t0);
            tau = p.les_tau_mode ? t0 + tau_t : tau_t;
        } else {
            const C pi[6] = {pxx, pyy, pzz, pxy, pxz, pyz};
            bool fb;
            tau = t0 + mlp_tau_t<C>(*p.mlp, pi, rho, drho, u, t0, C(p.cs18), &fb);
        }
    }
    if (COLL == 0) {
        const C inv_tau = C(1) / tau;
        const C fcoef = (C(1) - C(0.5) * inv_tau) * dt;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            C v = f[q] - neq[q] * inv_tau;
            if (p.has_force) v += fcoef * fi[q];
            f[q] = v;
        }
    } else {
        constexpr MrtTable<D> T{};
        const C s_nu = C(1) / tau;
        C r[Q];
#pragma unroll
        for (int k = 0; k < Q; ++k) {
            C mk = C(0);
#pragma unroll
            for (int i = 0; i < Q; ++i) {
                if (T.m[k][i] == 1) mk += neq[i];
                else if (T.m[k][i] == -1) mk -= neq[i];
                else if (T.m[k][i] != 0) mk += C(T.m[k][i]) * neq[i];
            }
            const C sk = L::viscous_row[k] ? s_nu : C(L::s_base[k]);
            C v = sk * mk;
            if (p.has_force) {
                C mf = C(0);
#pragma unroll
                for (int i = 0; i < Q; ++i) {
                    if (T.m[k][i] == 1) mf += fi[i];
                    else if (T.m[k][i] == -1) mf -= fi[i];
                    else if (T.m[k][i] != 0) mf += C(T.m[k][i]) * fi[i];
                }
                v -= (C(1) - C(0.5) * sk) * dt * mf;
            }
            r[k] = v * C(1.0 / double(T.n2[k]));
        }
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            C acc = C(0);
#pragma unroll
            for (int k = 0; k < Q; ++k) {
                if (T.m[k][i] == 1) acc += r[k];
                else if (T.m[k][i] == -1) acc -= r[k];
                else if (T.m[k][i] != 0) acc += C(T.m[k][i]) * r[k];
            }
            f[i] -= acc;
        }
    }
}

// workgroup -> brick group mapping that gives each of the 8 XCDs a contiguous range of bricks
// (Morton order), so neighbouring bricks meet in the same L2
__device__ __forceinline__ int64_t xcd_group(int64_t ngroups) {
    const int64_t nb = gridDim.x;
    const int64_t per = nb >> 3;   // grid is a multiple of 8
    const int64_t bid = blockIdx.x;
    const int64_t g = (bid & 7) * per + (bid >> 3);
    return g < ngroups ? g : -1;
}

// Fused pull-stream + collide over all nodes of a level (one wavefront per brick).
template <int D, typename T, typename C, bool SHIFT, int COLL>
__global__ __launch_bounds__(256) void k_stream_collide(BrickLevel<T> lv, const T *__restrict__ fin,
                                                        T *__restrict__ fout, T *__restrict__ aux_f,
                                                        StepParams p) {
    using L = Lattice<D>;
    using B = Brick<D>;
    constexpr int Q = L::Q;
    constexpr int E = B::edge;
    constexpr int NN = B::nnbr;
    const int64_t g = xcd_group((lv.nbrick + 3) >> 2);
    if (g < 0) return;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int64_t b = g * 4 + wave;
    if (b >= lv.nbrick) return;
    const uint64_t occ = lv.occ[b];
    if (!((occ >> lane) & 1ull)) return;
    const uint64_t am = lv.auxmask[b];
    const bool is_aux = (am >> lane) & 1ull;
    int64_t ai = 0;
    uint32_t missing = 0;
    if (is_aux) {
        ai = lv.auxbase[b] + __popcll(am & ((1ull << lane) - 1ull));
        missing = lv.aux_missing[ai];
    }
    int32_t nb[NN];
#pragma unroll
    for (int k = 0; k < NN; ++k) nb[k] = lv.nbr[b * NN + k];
    const int lx = B::lane_coord(lane, 0), ly = B::lane_coord(lane, 1);
    const int lz = D == 3 ? B::lane_coord(lane, 2) : 0;
    const int64_t ns = lv.nslot;
    C f[Q];
    f[0] = C(fin[b * 64 + lane]);
#pragma unroll
    for (int q = 1; q < Q; ++q) {
        const int sx = lx - L::cx[q], sy = ly - L::cy[q], sz = lz - L::cz[q];
        // the source lies in this brick or in the neighbour brick across the faces that c_q
        // crosses; select among those wave-uniform brick indices (indices fold after unrolling)
        const bool ex = L::cx[q] != 0 && (sx < 0 || sx >= E);
        const bool ey = L::cy[q] != 0 && (sy < 0 || sy >= E);
        const bool ez = D == 3 && L::cz[q] != 0 && (sz < 0 || sz >= E);
        constexpr int base = D == 3 ? 13 : 4;
        const int dx = -L::cx[q], dy = -3 * L::cy[q], dz = D == 3 ? -9 * L::cz[q] : 0;
        const int sb = ez ? (ey ? (ex ? nb[base + dx + dy + dz] : nb[base + dy + dz])
                                : (ex ? nb[base + dx + dz] : nb[base + dz]))
                          : (ey ? (ex ? nb[base + dx + dy] : nb[base + dy])
                                : (ex ? nb[base + dx] : nb[base]));
        const int sl = B::lane_of(sx & (E - 1), sy & (E - 1), sz & (E - 1));
        if (missing & (1u << q)) f[q] = aux_f[q * lv.naux + ai];
        else f[q] = C(fin[q * ns + (int64_t)sb * 64 + sl]);
    }
    if (is_aux) {
#pragma unroll
        for (int q = 0; q < Q; ++q) aux_f[q * lv.naux + ai] = T(f[q]);
        return;
    }
    C F[3] = {C(p.const_force[0]), C(p.const_force[1]), C(p.const_force[2])};
    C rho, u[3];
    collide<D, SHIFT, COLL, C>(f, F, p, rho, u);
#pragma unroll
    for (int q = 0; q < Q; ++q) fout[q * ns + b * 64 + lane] = T(f[q]);
}

// Collision of the aux nodes (deferred from the previous substep), writing post-collision
// values into the buffer the next fused kernel reads, plus rho/u (GridNodeLbm::rho_/velocity_).
template <int D, typename T, typename C, bool SHIFT, int COLL>
__global__ __launch_bounds__(256) void k_aux_collide(BrickLevel<T> lv, const T *__restrict__ aux_f,
                                                     const double *__restrict__ aux_force,
                                                     T *__restrict__ fa, double *__restrict__ aux_macro,
                                                     StepParams p) {
    constexpr int Q = Lattice<D>::Q;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= lv.naux) return;
    C f[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) f[q] = C(aux_f[q * lv.naux + i]);
    C F[3] = {C(p.const_force[0]), C(p.const_force[1]), C(p.const_force[2])};
    if (p.has_force && aux_force) {
        F[0] += C(aux_force[i * 3]); F[1] += C(aux_force[i * 3 + 1]); F[2] += C(aux_force[i * 3 + 2]);
    }
    C rho, u[3];
    collide<D, SHIFT, COLL, C>(f, F, p, rho, u);
    const int64_t slot = lv.aux_slot[i];
#pragma unroll
    for (int q = 0; q < Q; ++q) fa[q * lv.nslot + slot] = T(f[q]);
    aux_macro[i * 4] = double(rho);
    aux_macro[i * 4 + 1] = double(u[0]);
    aux_macro[i * 4 + 2] = double(u[1]);
    aux_macro[i * 4 + 3] = double(u[2]);
}

}  // namespace zlh
