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
#include "mrt_d3q19.hpp"

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
    int mlp_operands;      // 0 BF16, 1 FP16 operands of the FP32 matrix-core network
};

// Populations of a level are brick-major, [brick][q][64 lanes]: the Q x 64 values of one brick
// are one contiguous run (Q * 256 B in FP32), so a wave's pulls from its own and its 26
// neighbour bricks touch few pages and its store is a single run.  slot = brick * 64 + lane.
template <int Q>
__host__ __device__ __forceinline__ int64_t pop_at(int q, int64_t slot) {
    return ((slot >> 6) * Q + q) * 64 + (slot & 63);
}

// Storage form of ZLH_BF16_SHIFTED: f - w in bfloat16, converted to and from float with
// round-to-nearest-even (v_cvt_pk_bf16_f32).
typedef __bf16 bf16_t;

// streaming store of one population (2-byte storage goes out as its bit pattern)
template <typename T>
__device__ __forceinline__ void store_nt(T v, T *dst) {
    if constexpr (sizeof(T) == 2) __builtin_nontemporal_store(__builtin_bit_cast(uint16_t, v), reinterpret_cast<uint16_t *>(dst));
    else __builtin_nontemporal_store(v, dst);
}

template <typename T>
struct BrickLevel {
    int64_t nbrick;
    int64_t nslot;            // nbrick * 64
    int64_t naux;
    const uint64_t *occ;      // [nbrick]
    const int32_t *nbr;       // [nbrick][NN]
    const uint64_t *auxmask;  // [nbrick] nodes whose post-stream values go to the aux array
    const uint64_t *defmask;  // [nbrick] of those, the nodes whose collision waits for the aux collision
    const int32_t *auxbase;   // [nbrick]
    const uint32_t *aux_missing;  // [naux]
    const int32_t *aux_slot;  // [naux]
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
        C inv = recip(rho);
        u[0] = jx * inv; u[1] = jy * inv; u[2] = D == 3 ? jz * inv : C(0);
    } else {
        u[0] = jx; u[1] = jy; u[2] = D == 3 ? jz : C(0);
    }
}

// sub-grid model code of a collision variant: 0 none, 1 Smagorinsky with tau0 + tau_t (physical),
// 2 learned closure (BF16 operands in FP32), 3 Smagorinsky with the reference's tau_eff = tau_t,
// 4 learned closure (FP16 operands in FP32)
__host__ __device__ constexpr int les_code(int les, int les_tau_mode, int mlp_operands) {
    return les == 1 ? (les_tau_mode ? 1 : 3) : (les == 2 ? (mlp_operands ? 4 : 2) : les);
}

// relaxation time of a node: tau0, or with a sub-grid model (les_code) from the non-equilibrium
// stress pi = (xx, yy, zz, xy, xz, yz): Smagorinsky (les_models/lbm_les_smagorinsky.cpp:22-94;
// the reference sets tau_eff = tau_t, lbm_solver_interface.cpp:276-295) or the learned closure
template <typename C>
__device__ __forceinline__ C relaxation_time(int les, const C *pi, C rho, C drho, const C (&u)[3], const StepParams &p,
                                             const MlpOps *ops) {
    const C t0 = C(p.tau0);
    if (!les) return t0;
    if (les == 1 || les == 3) {
        const C pi2 = pi[0] * pi[0] + pi[1] * pi[1] + pi[2] * pi[2] + C(2) * (pi[3] * pi[3] + pi[4] * pi[4] + pi[5] * pi[5]);
        const C tau_t = C(0.5) * (root(t0 * t0 + C(p.cs18) * root(C(2) * pi2)) - t0);
        return les == 1 ? t0 + tau_t : tau_t;
    }
    if constexpr (std::is_same<C, float>::value) {
        // matrix-core network across the wavefront (callers keep every lane here); ops: the
        // operand images, in LDS in the bulk kernel (its address space must stay visible)
        if (les == 4) return t0 + mlp_tau_t_wave<OperandsF16>(*ops, pi, rho, drho, u, t0, C(p.cs18));
        return t0 + mlp_tau_t_wave<OperandsBf16>(*ops, pi, rho, drho, u, t0, C(p.cs18));
    } else {
        bool fb;
        return t0 + mlp_tau_t<C>(*p.mlp, pi, rho, drho, u, t0, C(p.cs18), &fb);
    }
}

// forcing correction of the non-equilibrium stress (CalSgsRelaxationTimeWithForce,
// les_models/lbm_les_smagorinsky.cpp:48-94)
template <int D, typename C>
__device__ __forceinline__ void stress_force_terms(C *pi, const C *F, const C (&u)[3], C dt) {
    pi[0] += dt * F[0] * u[0];
    pi[1] += dt * F[1] * u[1];
    pi[3] += C(0.5) * dt * (F[0] * u[0] + F[1] * u[1]);
    if (D == 3) {
        pi[2] += dt * F[2] * u[2];
        pi[4] += C(0.5) * dt * (F[0] * u[0] + F[2] * u[2]);
        pi[5] += C(0.5) * dt * (F[1] * u[1] + F[2] * u[2]);
    }
}

// non-equilibrium stress sum_q c c (f - feq) summed per direction (xx, yy, zz, xy, xz, yz)
template <int D, bool SHIFT, typename C>
__device__ __forceinline__ void neq_stress(const C (&f)[Lattice<D>::Q], C rho, C drho, const C (&u)[3], int compressible,
                                           C *pi) {
    using L = Lattice<D>;
    const C usq = u[0] * u[0] + u[1] * u[1] + (D == 3 ? u[2] * u[2] : C(0));
#pragma unroll
    for (int c = 0; c < 6; ++c) pi[c] = C(0);
#pragma unroll
    for (int q = 1; q < L::Q; ++q) {
        const C d = f[q] - feq_stored<D, SHIFT, C>(q, rho, drho, u[0], u[1], u[2], usq, compressible);
        const int ax = L::cx[q], ay = L::cy[q], az = L::cz[q];
        if (ax * ax) pi[0] += d;
        if (ay * ay) pi[1] += d;
        if (ax * ay) pi[3] += C(ax * ay) * d;
        if (D == 3) {
            if (az * az) pi[2] += d;
            if (ax * az) pi[4] += C(ax * az) * d;
            if (ay * az) pi[5] += C(ay * az) * d;
        }
    }
}

// MRT collision in moment space (LbmMrtCollisionOpt::CollisionOperator,
// lbm_collision_model_mrt.cpp:114-143: f* = f - M^-1 S M (f - feq) + M^-1 (I - S/2) M F_i dt).
// The moments m = M f come from pair sums/differences of opposite populations (for D3Q19 in the
// factored form of mrt_d3q19.hpp, which shares sums between rows); M feq, M F_i and
// the stress sum_q c c (f - feq) are closed forms in the moments (MomentSpace), so no per-
// direction equilibrium is formed; the post-collision moments return through
// f* = M^T diag(1/|M_k|^2) m*.  Stored (shifted) populations map to shifted moments
// m~ = m - M w, and M^-1 (M w) = w keeps the stored form exact.
template <int D, bool SHIFT, typename C, int VAR>
__device__ __forceinline__ void collide_mrt(C (&f)[Lattice<D>::Q], const C *F, const StepParams &p, C &rho_out,
                                            C (&u)[3], const MlpOps *ops) {
    using L = Lattice<D>;
    constexpr int Q = L::Q;
    constexpr MomentSpace<D> S{};
    constexpr int NP = S.NP;
    const int les = VAR >= 0 ? ((VAR & 15) >> 1) : les_code(p.les, p.les_tau_mode, p.mlp_operands);
    const bool forced = VAR >= 0 ? (VAR & 1) != 0 : p.has_force != 0;
    const C dt = C(p.dt);
    // the transforms run on the matrix cores in the kVarMrtMfma variants of the FP32 learned-
    // closure kernels (every lane of the wavefront collides there), on the vector units otherwise
    C mm[Q];
    if constexpr (mrt_on_mfma<C, VAR>()) {
        moments_mfma<Q>(*ops, f, mm);
    } else if constexpr (D == 3) {
        mrt19_moments<C>(f, mm);
    } else {
        C ps[NP], pd[NP];
#pragma unroll
        for (int j = 0; j < NP; ++j) {
            ps[j] = f[S.rep[j]] + f[S.opp[j]];
            pd[j] = f[S.rep[j]] - f[S.opp[j]];
        }
#pragma unroll
        for (int k = 0; k < Q; ++k) {
            C acc = S.even[k] ? C(S.m[k][0]) * f[0] : C(0);
#pragma unroll
            for (int j = 0; j < NP; ++j) {
                const int c = S.m[k][S.rep[j]];
                const C v = S.even[k] ? ps[j] : pd[j];
                if (c == 1) acc += v;
                else if (c == -1) acc -= v;
                else if (c != 0) acc += C(c) * v;
            }
            mm[k] = acc;
        }
    }
    // macroscopic variables (CalMacro*/CalMacroForce*, lbm_solver_interface.cpp:477-592): rows 0
    // and 3, 5 (, 7) are the density and the momentum
    C rho, drho, base;
    if (SHIFT) { drho = mm[0]; rho = C(1) + drho; base = drho; }
    else { rho = mm[0]; drho = rho - C(1); base = rho; }
    C j[3] = {mm[3], mm[5], D == 3 ? mm[7] : C(0)};
    if (forced) {
#pragma unroll
        for (int a = 0; a < D; ++a) j[a] += C(0.5) * F[a] * dt;
    }
    const C rhoc = p.compressible ? rho : C(1);
    {
        const C inv = recip(rhoc);
        u[0] = p.compressible ? j[0] * inv : j[0];
        u[1] = p.compressible ? j[1] * inv : j[1];
        u[2] = D == 3 ? (p.compressible ? j[2] * inv : j[2]) : C(0);
    }
    rho_out = rho;
    const C uu[3][3] = {{u[0] * u[0], u[0] * u[1], u[0] * u[2]},
                        {u[1] * u[0], u[1] * u[1], u[1] * u[2]},
                        {u[2] * u[0], u[2] * u[1], u[2] * u[2]}};
    // non-equilibrium moments of the rows that relax (s_k != 0)
    C nm[Q];
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        nm[k] = C(0);
        if (!L::viscous_row[k] && L::s_base[k] == 0.) continue;
        C poly = C(0);
#pragma unroll
        for (int x = 0; x < D; ++x) {
            if (S.b[k][x] != 0.) poly += C(S.b[k][x]) * u[x];
#pragma unroll
            for (int y = 0; y < D; ++y) {
                const double cxy = S.g[k][x][y] - (x == y ? 1.5 * S.a[k] : 0.);
                if (cxy != 0.) poly += C(cxy) * uu[x][y];
            }
        }
        C meq = rhoc * poly;
        if (S.a[k] != 0.) meq += C(S.a[k]) * base;
        nm[k] = mm[k] - meq;
    }
    C pi[6];
    if (les == 3) {
        // the reference's tau_eff = tau_t divides by a stress that is exactly zero at equilibrium;
        // the per-direction sum reproduces that zero (and the resulting non-finite state)
        neq_stress<D, SHIFT, C>(f, rho, drho, u, p.compressible, pi);
        if (forced) stress_force_terms<D, C>(pi, F, u, dt);
    } else if (les) {
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            C acc = C(0);
#pragma unroll
            for (int k = 0; k < Q; ++k)
                if (S.pi[c][k] != 0. && (L::viscous_row[k] || L::s_base[k] != 0.)) acc += C(S.pi[c][k]) * nm[k];
            pi[c] = acc;
        }
        if (forced) stress_force_terms<D, C>(pi, F, u, dt);
    }
    C uF = C(0);
    if (forced) uF = u[0] * F[0] + u[1] * F[1] + (D == 3 ? u[2] * F[2] : C(0));
    // post-collision moments over |M_k|^2; the rows with fixed rates first, so that only the
    // viscous rows' moments stay live across the sub-grid model
    C y[Q];
    auto post = [&](int k, C sk, bool relaxes) {
        C v = relaxes ? mm[k] - sk * nm[k] : mm[k];
        if (forced) {
            // Guo forcing moment b.F - 3 (u.F) a + 2 u.g.F
            C fm = C(0);
#pragma unroll
            for (int x = 0; x < D; ++x) {
                if (S.b[k][x] != 0.) fm += C(S.b[k][x]) * F[x];
#pragma unroll
                for (int z = 0; z < D; ++z)
                    if (S.g[k][x][z] != 0.) fm += C(2. * S.g[k][x][z]) * u[x] * F[z];
            }
            if (S.a[k] != 0.) fm -= C(3. * S.a[k]) * uF;
            v += (relaxes ? C(1) - C(0.5) * sk : C(1)) * dt * fm;
        }
        y[k] = v * C(1.0 / double(S.n2[k]));
    };
#pragma unroll
    for (int k = 0; k < Q; ++k)
        if (!L::viscous_row[k]) post(k, C(L::s_base[k]), L::s_base[k] != 0.);
    const C tau = les ? relaxation_time<C>(les, pi, rho, drho, u, p, ops) : C(p.tau0);
    const C s_nu = recip(tau);
#pragma unroll
    for (int k = 0; k < Q; ++k)
        if (L::viscous_row[k]) post(k, s_nu, true);
    if constexpr (mrt_on_mfma<C, VAR>()) {
        populations_mfma<Q>(*ops, y, f);
        return;
    }
    if constexpr (D == 3) {
        mrt19_populations<C>(y, f);
        return;
    }
    // back to populations: even rows give the pair mean, odd rows the pair difference
    {
        C acc = C(0);
#pragma unroll
        for (int k = 0; k < Q; ++k)
            if (S.even[k] && S.m[k][0] != 0) acc += C(S.m[k][0]) * y[k];
        f[0] = acc;
    }
#pragma unroll
    for (int jp = 0; jp < NP; ++jp) {
        C e = C(0), o = C(0);
#pragma unroll
        for (int k = 0; k < Q; ++k) {
            const int c = S.m[k][S.rep[jp]];
            if (c == 0) continue;
            C &acc = S.even[k] ? e : o;
            if (c == 1) acc += y[k];
            else if (c == -1) acc -= y[k];
            else acc += C(c) * y[k];
        }
        f[S.rep[jp]] = e + o;
        f[S.opp[jp]] = e - o;
    }
}

// One node collision on stored populations (in place): func_collision_node_
// (lbm_solver_interface.cpp:266-300) with SRT (lbm_collision_model_srt.cpp:20-39) or MRT
// (collide_mrt).  VAR = 2 * les_code + forced fixes the sub-grid model and the forcing at compile time
// (the bulk kernel then keeps only the registers its variant needs); VAR < 0 reads both from p.
template <int D, bool SHIFT, int COLL, typename C, int VAR = -1>
__device__ __forceinline__ void collide(C (&f)[Lattice<D>::Q], const C *F, const StepParams &p,
                                        C &rho_out, C (&u)[3], const MlpOps *ops = nullptr) {
    if constexpr (COLL == 1) {
        collide_mrt<D, SHIFT, C, VAR>(f, F, p, rho_out, u, ops);
        return;
    }
    using L = Lattice<D>;
    constexpr int Q = L::Q;
    const int les = VAR >= 0 ? ((VAR & 15) >> 1) : les_code(p.les, p.les_tau_mode, p.mlp_operands);
    const bool forced = VAR >= 0 ? (VAR & 1) != 0 : p.has_force != 0;
    const C dt = C(p.dt);
    C rho, drho;
    macro<D, SHIFT, C>(f, F, dt, forced, p.compressible, rho, drho, u);
    rho_out = rho;
    C usq = u[0] * u[0] + u[1] * u[1] + (D == 3 ? u[2] * u[2] : C(0));
    C neq[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q)
        neq[q] = f[q] - feq_stored<D, SHIFT, C>(q, rho, drho, u[0], u[1], u[2], usq, p.compressible);
    C tau = C(p.tau0);
    if (les) {
        C pi[6] = {C(0), C(0), C(0), C(0), C(0), C(0)};
#pragma unroll
        for (int q = 1; q < Q; ++q) {
            const int ax = L::cx[q], ay = L::cy[q], az = L::cz[q];
            if (ax * ax) pi[0] += neq[q];
            if (ay * ay) pi[1] += neq[q];
            if (ax * ay) pi[3] += C(ax * ay) * neq[q];
            if (D == 3) {
                if (az * az) pi[2] += neq[q];
                if (ax * az) pi[4] += C(ax * az) * neq[q];
                if (ay * az) pi[5] += C(ay * az) * neq[q];
            }
        }
        if (forced) stress_force_terms<D, C>(pi, F, u, dt);
        tau = relaxation_time<C>(les, pi, rho, drho, u, p, ops);
    }
    const C inv_tau = recip(tau);
    const C fcoef = (C(1) - C(0.5) * inv_tau) * dt;
    C uF = C(0);
    if (forced) uF = u[0] * F[0] + u[1] * F[1] + (D == 3 ? u[2] * F[2] : C(0));
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        C v = f[q] - neq[q] * inv_tau;
        if (forced) {
            // Guo term w [3 (c - u).F + 9 (c.u)(c.F)] (CalForcingTerming3D, lbm_solver_interface.cpp:443-470)
            const C cF = C(L::cx[q]) * F[0] + C(L::cy[q]) * F[1] + (D == 3 ? C(L::cz[q]) * F[2] : C(0));
            const C cu = C(L::cx[q]) * u[0] + C(L::cy[q]) * u[1] + (D == 3 ? C(L::cz[q]) * u[2] : C(0));
            v += fcoef * C(L::w[q]) * (C(3) * (cF - uF) + C(9) * cu * cF);
        }
        f[q] = v;
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

// Fused pull-stream + collide of one brick by one wavefront.
template <int D, typename T, typename C, bool SHIFT, int COLL, int VAR>
__device__ __forceinline__ void stream_collide_brick(const BrickLevel<T> &lv, const T *__restrict__ fin,
                                                     T *__restrict__ fout, T *__restrict__ aux_f, const StepParams &p,
                                                     int64_t b, const MlpOps *ops) {
    using L = Lattice<D>;
    using B = Brick<D>;
    constexpr int Q = L::Q;
    constexpr int E = B::edge;
    constexpr int NN = B::nnbr;
    static_assert(!mlp_on_mfma<C, VAR>(), "the FP32 learned closure runs in k_stream_collide_mlp");
    constexpr bool kWave = false;
    const int lane = threadIdx.x & 63;
    const uint64_t occ = lv.occ[b];
    const bool has_node = (occ >> lane) & 1ull;
    if (!kWave && !has_node) return;
    const uint64_t am = lv.auxmask[b];
    const bool is_aux = (am >> lane) & 1ull;
    const bool deferred = (lv.defmask[b] >> lane) & 1ull;
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
    C f[Q];
    const T *__restrict__ own = fin + b * Q * 64 + lane;
    f[0] = kWave && !has_node ? C(0) : C(own[0]);
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
        if (kWave && !has_node) f[q] = C(0);
        else if (missing & (1u << q)) f[q] = aux_f[q * lv.naux + ai];
        else f[q] = C(fin[((int64_t)sb * Q + q) * 64 + sl]);
    }
    if (is_aux) {
#pragma unroll
        for (int q = 0; q < Q; ++q) aux_f[q * lv.naux + ai] = T(f[q]);
        // a recorded node that nothing changes before its next collision collides here
        if (!kWave && deferred) return;
    }
    C F[3] = {C(p.const_force[0]), C(p.const_force[1]), C(p.const_force[2])};
    C rho, u[3];
    collide<D, SHIFT, COLL, C, VAR>(f, F, p, rho, u, ops);
    if (kWave && (!has_node || deferred)) return;
    // streaming stores: the next substep reads these values through other CUs' caches
    T *__restrict__ dst = fout + b * Q * 64 + lane;
#pragma unroll
    for (int q = 0; q < Q; ++q) store_nt(T(f[q]), dst + q * 64);
}

// Fused pull-stream + collide over all nodes of a level, one wavefront per brick, four bricks
// per workgroup.
template <int D, typename T, typename C, bool SHIFT, int COLL, int VAR>
__global__ __launch_bounds__(256) void k_stream_collide(BrickLevel<T> lv, const T *__restrict__ fin,
                                                        T *__restrict__ fout, T *__restrict__ aux_f,
                                                        StepParams p) {
    const int64_t g = xcd_group((lv.nbrick + 3) >> 2);
    if (g < 0) return;
    const int64_t b = g * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (b >= lv.nbrick) return;
    stream_collide_brick<D, T, C, SHIFT, COLL, VAR>(lv, fin, fout, aux_f, p, b, nullptr);
}

// Read-only brick tables of a level (the kernel takes them as no-alias parameters: it stores into
// the population arrays inside its loop, and only provably unclobbered tables are read through
// the scalar cache, whose loads do not wait on the vector memory counter).
struct BrickTables {
    const uint64_t *__restrict__ occ;
    const uint64_t *__restrict__ auxmask;
    const int32_t *__restrict__ auxbase;
    const int32_t *__restrict__ nbr;
    const uint32_t *__restrict__ aux_missing;
};

// Per-lane constants of the pull, fixed by the brick layout and computed once per pull: the
// byte offset 4 lane_of() of the source coordinate along each axis for c = +1, 0, -1, and
// whether the lane sits on the low / high face of each axis.  lane_of is a sum of one term per
// axis, so a source slot is the OR of three of these.
template <int D>
struct PullLane {
    uint32_t s4[3][3];   // [axis][1 - c]
    bool lo[3], hi[3];
};

template <int D>
__device__ __forceinline__ PullLane<D> pull_lane(int lane) {
    PullLane<D> P;
    const uint32_t l4 = 4u * (uint32_t)lane;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        constexpr uint32_t kNone = 0u;
        const uint32_t m4 = a < D ? lane_axis_mask4<D>(a) : kNone;
        const uint32_t own4 = l4 & m4;
        // k = 1 - c: the source coordinate along a is l - 1 (k = 0), l (1) or l + 1 (2)
        P.s4[a][0] = a < D ? lane_axis_step4(own4, m4, -1) : 0u;
        P.s4[a][1] = own4;
        P.s4[a][2] = a < D ? lane_axis_step4(own4, m4, 1) : 0u;
        P.lo[a] = a < D && own4 == 0u;
        P.hi[a] = a < D && own4 == m4;
    }
    return P;
}

// Pull of brick b (every lane, every direction) into a per-wave LDS buffer [Q][64] with one
// direct-to-LDS dword load per direction (gfx950 global_load_lds_dword: lane i lands at
// buf + 4 i).  Every lane loads from a valid address of fin: an absent neighbour brick is
// replaced by b, a deferred node's missing source loads whatever slot the stencil names, and a
// lane without a node loads the slot its stencil names inside brick b; the consumer drops those values (a deferred node's
// missing directions keep the boundary / transfer values already in the aux array).  Addresses are 32-bit byte offsets
// from the direction's base fin + 64 q.  Returns the lane's missing-source mask (deferred nodes).
template <int D, typename T>
__device__ __forceinline__ uint32_t pull_to_lds(const BrickTables &t, const T *__restrict__ fin, int64_t b, int lane,
                                                T *buf) {
    using L = Lattice<D>;
    using B = Brick<D>;
    constexpr int Q = L::Q;
    constexpr int NN = B::nnbr;
    static_assert(sizeof(T) == 4, "dword staging");
    const PullLane<D> P = pull_lane<D>(lane);   // per pull: live only while the addresses form
    // opaque per pull: the direction's LDS destinations are formed here, not hoisted out of the
    // caller's loop into 19 long-lived scalars
    asm volatile("" : "+s"(buf));
    // byte offsets of the neighbour bricks (uniform: scalar multiplies)
    constexpr uint32_t kBrickBytes = (uint32_t)(Q * 64 * sizeof(T));
    uint32_t nbo[NN];
#pragma unroll
    for (int k = 0; k < NN; ++k) {
        const int32_t v = t.nbr[b * NN + k];
        nbo[k] = (uint32_t)(v < 0 ? (int32_t)b : v) * kBrickBytes;
    }
    // a lane crosses into a neighbour brick only if it holds a node: a lane without one loads
    // the slot its stencil names inside its own brick, a line the brick's own pull fetches anyway
    const bool has_node = (t.occ[b] >> lane) & 1ull;
    bool clo[3], chi[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        clo[a] = P.lo[a] && has_node;
        chi[a] = P.hi[a] && has_node;
    }
    constexpr int base = D == 3 ? 13 : 4;
    // the own brick's offset in a vector register, so that each selection below takes one
    // uniform operand and this one
    uint32_t self = nbo[base];
    asm volatile("" : "+v"(self));
    // addresses by selection, no branches: a branch between two of these loads makes the
    // compiler wait on the vector memory counter there.  D2Q9 / D3Q19 directions cross at most
    // two axes: the brick is chosen along the first axis, then the second.
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int cx = L::cx[q], cy = L::cy[q], cz = D == 3 ? L::cz[q] : 0;
        const bool ex = cx == 1 ? clo[0] : (cx == -1 ? chi[0] : false);
        const bool ey = cy == 1 ? clo[1] : (cy == -1 ? chi[1] : false);
        const bool ez = cz == 1 ? clo[2] : (cz == -1 ? chi[2] : false);
        const int dx = -cx, dy = -3 * cy, dz = -9 * cz;
        uint32_t sel;
        if (cx != 0 && cy != 0)
            sel = ey ? (ex ? nbo[base + dx + dy] : nbo[base + dy]) : (ex ? nbo[base + dx] : self);
        else if (cx != 0 && cz != 0)
            sel = ez ? (ex ? nbo[base + dx + dz] : nbo[base + dz]) : (ex ? nbo[base + dx] : self);
        else if (cy != 0 && cz != 0)
            sel = ez ? (ey ? nbo[base + dy + dz] : nbo[base + dz]) : (ey ? nbo[base + dy] : self);
        else if (cx != 0)
            sel = ex ? nbo[base + dx] : self;
        else if (cy != 0)
            sel = ey ? nbo[base + dy] : self;
        else if (cz != 0)
            sel = ez ? nbo[base + dz] : self;
        else
            sel = self;
        const uint32_t off = sel + (P.s4[0][1 - cx] | P.s4[1][1 - cy] | P.s4[2][1 - cz]);
        // an opaque scalar base, so that the load takes the scalar-base + 32-bit-offset form
        const char *qbase = reinterpret_cast<const char *>(fin + q * 64);
        asm volatile("" : "+s"(qbase));
        const char *src = qbase + off;
        __builtin_amdgcn_global_load_lds((const void *)src, (__attribute__((address_space(3))) void *)(buf + q * 64),
                                         4, 0, 0);
    }
    uint32_t missing = 0;
    const uint64_t am = t.auxmask[b];
    if ((am >> lane) & 1ull) missing = t.aux_missing[t.auxbase[b] + __popcll(am & ((1ull << lane) - 1ull))];
    return missing;
}

// The same with the FP32 learned closure.  Each workgroup stages the network's operand images
// in LDS once (after its waves' first pulls are out); its four waves then take ROUNDS rounds of
// four adjacent bricks (the host picks 1 or 2), and each wave pulls its next brick into LDS
// while it collides the current one, so the pull's latency hides behind the network's
// arithmetic at the occupancy its registers allow.  Within
// a round the order is: read the staged brick, write deferred nodes' post-stream values, issue
// the next pull, collide, store; the store group is unconditional so that the next round's wait
// (in-order vector memory counter) covers the pull and not these stores.
// Register budget: 4 waves per SIMD unforced, 3 forced (whose Guo terms would spill at 4).
template <int D, typename T, typename C, bool SHIFT, int COLL, int VAR, int ROUNDS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu((VAR & 1) ? 3 : 4)))
void k_stream_collide_mlp(int64_t nbrick, int64_t naux, const uint64_t *__restrict__ occ,
                          const uint64_t *__restrict__ auxmask, const int32_t *__restrict__ auxbase,
                          const int32_t *__restrict__ nbr, const uint32_t *__restrict__ aux_missing,
                          const T *__restrict__ fin, T *__restrict__ fout, T *__restrict__ aux_f, StepParams p) {
    static_assert(mlp_on_mfma<C, VAR>(), "kernel of the FP32 learned closure");
    const BrickTables t{occ, auxmask, auxbase, nbr, aux_missing};
    using L = Lattice<D>;
    constexpr int Q = L::Q;
    __shared__ MlpOps ops;
    __shared__ T stage[4][Q * 64];
    constexpr int64_t per_group = 4 * ROUNDS;
    const int64_t g = xcd_group((nbrick + per_group - 1) / per_group);
    if (g < 0) return;                       // uniform per workgroup: before the barrier
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    T *buf = stage[wave];
    int64_t b = g * per_group + wave;
    // the first pull goes out before the network is staged, so that their latencies overlap
    uint32_t miss = 0;
    if (b < nbrick) miss = pull_to_lds<D, T>(t, fin, b, threadIdx.x & 63, buf);
    stage_mlp_ops<((VAR & 15) >> 1) == 4, mrt_on_mfma<C, VAR>()>(ops, p.mlp->ops);
    if (b >= nbrick) return;                 // uniform per wave; no barrier follows
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");   // the first pull has landed
#pragma unroll 1
    for (int r = 0; r < ROUNDS; ++r) {
        // the pull into buf has landed: the compiler does not order LDS reads after direct-to-LDS
        // loads of an earlier iteration.  In-order counter: the last Q vector memory operations
        // before this point are the previous round's stores, which may stay in flight.
        asm volatile("s_waitcnt vmcnt(%0)" ::"n"(Q) : "memory");
        int lane = threadIdx.x & 63;
        // opaque per brick: keeps the lane's stencil offsets from being hoisted out of the loop
        asm volatile("" : "+v"(lane));
        C f[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) f[q] = C(buf[q * 64 + lane]);
        const bool has_node = (t.occ[b] >> lane) & 1ull;
        const uint64_t am = t.auxmask[b];
        if ((am >> lane) & 1ull) {
            // deferred node: post-stream values to the aux array, except the missing sources
            // (their boundary / transfer values are there already): those go to the spare slot
            // at index Q naux.  Stores take a uniform base and a 32-bit byte offset (the host
            // keeps Q naux 4 below 2^32); a wavefront whose deferred lanes miss no source (a
            // uniform branch) stores without the per-direction selection.
            const uint32_t ai4 = 4u * (uint32_t)(t.auxbase[b] + __popcll(am & ((1ull << lane) - 1ull)));
            using gchar = __attribute__((address_space(1))) char;
            using gT = __attribute__((address_space(1))) T;
            gchar *base = (gchar *)aux_f;
            asm volatile("" : "+s"(base));
            // (each offset is made opaque so that the store keeps the base + 32-bit offset form)
            if (__builtin_amdgcn_ballot_w64(miss != 0u) == 0ull) {
#pragma unroll
                for (int q = 0; q < Q; ++q) {
                    uint32_t o = ai4 + (uint32_t)(4 * q * naux);
                    asm volatile("" : "+v"(o));
                    *(gT *)(base + o) = T(f[q]);
                }
            } else {
                const uint32_t spare4 = (uint32_t)(4 * Q * naux);
#pragma unroll
                for (int q = 0; q < Q; ++q) {
                    uint32_t o = ((miss >> q) & 1u) ? spare4 : ai4 + (uint32_t)(4 * q * naux);
                    asm volatile("" : "+v"(o));
                    *(gT *)(base + o) = T(f[q]);
                }
            }
        }
        // the buffer's reads have returned before it is refilled
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
        const int64_t bn = b + 4;
        const bool more = r + 1 < ROUNDS && bn < nbrick;
        uint32_t miss_next = 0;
        if (more) miss_next = pull_to_lds<D, T>(t, fin, bn, lane, buf);
        if (!has_node) {
#pragma unroll
            for (int q = 0; q < Q; ++q) f[q] = C(0);          // rest state (shifted storage)
        }
        C F[3] = {C(p.const_force[0]), C(p.const_force[1]), C(p.const_force[2])};
        C rho, u[3];
        collide<D, SHIFT, COLL, C, VAR>(f, F, p, rho, u, &ops);
        // every lane stores: empty slots are never read, and deferred slots are rewritten by the
        // next substep's aux collision before the fused kernel reads them
        T *__restrict__ dst = fout + b * Q * 64 + lane;
#pragma unroll
        for (int q = 0; q < Q; ++q) store_nt(T(f[q]), dst + q * 64);
        if (!more) break;
        b = bn;
        miss = miss_next;
    }
}

// Collision of the aux nodes (deferred from the previous substep), writing post-collision
// values into the buffer the next fused kernel reads, plus rho/u (GridNodeLbm::rho_/velocity_) of
// the nodes flagged in macro_flag (the fine-to-coarse sources, the only readers).
// VAR as in the fused kernel; with the FP32 learned closure the workgroup stages the network in
// LDS first.
template <int D, typename T, typename C, bool SHIFT, int COLL, int VAR>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(mlp_on_mfma<C, VAR>() ? ((VAR & 1) ? 3 : 4) : 1)))
void k_aux_collide(BrickLevel<T> lv, const int32_t *__restrict__ list, int64_t nlist,
                                                     const T *__restrict__ aux_f,
                                                     const double *__restrict__ aux_force,
                                                     T *__restrict__ fa, double *__restrict__ aux_macro,
                                                     const uint8_t *__restrict__ macro_flag, StepParams p) {
    constexpr int Q = Lattice<D>::Q;
    constexpr bool kWave = mlp_on_mfma<C, VAR>();
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    // lanes past the end run the collision on the rest state (FP32 learned closure is
    // wavefront-wide) and store nothing
    const bool active = t < nlist;
    // the aux index: entry t of the list of deferred nodes, or every recorded node in turn
    const int64_t i = active ? (list ? (int64_t)list[t] : t) : 0;
    if (!kWave && !active) return;
    C f[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) f[q] = active ? C(aux_f[q * lv.naux + i]) : C(SHIFT ? 0. : Lattice<D>::w[q]);
    // the network is staged after the node's loads are out, so that their latencies overlap
    const MlpOps *ops = nullptr;
    if constexpr (kWave) {
        __shared__ MlpOps s_ops;
        stage_mlp_ops<((VAR & 15) >> 1) == 4, mrt_on_mfma<C, VAR>()>(s_ops, p.mlp->ops);
        ops = &s_ops;
    }
    C F[3] = {C(p.const_force[0]), C(p.const_force[1]), C(p.const_force[2])};
    if (active && (VAR & 1) && aux_force) {
        F[0] += C(aux_force[i * 3]); F[1] += C(aux_force[i * 3 + 1]); F[2] += C(aux_force[i * 3 + 2]);
    }
    C rho, u[3];
    collide<D, SHIFT, COLL, C, VAR>(f, F, p, rho, u, ops);
    if (!active) return;
    const int64_t slot = lv.aux_slot[i];
#pragma unroll
    for (int q = 0; q < Q; ++q) fa[pop_at<Q>(q, slot)] = T(f[q]);
    if (macro_flag[i]) {
        aux_macro[i * 4] = double(rho);
        aux_macro[i * 4 + 1] = double(u[0]);
        aux_macro[i * 4 + 2] = double(u[1]);
        aux_macro[i * 4 + 3] = double(u[2]);
    }
}

}  // namespace zlh
