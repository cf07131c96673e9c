// mlp.hpp -- learned sub-grid closure: a 9 -> 32 -> 32 -> 1 MLP (SiLU, softplus) predicting the
// eddy viscosity from local non-equilibrium invariants.  No counterpart in the reference; it
// replaces the Smagorinsky relaxation time (les_models/lbm_les_smagorinsky.cpp:22-94) where its
// inputs lie inside the training range and falls back to Smagorinsky elsewhere.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <type_traits>

namespace zlh {

constexpr int kMlpF = 9;
constexpr int kMlpH = 32;
constexpr int kMlpParams = kMlpH * kMlpF + kMlpH + kMlpH * kMlpH + kMlpH + kMlpH + 1;   // 1409

// Output row held in accumulator r by lanes of half h (lane / 32) of a 32x32 MFMA result
// (v_mfma_f32_32x32x16_bf16: lane l holds C[i][l % 32] for i = 8 (r / 4) + 4 h + r % 4).
__host__ __device__ constexpr int mfma_row(int h, int r) { return (r >> 2) * 8 + h * 4 + (r & 3); }

// Per-lane operand images of the wavefront MFMA network, as BF16 and as FP16 bit patterns (built
// on the host; w1a/w2a BF16, h1a/h2a the same entries in FP16):
//   w1a[l][e]    = layer-1 A operand, K slot k = 8 (l / 32) + e of row l % 32: W1[l % 32][k] for
//                  the 9 features, b1[l % 32] in slot 9 (its B slot holds 1), zero above
//   w2a[s][l][e] = layer-2 A operand, K step s < 2: W2[l % 32][mfma_row(l / 32, 8 s + e)] (the
//                  layer-1 accumulator order, so activations stay in their registers)
//   w3r[h][r]    = w3[mfma_row(h, r)] (FP32)
//   b2r[h][r]    = b2[mfma_row(h, r)] (FP32): the layer-2 accumulators start from it
// plus the input normalisation, training range, output scale and b3, so that the fused kernel
// reads nothing of the network from global memory after its first store; and the operand images
// of the MRT moment transform on the matrix cores (moments_mfma, populations_mfma):
//   mfwd[s][l] = M[l % 32][2 s + l / 32]   (moment row, population column; zero outside Q x Q)
//   mbwd[s][l] = M[2 s + l / 32][l % 32]   (M^T)
constexpr int kMrtSteps = 10;   // K steps of two: (19 + 1) / 2 for D3Q19, 5 used for D2Q9
struct alignas(16) MlpOps {
    uint16_t w1a[64][8];
    uint16_t w2a[2][64][8];
    uint16_t h1a[64][8];
    uint16_t h2a[2][64][8];
    float w3r[2][16];
    float b2r[2][16];
    float mean[kMlpF];
    float inv_std[kMlpF];
    float fmin[kMlpF];
    float fmax[kMlpF];
    float nu_scale;
    float b3;
    float pad[2];
    float mfwd[kMrtSteps][64];   // last: only the kernels with the transform on the matrix cores stage them
    float mbwd[kMrtSteps][64];
};

// device-resident closure: operand images, parameters (flat order W1, b1, W2, b2, w3, b3), input
// normalisation, training range of the raw features and the output scale
struct MlpDev {
    MlpOps ops;
    float p[kMlpParams];
    float mean[kMlpF];
    float inv_std[kMlpF];
    float fmin[kMlpF];
    float fmax[kMlpF];
    float nu_scale;
};

__device__ __forceinline__ float sigmoidf_(float z) { return 1.f / (1.f + __expf(-z)); }
__device__ __forceinline__ float softplusf_(float z) { return z > 20.f ? z : log1pf(__expf(z)); }

// 1 / x: exact division in FP64; in FP32 the hardware reciprocal (about 1 ulp, one instruction
// where the IEEE-rounded division takes about ten)
template <typename C>
__device__ __forceinline__ C recip(C x) {
    if constexpr (std::is_same<C, float>::value) return __builtin_amdgcn_rcpf(x);
    else return C(1) / x;
}

// square root: correctly rounded in FP64; in FP32 the hardware square root (about 1 ulp, one
// instruction where the correctly rounded sqrtf takes about fifteen)
template <typename C>
__device__ __forceinline__ C root(C x) {
    if constexpr (std::is_same<C, float>::value) return __builtin_amdgcn_sqrtf(x);
    else return sqrt(x);
}

// raw closure inputs from the non-equilibrium stress Pi (xx, yy, zz, xy, xz, yz), density and
// velocity: P = Pi / rho, a = sqrt(2 P:P), I1 = tr P, I2 = (I1^2 - P:P) / 2, I3 = det P, |u|^2,
// rho - 1, tau0, a tau0, a^2
template <typename C>
__device__ __forceinline__ void closure_features(const C *pi, C rho, C drho, const C *u, C tau0, float *x) {
    const C ir = recip(rho);
    const C pxx = pi[0] * ir, pyy = pi[1] * ir, pzz = pi[2] * ir, pxy = pi[3] * ir, pxz = pi[4] * ir, pyz = pi[5] * ir;
    const C pp = pxx * pxx + pyy * pyy + pzz * pzz + C(2) * (pxy * pxy + pxz * pxz + pyz * pyz);
    const C a = root(C(2) * pp);
    const C i1 = pxx + pyy + pzz;
    const C i2 = C(0.5) * (i1 * i1 - pp);
    const C i3 = pxx * (pyy * pzz - pyz * pyz) - pxy * (pxy * pzz - pyz * pxz) + pxz * (pxy * pyz - pyy * pxz);
    x[0] = float(a);
    x[1] = float(i1);
    x[2] = float(i2);
    x[3] = float(i3);
    x[4] = float(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    x[5] = float(drho);
    x[6] = float(tau0);
    x[7] = float(a * tau0);
    x[8] = float(a * a);
}

// network output nu_t / nu_scale for normalised inputs (FP32 on the vector units)
__device__ __forceinline__ float mlp_forward(const MlpDev &m, const float *xn) {
    const float *W1 = m.p, *b1 = W1 + kMlpH * kMlpF, *W2 = b1 + kMlpH, *b2 = W2 + kMlpH * kMlpH;
    const float *w3 = b2 + kMlpH, b3 = w3[kMlpH];
    float h1[kMlpH];
#pragma unroll
    for (int j = 0; j < kMlpH; ++j) {
        float z = b1[j];
#pragma unroll
        for (int i = 0; i < kMlpF; ++i) z += W1[j * kMlpF + i] * xn[i];
        h1[j] = z * sigmoidf_(z);
    }
    float out = b3;
#pragma unroll 4
    for (int j = 0; j < kMlpH; ++j) {
        float z = b2[j];
#pragma unroll
        for (int i = 0; i < kMlpH; ++i) z += W2[j * kMlpH + i] * h1[i];
        out += w3[j] * (z * sigmoidf_(z));
    }
    return softplusf_(out);
}

// relaxation-time increment of the closure; Smagorinsky (with the physical tau mode) where any
// raw input lies outside the training range
template <typename C>
__device__ __forceinline__ C mlp_tau_t(const MlpDev &m, const C *pi, C rho, C drho, const C *u, C tau0, C cs18,
                                      bool *fallback) {
    float x[kMlpF];
    closure_features<C>(pi, rho, drho, u, tau0, x);
    bool inside = true;
#pragma unroll
    for (int i = 0; i < kMlpF; ++i) inside = inside && x[i] >= m.fmin[i] && x[i] <= m.fmax[i];
    if (!inside) {
        *fallback = true;
        const C q = root(C(2) * (pi[0] * pi[0] + pi[1] * pi[1] + pi[2] * pi[2] +
                                 C(2) * (pi[3] * pi[3] + pi[4] * pi[4] + pi[5] * pi[5])));
        return C(0.5) * (root(tau0 * tau0 + cs18 * q) - tau0);
    }
    *fallback = false;
    float xn[kMlpF];
#pragma unroll
    for (int i = 0; i < kMlpF; ++i) xn[i] = (x[i] - m.mean[i]) * m.inv_std[i];
    return C(3) * C(m.nu_scale * mlp_forward(m, xn));
}

// ---------------------------------------------------------------------------------------------
// FP32 solver: the network of 64 nodes (one per lane) as matrix-core products with BF16 or FP16
// operands and FP32 accumulation.  Layer l computes H_l^T [32 x 64 nodes] = W_l [32 x K] *
// H_{l-1}^T in two 32-node column blocks; every lane of the wavefront must call it (EXEC all
// ones).  v_mfma_f32_32x32x16_bf16 and _f16 share one lane layout.
// ---------------------------------------------------------------------------------------------
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8_t __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

// kernel variants: VAR = 2 sub-grid code + forced (+ kVarMrtMfma); sub-grid codes of the learned
// closure: 2 with BF16 operands, 4 with FP16
constexpr int kVarMrtMfma = 16;   // the FP32 learned-closure kernels' MRT transform on the matrix cores
template <typename C, int VAR>
__host__ __device__ constexpr bool mlp_on_mfma() {
    return std::is_same<C, float>::value && (VAR < 0 || ((VAR & 15) >> 1) == 2 || ((VAR & 15) >> 1) == 4);
}
template <typename C, int VAR>
__host__ __device__ constexpr bool mrt_on_mfma() {
    return VAR >= 0 && (VAR & kVarMrtMfma) != 0 && mlp_on_mfma<C, VAR>();
}

struct OperandsBf16 {
    typedef bf16x8_t vec;
    typedef __bf16 elem;
    static __device__ __forceinline__ vec load(const MlpOps &o, const uint16_t *bits) {
        return __builtin_bit_cast(vec, *reinterpret_cast<const u16x8_t *>(bits));
    }
    static __device__ __forceinline__ const uint16_t *w1(const MlpOps &o, int lane) { return o.w1a[lane]; }
    static __device__ __forceinline__ const uint16_t *w2(const MlpOps &o, int s, int lane) { return o.w2a[s][lane]; }
    static __device__ __forceinline__ f32x16_t mfma(vec a, vec b, f32x16_t c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
    }
};
struct OperandsF16 {
    typedef f16x8_t vec;
    typedef _Float16 elem;
    static __device__ __forceinline__ vec load(const MlpOps &o, const uint16_t *bits) {
        return __builtin_bit_cast(vec, *reinterpret_cast<const u16x8_t *>(bits));
    }
    static __device__ __forceinline__ const uint16_t *w1(const MlpOps &o, int lane) { return o.h1a[lane]; }
    static __device__ __forceinline__ const uint16_t *w2(const MlpOps &o, int s, int lane) { return o.h2a[s][lane]; }
    static __device__ __forceinline__ f32x16_t mfma(vec a, vec b, f32x16_t c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
    }
};

// SiLU of the matrix-core form, on pre-activations scaled by c = -log2(e) (the host scales W1, b1
// and b2 in the operand images): for t = c z, t / (1 + 2^t) = c silu(z).  Hardware exp2 and
// reciprocal, about 1 ulp each, far below the operands' rounding; a large t gives 2^t -> inf,
// rcp -> 0.  The next layer takes the scaled activations as they are (its weights are unscaled;
// the output weights carry 1/c).
constexpr double kSiluScale = -1.4426950408889634;
__device__ __forceinline__ float silu_scaled(float t) {
    return t * __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(t));
}

// softplus of the matrix-core form: max(z, 0) + ln 2 log2(1 + 2^(-|z| log2 e)) with the hardware
// exp2 and log2 (about 1 ulp each).  Rounding 1 + e to FP32 costs at most 2^-24 absolute where
// e is small, far below the operands' rounding; the library log1pf takes about 110 instructions.
__device__ __forceinline__ float softplus_fast(float z) {
    const float e = __builtin_amdgcn_exp2f(-1.4426950408889634f * __builtin_fabsf(z));
    return __builtin_fmaxf(z, 0.f) + 0.6931471805599453f * __builtin_amdgcn_logf(1.f + e);
}

// The MRT moment transform m = M f and its inverse f = M^T y as batched 19 x 19 contractions on
// the matrix cores (v_mfma_f32_32x32x2f32, FP32 operands, products exact, FP32 sums): the 64 nodes
// of the wavefront are two 32-node column blocks, a K step takes populations 2s and 2s + 1.  The
// B operand of block 0 holds population 2s of node l (lanes 0..31) and 2s + 1 of node l - 32
// (lanes 32..63); block 1 the same for nodes 32..63: v_permlane32_swap(f[2s], f[2s + 1]) gives
// both.  The results (rows in the accumulator order of a 32 x 32 product) return to one node per
// lane by v_permlane32_swap(block 0, block 1) of each accumulator register: then register r of
// the first result holds row 8 (r / 4) + r % 4 and of the second row 8 (r / 4) + 4 + r % 4.
// Every lane of the wavefront must call them (EXEC all ones).
__device__ __forceinline__ constexpr int mrt_reg(int k) { return 4 * (k >> 3) + (k & 3); }
__device__ __forceinline__ constexpr bool mrt_upper(int k) { return (k & 4) != 0; }

template <int Q>
__device__ __forceinline__ void rows_to_lanes(const f32x16_t &d0, const f32x16_t &d1, float (&out)[Q]) {
    float lo[16], hi[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        // only the registers that carry rows below Q
        const bool need = 8 * (r >> 2) + (r & 3) < Q;
        if (!need) { lo[r] = hi[r] = 0.f; continue; }
        const auto s = __builtin_amdgcn_permlane32_swap(__float_as_uint(d0[r]), __float_as_uint(d1[r]), false, false);
        lo[r] = __uint_as_float(s[0]);
        hi[r] = __uint_as_float(s[1]);
    }
#pragma unroll
    for (int k = 0; k < Q; ++k) out[k] = mrt_upper(k) ? hi[mrt_reg(k)] : lo[mrt_reg(k)];
}

template <int Q>
__device__ __forceinline__ void contract_mfma(const float (&x)[Q], const float (*a)[64], float (&out)[Q]) {
    constexpr int NK = (Q + 1) / 2;
    static_assert(NK <= kMrtSteps, "transform operand images cover the lattice");
    const int lane = __lane_id();
    f32x16_t d0 = {}, d1 = {};
#pragma unroll
    for (int s = 0; s < NK; ++s) {
        const float lo = x[2 * s];
        const float hi = 2 * s + 1 < Q ? x[2 * s + 1] : 0.f;
        const auto b = __builtin_amdgcn_permlane32_swap(__float_as_uint(lo), __float_as_uint(hi), false, false);
        const float av = a[s][lane];
        d0 = __builtin_amdgcn_mfma_f32_32x32x2f32(av, __uint_as_float(b[0]), d0, 0, 0, 0);
        d1 = __builtin_amdgcn_mfma_f32_32x32x2f32(av, __uint_as_float(b[1]), d1, 0, 0, 0);
    }
    rows_to_lanes<Q>(d0, d1, out);
}

// m = M f of this lane's node (shifted populations give shifted moments)
template <int Q>
__device__ __forceinline__ void moments_mfma(const MlpOps &o, const float (&f)[Q], float (&m)[Q]) {
    contract_mfma<Q>(f, o.mfwd, m);
}
// f = M^T y of this lane's node
template <int Q>
__device__ __forceinline__ void populations_mfma(const MlpOps &o, const float (&y)[Q], float (&f)[Q]) {
    contract_mfma<Q>(y, o.mbwd, f);
}

// value held by lane l ^ 32 (v_permlane32_swap exchanges the halves of a register pair; unlike
// ds_bpermute it is no LDS instruction, so it never waits on direct-to-LDS loads in flight)
__device__ __forceinline__ float swap_halves(float v) {
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    return __uint_as_float(__lane_id() < 32 ? r[1] : r[0]);
}

// network output (nu_t / nu_scale) of this lane's normalised inputs; o may live in LDS
template <typename OT>
__device__ __forceinline__ float mlp_forward_wave(const MlpOps &o, const float *xn, float b3) {
    typedef typename OT::vec vec;
    typedef typename OT::elem elem;
    const int lane = __lane_id();
    const int h = lane >> 5;
    // layer-1 B operands: column block 0 = nodes 0..31, block 1 = nodes 32..63; lane l carries
    // K slots 8 h .. 8 h + 7 of node l % 32 (block 0) or 32 + l % 32 (block 1); slot 9 = 1 (bias).
    // Slot e takes feature e in the lower half and t_e (feature 8, the bias 1, zeros) in the upper
    // half: v_permlane32_swap(xn[e], t_e) leaves exactly block 0's operand in the first register
    // and block 1's in the second.
    vec x0, x1;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const float t = e == 0 ? xn[8] : (e == 1 ? 1.f : 0.f);
        const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(xn[e]), __float_as_uint(t), false, false);
        x0[e] = (elem)__uint_as_float(r[0]);
        x1[e] = (elem)__uint_as_float(r[1]);
    }
    const f32x16_t zero = {};
    f32x16_t z0, z1;
    {
        const vec a1 = OT::load(o, OT::w1(o, lane));
        z0 = OT::mfma(a1, x0, zero);
        z1 = OT::mfma(a1, x1, zero);
    }
    // SiLU; accumulator r of this lane is hidden unit mfma_row(h, r), which is also the K slot
    // 8 h + r % 8 of layer-2 K step r / 8, so the activations stay in place
    vec g0[2], g1[2];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        g0[r >> 3][r & 7] = (elem)silu_scaled(z0[r]);
        g1[r >> 3][r & 7] = (elem)silu_scaled(z1[r]);
    }
    // layer 2: the accumulators start from the (FP32, scaled) bias of their rows
    f32x16_t y0, y1;
#pragma unroll
    for (int r = 0; r < 16; ++r) y0[r] = y1[r] = o.b2r[h][r];
#pragma unroll
    for (int st = 0; st < 2; ++st) {
        const vec a2 = OT::load(o, OT::w2(o, st, lane));
        y0 = OT::mfma(a2, g0[st], y0);
        y1 = OT::mfma(a2, g1[st], y1);
    }
    float p0 = 0.f, p1 = 0.f;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const float ww = o.w3r[h][r];
        p0 += ww * silu_scaled(y0[r]);
        p1 += ww * silu_scaled(y1[r]);
    }
    // node l % 32 + 32 h sums the two halves of its block: after v_permlane32_swap(p0, p1) the
    // lower lanes hold block 0's halves (p0 of lanes l and l + 32) and the upper lanes block 1's
    const auto pr = __builtin_amdgcn_permlane32_swap(__float_as_uint(p0), __float_as_uint(p1), false, false);
    return softplus_fast(__uint_as_float(pr[0]) + __uint_as_float(pr[1]) + b3);
}

// mlp_tau_t for the FP32 solver, evaluated by the whole wavefront
template <typename OT>
__device__ __forceinline__ float mlp_tau_t_wave(const MlpOps &o, const float *pi, float rho, float drho,
                                                const float *u, float tau0, float cs18) {
    float x[kMlpF], xn[kMlpF];
    closure_features<float>(pi, rho, drho, u, tau0, x);
    bool inside = true;
#pragma unroll
    for (int i = 0; i < kMlpF; ++i) {
        inside = inside && x[i] >= o.fmin[i] && x[i] <= o.fmax[i];
        xn[i] = (x[i] - o.mean[i]) * o.inv_std[i];
    }
    const float net = mlp_forward_wave<OT>(o, xn, o.b3);
    // the Smagorinsky fallback only where some lane of the wavefront needs it
    float smag = 0.f;
    if (__ballot(!inside)) {
        const float q = root(2.f * (pi[0] * pi[0] + pi[1] * pi[1] + pi[2] * pi[2] +
                                    2.f * (pi[3] * pi[3] + pi[4] * pi[4] + pi[5] * pi[5])));
        smag = 0.5f * (root(tau0 * tau0 + cs18 * q) - tau0);
    }
    return inside ? 3.f * o.nu_scale * net : smag;
}

// copy of the operand images into LDS by one workgroup (followed by a barrier)
// Copies into LDS the parts of the operand images a kernel reads: one operand form's network
// images (BF16 w1a, w2a or FP16 h1a, h2a), the scalars, and with MRT the transform's images; the
// other parts of dst stay unwritten and are never read.
template <bool FP16, bool MRT>
__device__ __forceinline__ void stage_mlp_ops(MlpOps &dst, const MlpOps &src) {
    static_assert(sizeof(MlpOps) % 16 == 0 && offsetof(MlpOps, w3r) % 16 == 0 && offsetof(MlpOps, mfwd) % 16 == 0,
                  "operand images are copied in 16-byte units");
    static_assert(offsetof(MlpOps, w2a) == offsetof(MlpOps, w1a) + sizeof(src.w1a) &&
                  offsetof(MlpOps, h2a) == offsetof(MlpOps, h1a) + sizeof(src.h1a), "each form's images are adjacent");
    constexpr int img0 = (int)((FP16 ? offsetof(MlpOps, h1a) : offsetof(MlpOps, w1a)) / 16);
    constexpr int nimg = (int)((sizeof(src.w1a) + sizeof(src.w2a)) / 16);
    constexpr int sc0 = (int)(offsetof(MlpOps, w3r) / 16);
    constexpr int nsc = (int)(((MRT ? sizeof(MlpOps) : offsetof(MlpOps, mfwd)) - offsetof(MlpOps, w3r)) / 16);
    const uint4 *s = reinterpret_cast<const uint4 *>(&src);
    uint4 *d = reinterpret_cast<uint4 *>(&dst);
    for (int i = threadIdx.x; i < nimg + nsc; i += blockDim.x) {
        const int j = i < nimg ? img0 + i : sc0 + (i - nimg);
        d[j] = s[j];
    }
    __syncthreads();
}

// batch inference on the MFMA network (FP32 solver form); every lane runs, inactive ones on zeros
template <typename OT>
__global__ __launch_bounds__(256) void k_mlp_eval_wave(const MlpDev *m, const float *X, int64_t n, float *y) {
    __shared__ MlpOps ops;
    stage_mlp_ops<std::is_same<OT, OperandsF16>::value, false>(ops, m->ops);
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float xn[kMlpF];
    for (int i = 0; i < kMlpF; ++i) xn[i] = k < n ? X[k * kMlpF + i] : 0.f;
    const float out = mlp_forward_wave<OT>(ops, xn, ops.b3);
    if (k < n) y[k] = out;
}

// ---------------------------------------------------------------------------------------------
// training: one thread per sample, forward + backward in FP32, wavefront-reduced gradients
// accumulated with atomics; Adam update over the flat parameter vector
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ float wave_sum_f(float v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ uint32_t hash32(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
    return x;
}

// gradient of mean((softplus(net(x)) - t)^2) over a batch; idx == nullptr uses samples 0..batch-1,
// otherwise sample k of the batch is (hash(seed, k) mod n)
__global__ __launch_bounds__(256) void k_mlp_grad(const float *__restrict__ params, const float *__restrict__ X,
                                                  const float *__restrict__ T, int64_t n, int batch, uint32_t seed,
                                                  int random, float *__restrict__ grad, float *__restrict__ loss) {
    __shared__ float sp[kMlpParams];
    for (int i = threadIdx.x; i < kMlpParams; i += blockDim.x) sp[i] = params[i];
    __syncthreads();
    const float *W1 = sp, *b1 = W1 + kMlpH * kMlpF, *W2 = b1 + kMlpH, *b2 = W2 + kMlpH * kMlpH;
    const float *w3 = b2 + kMlpH;
    const float b3 = w3[kMlpH];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = k < batch;
    int64_t s = 0;
    if (active) s = random ? (int64_t)(hash32(seed * 2654435761u + (uint32_t)k) % (uint64_t)n) : k;
    float x[kMlpF], z1[kMlpH], h1[kMlpH], z2[kMlpH];
    for (int i = 0; i < kMlpF; ++i) x[i] = active ? X[s * kMlpF + i] : 0.f;
    for (int j = 0; j < kMlpH; ++j) {
        float z = b1[j];
        for (int i = 0; i < kMlpF; ++i) z += W1[j * kMlpF + i] * x[i];
        z1[j] = z;
        h1[j] = z * sigmoidf_(z);
    }
    float z3 = b3;
    for (int j = 0; j < kMlpH; ++j) {
        float z = b2[j];
        for (int i = 0; i < kMlpH; ++i) z += W2[j * kMlpH + i] * h1[i];
        z2[j] = z;
        z3 += w3[j] * (z * sigmoidf_(z));
    }
    const float y = softplusf_(z3);
    const float t = active ? T[s] : 0.f;
    const float inv_b = 1.f / (float)batch;
    const float dz3 = active ? 2.f * (y - t) * inv_b * sigmoidf_(z3) : 0.f;
    const float lsum = wave_sum_f(active ? (y - t) * (y - t) * inv_b : 0.f);
    const int lane = threadIdx.x & 63;
    if (lane == 0) atomicAdd(loss, lsum);
    // layer 3
    float *gW1 = grad, *gb1 = gW1 + kMlpH * kMlpF, *gW2 = gb1 + kMlpH, *gb2 = gW2 + kMlpH * kMlpH;
    float *gw3 = gb2 + kMlpH;
    float g = wave_sum_f(dz3);
    if (lane == 0) atomicAdd(&gw3[kMlpH], g);
    float dz2[kMlpH];
    for (int j = 0; j < kMlpH; ++j) {
        const float sg = sigmoidf_(z2[j]);
        const float h2 = z2[j] * sg;
        g = wave_sum_f(dz3 * h2);
        if (lane == 0) atomicAdd(&gw3[j], g);
        dz2[j] = dz3 * w3[j] * sg * (1.f + z2[j] * (1.f - sg));
        g = wave_sum_f(dz2[j]);
        if (lane == 0) atomicAdd(&gb2[j], g);
    }
    for (int j = 0; j < kMlpH; ++j)
        for (int i = 0; i < kMlpH; ++i) {
            g = wave_sum_f(dz2[j] * h1[i]);
            if (lane == 0) atomicAdd(&gW2[j * kMlpH + i], g);
        }
    for (int i = 0; i < kMlpH; ++i) {
        float dh = 0.f;
        for (int j = 0; j < kMlpH; ++j) dh += dz2[j] * W2[j * kMlpH + i];
        const float sg = sigmoidf_(z1[i]);
        const float dz1 = dh * sg * (1.f + z1[i] * (1.f - sg));
        g = wave_sum_f(dz1);
        if (lane == 0) atomicAdd(&gb1[i], g);
        for (int f = 0; f < kMlpF; ++f) {
            g = wave_sum_f(dz1 * x[f]);
            if (lane == 0) atomicAdd(&gW1[i * kMlpF + f], g);
        }
    }
}

// ---------------------------------------------------------------------------------------------
// training on the matrix cores: forward and backward of 64 samples per wavefront as products
// with BF16 operands and FP32 accumulation (v_mfma_f32_32x32x16_bf16).  Same gradient, batch
// selection and loss as k_mlp_grad.  The forward has mlp_forward_wave's layout without the SiLU
// scaling (the parameters change every step, so the operands are rounded here from FP32).  The
// weight gradients contract over samples, which sit on lanes, so the backward turns the layer
// deltas and activations around through LDS:
//   dW2 [32 x 32]        = D2 [32 x 64 samples] . H1^T     4 K steps of 16 samples
//   D1^T [32 x 64]       = W2^T . D2                       D2 kept in accumulator order, as the
//                                                          forward keeps H1 for layer 2
//   [dW1 | db1] [32 x 10] = D1 [32 x 64] . [X | 1]          4 K steps; column 9 is the bias
// Rounded to BF16: X, W1, b1, W2, b2, H1, D2 (as operand of both products) and D1.  db2, dw3,
// db3 and the loss accumulate per lane in FP32 from unrounded values.  A wavefront keeps its
// sums over all the batches it takes in its accumulators; the workgroup adds its four
// wavefronts' sums in LDS and then into grad, one atomic per parameter.
// ---------------------------------------------------------------------------------------------
constexpr int kGradTileStride = 72;   // LDS row stride of the transposed tiles, BF16 elements (144 B)
constexpr int kGradMaxBlocks = 256;

struct alignas(16) GradTiles {
    __bf16 d[kMlpH][kGradTileStride];   // layer delta [hidden][sample]: D2, then D1
    __bf16 h[kMlpH][kGradTileStride];   // layer-1 activations [hidden][sample]
    __bf16 x[16][kGradTileStride];      // inputs [feature][sample]; row 9 = 1, rows 10..15 = 0
};

__device__ __forceinline__ float dsilu(float z, float sg) { return sg * (1.f + z * (1.f - sg)); }

__global__ __launch_bounds__(256) void k_mlp_grad_mfma(const float *__restrict__ params, const float *__restrict__ X,
                                                       const float *__restrict__ T, int64_t n, int batch,
                                                       uint32_t seed, int random, float *__restrict__ grad,
                                                       float *__restrict__ loss) {
    __shared__ float sp[kMlpParams];
    __shared__ float acc[kMlpParams + 1];   // workgroup sums: gradient, then the loss
    __shared__ GradTiles tiles[4];
    const int lane = __lane_id(), h = lane >> 5, col = lane & 31;
    GradTiles &tl = tiles[threadIdx.x >> 6];
    for (int i = threadIdx.x; i < kMlpParams + 1; i += blockDim.x) {
        if (i < kMlpParams) sp[i] = params[i];
        acc[i] = 0.f;
    }
    for (int r = kMlpF; r < 16; ++r) tl.x[r][lane] = (__bf16)(r == kMlpF ? 1.f : 0.f);
    __syncthreads();
    const float *W1 = sp, *b1 = W1 + kMlpH * kMlpF, *W2 = b1 + kMlpH, *b2 = W2 + kMlpH * kMlpH;
    const float *w3 = b2 + kMlpH;
    const float b3 = w3[kMlpH];
    // A operands: row `col` of the weight matrix, K slots 8 h + e (the layout of mlp_forward_wave's operand images)
    bf16x8_t a1, a2[3], at[2];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int k = 8 * h + e;
        a1[e] = (__bf16)(k < kMlpF ? W1[col * kMlpF + k] : k == kMlpF ? b1[col] : 0.f);
#pragma unroll
        for (int st = 0; st < 2; ++st) {
            a2[st][e] = (__bf16)W2[col * kMlpH + mfma_row(h, 8 * st + e)];
            at[st][e] = (__bf16)W2[mfma_row(h, 8 * st + e) * kMlpH + col];   // W2^T
        }
        a2[2][e] = (__bf16)(k == 0 ? b2[col] : 0.f);
    }
    float w3r[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) w3r[r] = w3[mfma_row(h, r)];
    const float inv_b = 1.f / (float)batch;
    const f32x16_t zero = {};
    f32x16_t gw2 = zero, gw1 = zero;
    float gw3[16], gb2[16], gb3 = 0.f, lsum = 0.f;
#pragma unroll
    for (int r = 0; r < 16; ++r) gw3[r] = gb2[r] = 0.f;
    // every wavefront of the workgroup runs the same trips (the barriers need it); lanes past the
    // batch carry zero inputs and a zero output delta, so they add nothing
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < batch; base += (int64_t)gridDim.x * blockDim.x) {
        const int64_t k = base + threadIdx.x;
        const bool active = k < batch;
        int64_t s = 0;
        if (active) s = random ? (int64_t)(hash32(seed * 2654435761u + (uint32_t)k) % (uint64_t)n) : k;
        float xn[kMlpF], px[kMlpF];
#pragma unroll
        for (int i = 0; i < kMlpF; ++i) xn[i] = active ? X[s * kMlpF + i] : 0.f;
        const float t = active ? T[s] : 0.f;
#pragma unroll
        for (int i = 0; i < kMlpF; ++i) px[i] = swap_halves(xn[i]);
        // layer 1: column block c holds samples 32 c .. 32 c + 31 of the wavefront
        bf16x8_t xb[2];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const float hi = e == 1 ? 1.f : 0.f;
            xb[0][e] = (__bf16)(e == 0 ? (h ? px[8] : xn[0]) : (h ? hi : xn[e]));
            xb[1][e] = (__bf16)(e == 0 ? (h ? xn[8] : px[0]) : (h ? hi : px[e]));
        }
        f32x16_t z1[2], z2[2];
        bf16x8_t g[2][3];
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            z1[c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, xb[c], zero, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 16; ++r) g[c][r >> 3][r & 7] = (__bf16)(z1[c][r] * sigmoidf_(z1[c][r]));
#pragma unroll
            for (int e = 0; e < 8; ++e) g[c][2][e] = (__bf16)(h == 0 && e == 0 ? 1.f : 0.f);
            z2[c] = zero;
#pragma unroll
            for (int st = 0; st < 3; ++st) z2[c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a2[st], g[c][st], z2[c], 0, 0, 0);
        }
        // output, loss and its derivative for this lane's sample (lane l holds sample l)
        float p[2] = {0.f, 0.f}, s2[2][16];
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                s2[c][r] = sigmoidf_(z2[c][r]);
                p[c] += w3r[r] * z2[c][r] * s2[c][r];
            }
        const float z3 = (h ? p[1] : p[0]) + swap_halves(h ? p[0] : p[1]) + b3;
        const float y = softplusf_(z3);
        const float dz3 = active ? 2.f * (y - t) * inv_b * sigmoidf_(z3) : 0.f;
        lsum += active ? (y - t) * (y - t) * inv_b : 0.f;
        gb3 += dz3;
        // output delta of the samples of this lane's two columns
        const float other = swap_halves(dz3);
        const float d3[2] = {h ? other : dz3, h ? dz3 : other};
        bf16x8_t bd[2][2];
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = mfma_row(h, r);
                gw3[r] += d3[c] * z2[c][r] * s2[c][r];
                const float dz = d3[c] * w3r[r] * dsilu(z2[c][r], s2[c][r]);
                gb2[r] += dz;
                const __bf16 db = (__bf16)dz;
                bd[c][r >> 3][r & 7] = db;
                tl.d[row][32 * c + col] = db;
                tl.h[row][32 * c + col] = g[c][r >> 3][r & 7];
            }
#pragma unroll
        for (int i = 0; i < kMlpF; ++i) tl.x[i][lane] = (__bf16)xn[i];
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
            const bf16x8_t A = *reinterpret_cast<const bf16x8_t *>(&tl.d[col][16 * ks + 8 * h]);
            const bf16x8_t B = *reinterpret_cast<const bf16x8_t *>(&tl.h[col][16 * ks + 8 * h]);
            gw2 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A, B, gw2, 0, 0, 0);
        }
        f32x16_t e1[2];
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            e1[c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(at[0], bd[c][0], zero, 0, 0, 0);
            e1[c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(at[1], bd[c][1], e1[c], 0, 0, 0);
        }
        __syncthreads();   // D2 read by every lane before D1 takes its place
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int r = 0; r < 16; ++r)
                tl.d[mfma_row(h, r)][32 * c + col] = (__bf16)(e1[c][r] * dsilu(z1[c][r], sigmoidf_(z1[c][r])));
        __syncthreads();
        const int xr = col < 16 ? col : 15;   // columns 16..31 of [X | 1] are zero (row 15)
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
            const bf16x8_t A = *reinterpret_cast<const bf16x8_t *>(&tl.d[col][16 * ks + 8 * h]);
            const bf16x8_t B = *reinterpret_cast<const bf16x8_t *>(&tl.x[xr][16 * ks + 8 * h]);
            gw1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A, B, gw1, 0, 0, 0);
        }
        __syncthreads();   // tiles read before the next batch overwrites them
    }
    float *aW1 = acc, *ab1 = aW1 + kMlpH * kMlpF, *aW2 = ab1 + kMlpH, *ab2 = aW2 + kMlpH * kMlpH;
    float *aw3 = ab2 + kMlpH;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int row = mfma_row(h, r);
        atomicAdd(&aW2[row * kMlpH + col], gw2[r]);
        if (col < kMlpF) atomicAdd(&aW1[row * kMlpF + col], gw1[r]);
        else if (col == kMlpF) atomicAdd(&ab1[row], gw1[r]);
        atomicAdd(&aw3[row], gw3[r]);
        atomicAdd(&ab2[row], gb2[r]);
    }
    gb3 = wave_sum_f(gb3);
    lsum = wave_sum_f(lsum);
    if (lane == 0) {
        atomicAdd(&aw3[kMlpH], gb3);
        atomicAdd(&acc[kMlpParams], lsum);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kMlpParams; i += blockDim.x) atomicAdd(&grad[i], acc[i]);
    if (threadIdx.x == 0) atomicAdd(loss, acc[kMlpParams]);
}

__global__ void k_adam(float *params, const float *grad, float *m, float *v, int n, float lr, float b1, float b2,
                       float eps, int t) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = grad[i];
    const float mi = b1 * m[i] + (1.f - b1) * g;
    const float vi = b2 * v[i] + (1.f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / (1.f - powf(b1, (float)t));
    const float vh = vi / (1.f - powf(b2, (float)t));
    params[i] -= lr * mh / (sqrtf(vh) + eps);
}

// batch inference (normalised inputs) for tests and evaluation
__global__ void k_mlp_eval(const MlpDev *m, const float *X, int64_t n, float *y) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    float xn[kMlpF];
    for (int i = 0; i < kMlpF; ++i) xn[i] = X[k * kMlpF + i];
    y[k] = mlp_forward(*m, xn);
}

}  // namespace zlh
