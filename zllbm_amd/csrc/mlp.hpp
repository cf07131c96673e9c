// mlp.hpp -- learned sub-grid closure: a 9 -> 32 -> 32 -> 1 MLP (SiLU, softplus) predicting the
// eddy viscosity from local non-equilibrium invariants.  No counterpart in the reference; it
// replaces the Smagorinsky relaxation time (les_models/lbm_les_smagorinsky.cpp:22-94) where its
// inputs lie inside the training range and falls back to Smagorinsky elsewhere.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace zlh {

constexpr int kMlpF = 9;
constexpr int kMlpH = 32;
constexpr int kMlpParams = kMlpH * kMlpF + kMlpH + kMlpH * kMlpH + kMlpH + kMlpH + 1;   // 1409

// device-resident closure: parameters (flat order W1, b1, W2, b2, w3, b3), input normalisation,
// training range of the raw features and the output scale
struct MlpDev {
    float p[kMlpParams];
    float mean[kMlpF];
    float inv_std[kMlpF];
    float fmin[kMlpF];
    float fmax[kMlpF];
    float nu_scale;
};

__device__ __forceinline__ float sigmoidf_(float z) { return 1.f / (1.f + __expf(-z)); }
__device__ __forceinline__ float softplusf_(float z) { return z > 20.f ? z : log1pf(__expf(z)); }

// raw closure inputs from the non-equilibrium stress Pi (xx, yy, zz, xy, xz, yz), density and
// velocity: P = Pi / rho, a = sqrt(2 P:P), I1 = tr P, I2 = (I1^2 - P:P) / 2, I3 = det P, |u|^2,
// rho - 1, tau0, a tau0, a^2
template <typename C>
__device__ __forceinline__ void closure_features(const C *pi, C rho, C drho, const C *u, C tau0, float *x) {
    const C ir = C(1) / rho;
    const C pxx = pi[0] * ir, pyy = pi[1] * ir, pzz = pi[2] * ir, pxy = pi[3] * ir, pxz = pi[4] * ir, pyz = pi[5] * ir;
    const C pp = pxx * pxx + pyy * pyy + pzz * pzz + C(2) * (pxy * pxy + pxz * pxz + pyz * pyz);
    const C a = sqrt(C(2) * pp);
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
        const C q = sqrt(C(2) * (pi[0] * pi[0] + pi[1] * pi[1] + pi[2] * pi[2] +
                                 C(2) * (pi[3] * pi[3] + pi[4] * pi[4] + pi[5] * pi[5])));
        return C(0.5) * (sqrt(tau0 * tau0 + cs18 * q) - tau0);
    }
    *fallback = false;
    float xn[kMlpF];
#pragma unroll
    for (int i = 0; i < kMlpF; ++i) xn[i] = (x[i] - m.mean[i]) * m.inv_std[i];
    return C(3) * C(m.nu_scale * mlp_forward(m, xn));
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
