// lattice.h -- D2Q9 / D3Q19 constants for the device kernels.
//
// Direction order of the reference (lbm_d2q9.h:53-61, lbm_d3q19.h:18-39,73-85); the MRT moment
// basis is generated from its polynomial rows (the d'Humieres basis of kMatrixMMrt_,
// lbm_d2q9.h:70-80, lbm_d3q19.h:98-118).  The rows are mutually orthogonal, so the inverse
// transform is M^T diag(1/|row|^2), which keeps both transforms as small-integer sums.
#pragma once
#include <cstdint>

namespace zlh {

template <int D> struct Lattice;

template <> struct Lattice<2> {
    static constexpr int Q = 9;
    static constexpr int cx[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
    static constexpr int cy[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
    static constexpr int cz[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    static constexpr double w[9] = {4. / 9., 1. / 9., 1. / 9., 1. / 9., 1. / 9.,
                                    1. / 36., 1. / 36., 1. / 36., 1. / 36.};
    // base relaxation rates (k0VectorSMrt_, lbm_d2q9.h:92); viscous rows set per node
    static constexpr double s_base[9] = {0., 1.4, 1.4, 0., 1.2, 0., 1.2, 0., 0.};
    static constexpr bool viscous_row[9] = {false, false, false, false, false, false, false, true, true};
    static constexpr int moment(int k, int i) {
        int x = cx[i], y = cy[i];
        int c2 = x * x + y * y;
        switch (k) {
            case 0: return 1;
            case 1: return 3 * c2 - 4;
            case 2: return (9 * c2 * c2 - 21 * c2) / 2 + 4;
            case 3: return x;
            case 4: return (3 * c2 - 5) * x;
            case 5: return y;
            case 6: return (3 * c2 - 5) * y;
            case 7: return x * x - y * y;
            default: return x * y;
        }
    }
};

template <> struct Lattice<3> {
    static constexpr int Q = 19;
    static constexpr int cx[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
    static constexpr int cy[19] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
    static constexpr int cz[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
    static constexpr double w[19] = {1. / 3., 1. / 18., 1. / 18., 1. / 18., 1. / 18., 1. / 18., 1. / 18.,
                                     1. / 36., 1. / 36., 1. / 36., 1. / 36., 1. / 36., 1. / 36.,
                                     1. / 36., 1. / 36., 1. / 36., 1. / 36., 1. / 36., 1. / 36.};
    // base relaxation rates (k0VectorSMrt_, lbm_d3q19.h:159-160); viscous rows 9,11,13,14,15
    static constexpr double s_base[19] = {0., 1.5, 1.4, 0., 1.2, 0., 1.2, 0., 1.2, 0., 1.4, 0., 1.4,
                                          0., 0., 0., 1.98, 1.98, 1.98};
    static constexpr bool viscous_row[19] = {false, false, false, false, false, false, false, false,
                                             false, true, false, true, false, true, true, true,
                                             false, false, false};
    static constexpr int moment(int k, int i) {
        int x = cx[i], y = cy[i], z = cz[i];
        int c2 = x * x + y * y + z * z;
        switch (k) {
            case 0: return 1;
            case 1: return 19 * c2 - 30;
            case 2: return (21 * c2 * c2 - 53 * c2 + 24) / 2;
            case 3: return x;
            case 4: return (5 * c2 - 9) * x;
            case 5: return y;
            case 6: return (5 * c2 - 9) * y;
            case 7: return z;
            case 8: return (5 * c2 - 9) * z;
            case 9: return 3 * x * x - c2;
            case 10: return (3 * c2 - 5) * (3 * x * x - c2);
            case 11: return y * y - z * z;
            case 12: return (3 * c2 - 5) * (y * y - z * z);
            case 13: return x * y;
            case 14: return y * z;
            case 15: return x * z;
            case 16: return (y * y - z * z) * x;
            case 17: return (z * z - x * x) * y;
            default: return (x * x - y * y) * z;
        }
    }
};

template <int D>
constexpr int moment_norm2(int k) {
    int s = 0;
    for (int i = 0; i < Lattice<D>::Q; ++i) s += Lattice<D>::moment(k, i) * Lattice<D>::moment(k, i);
    return s;
}

template <int D>
constexpr int opposite(int q) {
    for (int j = 0; j < Lattice<D>::Q; ++j)
        if (Lattice<D>::cx[j] == -Lattice<D>::cx[q] && Lattice<D>::cy[j] == -Lattice<D>::cy[q] &&
            Lattice<D>::cz[j] == -Lattice<D>::cz[q])
            return j;
    return -1;
}

// Moment-space tables of the MRT basis (all compile-time):
//   pairs (rep, opp) of opposite directions and the parity of each row under c -> -c (even rows
//   see only f_q + f_opp, odd rows only f_q - f_opp);
//   equilibrium moments of feq = w rho (1 + 3 c.u + 4.5 (c.u)^2 - 1.5 u^2):
//     M feq = rho (a + b.u + u.C.u), a = M w, b_a = 3 M (w c_a), g_ab = 4.5 M (w c_a c_b),
//     C_ab = g_ab - 1.5 delta_ab a;
//   Guo forcing moments M F_i = b.F - 3 (u.F) a + 2 u.g.F;
//   second moments of a population vector from its moments: sum_q c_a c_b f_q = sum_k pi_ab,k m_k
//     (pi_ab,k = sum_q c_a c_b M_kq / |M_k|^2, as M^-1 = M^T diag(1/|M_k|^2)).
template <int D>
struct MomentSpace {
    using L = Lattice<D>;
    static constexpr int Q = L::Q;
    static constexpr int NP = (Q - 1) / 2;
    int rep[NP], opp[NP];
    bool even[Q];
    int m[Q][Q];
    int n2[Q];
    double a[Q], b[Q][3], g[Q][3][3], pi[6][Q];
    constexpr MomentSpace() : rep(), opp(), even(), m(), n2(), a(), b(), g(), pi() {
        int np = 0;
        for (int q = 1; q < Q; ++q) {
            const int o = opposite<D>(q);
            if (q < o) { rep[np] = q; opp[np] = o; ++np; }
        }
        const int ab[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};
        for (int k = 0; k < Q; ++k) {
            int s = 0;
            even[k] = true;
            for (int q = 0; q < Q; ++q) {
                m[k][q] = L::moment(k, q);
                s += m[k][q] * m[k][q];
                if (L::moment(k, opposite<D>(q)) != m[k][q]) even[k] = false;
            }
            n2[k] = s;
            for (int q = 0; q < Q; ++q) {
                const int c[3] = {L::cx[q], L::cy[q], L::cz[q]};
                const double mw = m[k][q] * L::w[q];
                a[k] += mw;
                for (int x = 0; x < 3; ++x) {
                    b[k][x] += 3. * mw * c[x];
                    for (int y = 0; y < 3; ++y) g[k][x][y] += 4.5 * mw * c[x] * c[y];
                }
                for (int j = 0; j < 6; ++j) pi[j][k] += double(c[ab[j][0]] * c[ab[j][1]] * m[k][q]);
            }
            for (int j = 0; j < 6; ++j) pi[j][k] /= double(s);
        }
    }
};

// brick geometry: 4x4x4 (3D) or 8x8 (2D) nodes = 64 lanes = one wavefront, lane = low 6 Morton bits
template <int D> struct Brick;
template <> struct Brick<2> {
    static constexpr int edge = 8;
    static constexpr int nnbr = 9;
    __host__ __device__ static constexpr int lane_coord(int lane, int axis) {
        return ((lane >> axis) & 1) | (((lane >> (2 + axis)) & 1) << 1) | (((lane >> (4 + axis)) & 1) << 2);
    }
    __host__ __device__ static constexpr int lane_of(int x, int y, int z) {
        return (x & 1) | ((y & 1) << 1) | (((x >> 1) & 1) << 2) | (((y >> 1) & 1) << 3) |
               (((x >> 2) & 1) << 4) | (((y >> 2) & 1) << 5);
    }
};
template <> struct Brick<3> {
    static constexpr int edge = 4;
    static constexpr int nnbr = 27;
    __host__ __device__ static constexpr int lane_coord(int lane, int axis) {
        return ((lane >> axis) & 1) | (((lane >> (3 + axis)) & 1) << 1);
    }
    __host__ __device__ static constexpr int lane_of(int x, int y, int z) {
        return (x & 1) | ((y & 1) << 1) | ((z & 1) << 2) | (((x >> 1) & 1) << 3) |
               (((y >> 1) & 1) << 4) | (((z >> 1) & 1) << 5);
    }
};

// The lane bits of one axis ("dilated" coordinate: the axis' coordinate bits at their interleaved
// positions) and their step by one along that axis, modulo the brick edge, without
// de-interleaving: a borrow or carry passes through the other axes' bits when they are forced to
// 0 (decrement) or 1 (increment).  Everything is scaled by 4 (byte offsets of 4-byte lanes).
template <int D>
__host__ __device__ constexpr uint32_t lane_axis_mask4(int axis) {
    uint32_t m = 0;
    for (int bit = 0; bit < 6; ++bit)
        if (axis < D && Brick<D>::lane_coord(1 << bit, axis) != 0) m |= 1u << bit;
    return 4u * m;
}
// 4 x the lane bits of the coordinate c - 1 (step < 0) or c + 1 (step > 0) along the axis whose
// mask is m4, from own4 = 4 x (the lane's bits of that axis)
__host__ __device__ inline uint32_t lane_axis_step4(uint32_t own4, uint32_t m4, int step) {
    return step < 0 ? ((own4 - 4u) & m4) : (((own4 | ~m4) + 4u) & m4);
}

}  // namespace zlh
