// mrt_d3q19.hpp -- the D3Q19 MRT transforms m = M f and f = M^T y in factored form.
//
// M is the moment basis of Lattice<3>::moment (d'Humieres et al. 2002, the basis of
// lbm_d3q19.h / lbm_collision_model_mrt.cpp:63-82, 114-143): rows rho, e, eps, j_x, q_x, j_y,
// q_y, j_z, q_z, 3p_xx, 3pi_xx, p_ww, pi_ww, p_xy, p_yz, p_xz, m_x, m_y, m_z.  The direction
// pairs are (1,2) (3,4) (5,6) along the axes and (7,10) (8,9) (11,14) (12,13) (15,18) (16,17)
// on the edges.  Every row is a small-integer combination of a few pair sums (even rows) or
// pair differences (odd rows) that several rows share:
//   face sums F = sX + sY + sZ, edge-plane sums sXY, sXZ, sYZ and their total E;
//   per axis, the edge differences that carry c_a (ex, ey, ez) and their counterparts with the
//   sign of the third-order rows.
// Computed this way the forward transform takes 62 additions and multiply-adds per node and the
// back transform about 60, against about 120 and 80 for row-by-row sums over the pairs.  The
// results equal M f and M^T y up to FP rounding (different association); tests/test_mrt_factored
// checks both against the integer matrix.
#pragma once

#ifndef ZLH_HD
#define ZLH_HD __host__ __device__ __forceinline__
#endif

namespace zlh {

template <typename C>
ZLH_HD void mrt19_moments(const C (&f)[19], C (&m)[19]) {
    // pair sums and differences, each pair as (direction, opposite)
    const C sX = f[1] + f[2], sY = f[3] + f[4], sZ = f[5] + f[6];
    const C dX = f[1] - f[2], dY = f[3] - f[4], dZ = f[5] - f[6];
    const C sXYp = f[7] + f[10], sXYm = f[8] + f[9], sXZp = f[11] + f[14], sXZm = f[12] + f[13];
    const C sYZp = f[15] + f[18], sYZm = f[16] + f[17];
    const C dXYp = f[7] - f[10], dXYm = f[8] - f[9], dXZp = f[11] - f[14], dXZm = f[12] - f[13];
    const C dYZp = f[15] - f[18], dYZm = f[16] - f[17];
    // even rows
    const C F = sX + sY + sZ;
    const C sXY = sXYp + sXYm, sXZ = sXZp + sXZm, sYZ = sYZp + sYZm;
    const C E = sXY + sXZ + sYZ;
    m[0] = f[0] + F + E;
    m[1] = C(-30) * f[0] + C(-11) * F + C(8) * E;
    m[2] = C(12) * f[0] + C(-4) * F + E;
    const C a9 = C(3) * sX - F, b9 = E - C(3) * sYZ;
    m[9] = a9 + b9;
    m[10] = C(-2) * a9 + b9;
    const C a11 = sY - sZ, b11 = sXY - sXZ;
    m[11] = a11 + b11;
    m[12] = C(-2) * a11 + b11;
    m[13] = sXYp - sXYm;
    m[14] = sYZp - sYZm;
    m[15] = sXZp - sXZm;
    // odd rows: the edge pairs' contributions along x, y, z and their third-order counterparts
    const C gxy = dXYp - dXYm, gxz = dXZp - dXZm;        // x components of the xy / xz edges
    const C hxy = dXYp + dXYm, hyz = dYZp - dYZm;        // y components of the xy / yz edges
    const C kxz = dXZp + dXZm, kyz = dYZp + dYZm;        // z components of the xz / yz edges
    const C ex = gxy + gxz, ey = hxy + hyz, ez = kxz + kyz;
    m[3] = dX + ex;
    m[4] = C(-4) * dX + ex;
    m[5] = dY + ey;
    m[6] = C(-4) * dY + ey;
    m[7] = dZ + ez;
    m[8] = C(-4) * dZ + ez;
    m[16] = gxy - gxz;
    m[17] = hyz - hxy;
    m[18] = kxz - kyz;
}

// f = M^T y (y: the post-collision moments divided by |M_k|^2)
template <typename C>
ZLH_HD void mrt19_populations(const C (&y)[19], C (&f)[19]) {
    f[0] = y[0] + C(-30) * y[1] + C(12) * y[2];
    // faces: c^2 = 1
    const C A = y[0] + C(-11) * y[1] + C(-4) * y[2];
    const C P = y[9] + C(-2) * y[10];                   // 3p_xx / 3pi_xx of a face along x
    const C R = y[11] + C(-2) * y[12];                  // p_ww / pi_ww of a face along y
    const C eX = A + C(2) * P, Bm = A - P;
    const C eY = Bm + R, eZ = Bm - R;
    const C oX = y[3] + C(-4) * y[4], oY = y[5] + C(-4) * y[6], oZ = y[7] + C(-4) * y[8];
    f[1] = eX + oX;
    f[2] = eX - oX;
    f[3] = eY + oY;
    f[4] = eY - oY;
    f[5] = eZ + oZ;
    f[6] = eZ - oZ;
    // edges: c^2 = 2
    const C G = y[0] + C(8) * y[1] + y[2];
    const C Q9 = y[9] + y[10], Q11 = y[11] + y[12];
    const C GQ9 = G + Q9;
    const C Hxy = GQ9 + Q11, Hxz = GQ9 - Q11, Hyz = G + C(-2) * Q9;
    const C Jx = y[3] + y[4], Jy = y[5] + y[6], Jz = y[7] + y[8];
    // xy plane: 7 = (1,1,0), 8 = (-1,1,0)
    {
        const C e7 = Hxy + y[13], e8 = Hxy - y[13];
        const C U = Jx + y[16], V = Jy - y[17];
        const C o7 = U + V, o8 = V - U;
        f[7] = e7 + o7;
        f[10] = e7 - o7;
        f[8] = e8 + o8;
        f[9] = e8 - o8;
    }
    // xz plane: 11 = (1,0,1), 12 = (-1,0,1)
    {
        const C e11 = Hxz + y[15], e12 = Hxz - y[15];
        const C U = Jx - y[16], V = Jz + y[18];
        const C o11 = U + V, o12 = V - U;
        f[11] = e11 + o11;
        f[14] = e11 - o11;
        f[12] = e12 + o12;
        f[13] = e12 - o12;
    }
    // yz plane: 15 = (0,1,1), 16 = (0,-1,1)
    {
        const C e15 = Hyz + y[14], e16 = Hyz - y[14];
        const C U = Jy + y[17], V = Jz - y[18];
        const C o15 = U + V, o16 = V - U;
        f[15] = e15 + o15;
        f[18] = e15 - o15;
        f[16] = e16 + o16;
        f[17] = e16 - o16;
    }
}

}  // namespace zlh
