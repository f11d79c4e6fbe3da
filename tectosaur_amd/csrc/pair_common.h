// Device code shared by the pair integrators (pairs.hip, interior.hip):
// the source/observation triangle record, its set-up with rotation and
// flip, and the 3x3 kernel tensors.
#pragma once

#include "tct_common.h"

namespace tct {

struct Tri {
    double v[3][3];              // vertices in the rotated frame
    double n[3];                 // unit normal
    double jac;                  // |unscaled normal| = 2 * area
    double curl[3][3];           // surface curl of the basis functions
};

__device__ __forceinline__ int pmod3(int i) { return ((i % 3) + 3) % 3; }

// tri_info of tectosaur/integral_primitives.cl:56-93 (+ surf_curl_basis
// :95-112): vertex c of the rotated triangle is original vertex
// (flip(c) + clicks) mod 3; the flip also negates the unscaled normal.
__device__ __forceinline__ void load_tri(Tri &t, const double *__restrict__ pts,
                                         const int32_t *__restrict__ tris, int idx, int clicks,
                                         int flip) {
    double v[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        int fc = c;
        if (flip) fc = (c == 0) ? 1 : (c == 1 ? 0 : 2);
        const int p = tris[3 * idx + pmod3(fc + clicks)];
#pragma unroll
        for (int d = 0; d < 3; ++d) v[c][d] = pts[3 * p + d];
    }
    double s20[3], s21[3], un[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        s20[d] = v[2][d] - v[0][d];
        s21[d] = v[2][d] - v[1][d];
    }
    un[0] = s20[1] * s21[2] - s20[2] * s21[1];
    un[1] = s20[2] * s21[0] - s20[0] * s21[2];
    un[2] = s20[0] * s21[1] - s20[1] * s21[0];
    if (flip)
#pragma unroll
        for (int d = 0; d < 3; ++d) un[d] = -un[d];
    t.jac = sqrt(un[0] * un[0] + un[1] * un[1] + un[2] * un[2]);
    const double inv = 1.0 / t.jac;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        t.n[d] = un[d] * inv;
#pragma unroll
        for (int c = 0; c < 3; ++c) t.v[c][d] = v[c][d];
    }
    // grad of (1-x-y, x, y) = (-1,-1), (1,0), (0,1); curl_b = (g_b0 e2 - g_b1 e1) / J
    // with e1 = v1 - v0, e2 = v2 - v0.
#pragma unroll
    for (int s = 0; s < 3; ++s) {
        const double e1 = v[1][s] - v[0][s], e2 = v[2][s] - v[0][s];
        t.curl[0][s] = (-e2 + e1) * inv;
        t.curl[1][s] = e2 * inv;
        t.curl[2][s] = -e1 * inv;
    }
}

// Number of accumulators per kernel class.
template <int KID>
constexpr int n_acc() {
    return KID <= TCT_H ? 81 : (KID == TCT_RH ? 7 : 30);
}

// 3x3 kernel tensor (tectosaur/kernels/__init__.py tensor_code), m = obs
// normal, n = src normal, D = y - x.
// invr = 1/r, or 0 to drop the point (every entry is then 0).
template <int KID>
__device__ __forceinline__ void tensor_invr(const double D[3], double invr, const double m[3],
                                            const double n[3], const ElasticConsts &c,
                                            double K[9]) {
    const double invr2 = invr * invr;
    const double invr3 = invr2 * invr;
    if constexpr (KID == TCT_U) {
        const double Q1 = c.CsU0 * invr, Q2 = c.CsU1 * invr3;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) K[i * 3 + j] = Q2 * D[i] * D[j] + (i == j ? Q1 : 0.0);
    } else if constexpr (KID == TCT_T || KID == TCT_A) {
        // K = A(-(N.D) I + s W) + C (N.D) D D^T with W the cross-product
        // matrix of N x D (sign s = +1 for T, -1 for A).
        const double *N = KID == TCT_T ? n : m;
        const double s = KID == TCT_T ? 1.0 : -1.0;
        const double rn = N[0] * D[0] + N[1] * D[1] + N[2] * D[2];
        const double A = s * c.CsT0 * invr3;
        const double C = -s * c.CsT1 * invr3 * invr2;
        const double cx = N[1] * D[2] - N[2] * D[1];
        const double cy = N[2] * D[0] - N[0] * D[2];
        const double cz = N[0] * D[1] - N[1] * D[0];
        const double As = A * s;
        K[0] = -A * rn + C * rn * D[0] * D[0];
        K[1] = -As * cz + C * rn * D[0] * D[1];
        K[2] = As * cy + C * rn * D[0] * D[2];
        K[3] = As * cz + C * rn * D[1] * D[0];
        K[4] = -A * rn + C * rn * D[1] * D[1];
        K[5] = -As * cx + C * rn * D[1] * D[2];
        K[6] = -As * cy + C * rn * D[2] * D[0];
        K[7] = As * cx + C * rn * D[2] * D[1];
        K[8] = -A * rn + C * rn * D[2] * D[2];
    } else {
        const double Dr[3] = {D[0] * invr, D[1] * invr, D[2] * invr};
        const double rn = n[0] * Dr[0] + n[1] * Dr[1] + n[2] * Dr[2];
        const double rm = m[0] * Dr[0] + m[1] * Dr[1] + m[2] * Dr[2];
        const double mn = m[0] * n[0] + m[1] * n[1] + m[2] * n[2];
        const double Q = c.CsH0 * invr3;
        const double A = Q * 3.0 * rn;
        const double B = Q * c.CsH1;
        const double C = Q * c.CsH3;
        double MT[3], NT[3], DT[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            MT[k] = Q * c.CsH2 * n[k] + A * c.CsH1 * Dr[k];
            NT[k] = B * m[k] + C * Dr[k] * rm;
            DT[k] = B * 3.0 * n[k] * rm + C * Dr[k] * mn + A * (c.nu * m[k] - 5.0 * Dr[k] * rm);
        }
        const double ST = A * c.nu * rm + B * mn;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                K[i * 3 + j] = n[i] * NT[j] + m[i] * MT[j] + Dr[i] * DT[j] + (i == j ? ST : 0.0);
    }
}

template <int KID>
__device__ __forceinline__ void tensor(const double D[3], double r2, const double m[3],
                                       const double n[3], const ElasticConsts &c,
                                       double K[9]) {
    tensor_invr<KID>(D, rsqrt(r2), m, n, c, K);
}

// Symmetric index of (p, t) in 6-vector order xx, xy, xz, yy, yz, zz.
__device__ __forceinline__ constexpr int sym6(int p, int t) {
    return p <= t ? (p == 0 ? t : (p == 1 ? 2 + t : 5)) : (t == 0 ? p : (t == 1 ? 2 + p : 5));
}

// Levi-Civita symbol of (i, j, k) in {0,1,2}^3.
__device__ __forceinline__ int levi(int i, int j, int k) { return (i - j) * (j - k) * (k - i) / 2; }

}  // namespace tct
