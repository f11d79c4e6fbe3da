// Shared pieces of the farfield kernels (direct N-body and treecode):
// kernel constants, the masked reciprocal square root, source-record
// sizes and the per-pair kernel bodies.
#pragma once
#include "tct_common.h"

namespace tct {

constexpr int G = TCT_GROUP;  // points per origin group / source tile
constexpr int R = TCT_REC;    // words per source record

// Kernel constants after folding the leading constant into the strengths.
// eps2: pairs with r2 <= eps2 are skipped.  The FP64 path uses 0 (the
// reference's exact r2 == 0 test).  The FP32 path uses a bound on the
// rounding of a point stored relative to two different group origins, so a
// point paired with itself is skipped even when obs and src are packed
// differently; real point spacings are many orders of magnitude larger.
template <typename T>
struct KConst {
    T c0, c1, c2, c3, c4, eps2;
};

// Without the test, for tiles where no pair can reach r2 <= eps2.
__device__ __forceinline__ float rsq_plain(float r2) { return __builtin_amdgcn_rsqf(r2); }
__device__ __forceinline__ double rsq_plain(double r2) { return rsq_pos(r2); }

__device__ __forceinline__ float rsq_masked(float r2, float eps2) {
    float v = __builtin_amdgcn_rsqf(r2);
    return r2 > eps2 ? v : 0.0f;
}
__device__ __forceinline__ double rsq_masked(double r2, double eps2) {
    double v = rsq_pos(r2);
    return r2 > eps2 ? v : 0.0;
}

template <int KID>
constexpr int rec_words() { return R; }
template <int KID>
constexpr int n_out() { return (KID == TCT_RA || KID == TCT_RH) ? 9 : 3; }
template <int KID>
constexpr bool needs_obs_normal() { return KID == TCT_A || KID == TCT_H || KID == TCT_RA; }
// Sums that are mapped to the outputs once per source tile (fold_mterm)
// rather than per pair, as per-lane scalars (am): terms whose vector is the
// observation normal m, multiplied by m once per tile -- A has one
// (-CsT0/CsT1 (D.s)/r^3), H two (e0 / r^3 and 3 CsH1 (n.D)(s.D)/r^5) -- and
// RH's Phi = sum G / r (9), to which the linear map P(G) = CsRH1 tr(G) I +
// CsRH2 G^T - 2 G is applied once per tile, so the source record carries G
// only.
template <int KID>
constexpr int n_mterm() { return KID == TCT_H ? 2 : (KID == TCT_A ? 1 : (KID == TCT_RH ? 9 : 0)); }
template <int KID>
constexpr int n_mterm_arr() { return n_mterm<KID>() > 0 ? n_mterm<KID>() : 1; }

// Fused multiply-add in the precision of T.
__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }

// a . b for 3-vectors: one multiply and two FMAs, always in this order.
template <typename T>
__device__ __forceinline__ T dot3(T a0, T a1, T a2, T b0, T b1, T b2) {
#pragma clang fp contract(off)
    return fma_t(a2, b2, fma_t(a1, b1, a0 * b0));
}

// One (observation point, source point) pair, accumulate into acc[n_out].
// x: observation point (already shifted into the source tile frame);
// m: observation unit normal; q: the source record (wave-uniform).
//   U/T/A/H: q = [y n s e0 e1 -], s the strength scaled by the kernel
//            constant, e0/e1 per-source dot products;
//   RT:      q = [y n s e], s = CsRT2 w J in(q), e = CsRT1 w J a with
//            a_t = sum_{d,I} eps_{I t d} S_{dI} (S the source surface curl);
//   RA:      q = [y - s -], s = w J in(q);
//   RH:      q = [y G(9)], G = w J CsRH0 S (G[j][m] at q[3 + 3 j + m]).
// The FP32 N-body's pipelined loop copies only the leading words each body
// reads (used_words in far_p2p.hip: U/A/RA 9, T 10, H 11, RT 12, RH 12); a
// body that reads a later word must raise that count.
// Every fused multiply-add is written out (fma_t) and the compiler may form
// no other (contract(off)): left to itself it fused a + b*c + d*e on one
// product in some unrolled copies and on the other in others, so a point's
// sum depended on the launch shape.  Written this way every copy performs
// the same operations, and the results depend only on the source chunks.
// Every accumulated term is its own FMA: a sum of products added to an
// accumulator would cost a multiply and an add more per component.
template <int KID, typename T, bool MASK = true>
__device__ __forceinline__ void pair_eval(const T x[3], const T m[3], const T *__restrict__ q,
                                          const KConst<T> &k, T *acc, T *am) {
#pragma clang fp contract(off)
    const T D0 = q[0] - x[0], D1 = q[1] - x[1], D2 = q[2] - x[2];
    const T r2 = dot3(D0, D1, D2, D0, D1, D2);
    const T invr = MASK ? rsq_masked(r2, k.eps2) : rsq_plain(r2);
    const T w = invr * invr;
    if constexpr (KID == TCT_U) {
        // U: (3-4nu)/r s + (D.s) D / r^3, strengths scaled by CsU1.
        const T s0 = q[6], s1 = q[7], s2 = q[8];
        const T ds = dot3(D0, D1, D2, s0, s1, s2);
        const T a = k.c0 * invr;
        const T b = (invr * w) * ds;
        acc[0] = fma_t(a, s0, acc[0]);
        acc[0] = fma_t(b, D0, acc[0]);
        acc[1] = fma_t(a, s1, acc[1]);
        acc[1] = fma_t(b, D1, acc[1]);
        acc[2] = fma_t(a, s2, acc[2]);
        acc[2] = fma_t(b, D2, acc[2]);
    } else if constexpr (KID == TCT_T) {
        // T: A(-(n.D) s + (n x D) x s) - C (n.D)(D.s) D with
        // (n x D) x s = D (n.s) - n (D.s); strengths scaled by CsT1,
        // k.c0 = CsT0 / CsT1, e0 = n.s.
        const T n0 = q[3], n1 = q[4], n2 = q[5];
        const T s0 = q[6], s1 = q[7], s2 = q[8];
        const T q3 = invr * w;
        const T al = k.c0 * q3;
        const T rn = dot3(n0, n1, n2, D0, D1, D2);
        const T ds = dot3(D0, D1, D2, s0, s1, s2);
        const T be = -al * rn;
        const T ga = -al * ds;
        const T de = fma_t(al, q[9], -((q3 * w) * rn) * ds);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const T sd = d == 0 ? s0 : (d == 1 ? s1 : s2);
            const T nd = d == 0 ? n0 : (d == 1 ? n1 : n2);
            const T Dd = d == 0 ? D0 : (d == 1 ? D1 : D2);
            acc[d] = fma_t(be, sd, acc[d]);
            acc[d] = fma_t(ga, nd, acc[d]);
            acc[d] = fma_t(de, Dd, acc[d]);
        }
    } else if constexpr (KID == TCT_A) {
        // A: the adjoint form with the observation normal m.
        const T s0 = q[6], s1 = q[7], s2 = q[8];
        const T q3 = invr * w;
        const T al = k.c0 * q3;
        const T rn = dot3(m[0], m[1], m[2], D0, D1, D2);
        const T ds = dot3(D0, D1, D2, s0, s1, s2);
        const T sm = dot3(m[0], m[1], m[2], s0, s1, s2);
        const T be = al * rn;
        const T de = fma_t(al, sm, ((q3 * w) * rn) * ds);
        acc[0] = fma_t(be, s0, acc[0]);
        acc[0] = fma_t(de, D0, acc[0]);
        acc[1] = fma_t(be, s1, acc[1]);
        acc[1] = fma_t(de, D1, acc[1]);
        acc[2] = fma_t(be, s2, acc[2]);
        acc[2] = fma_t(de, D2, acc[2]);
        am[0] = fma_t(q3, ds, am[0]);   // m term: -al (D.s) m
    } else if constexpr (KID == TCT_H) {
        // H with strengths scaled by CsH0; k = (CsH1, CsH3, 3 CsH1, nu, 3 nu),
        // e0 = CsH2 (n.s), e1 = 3 CsH1 (n.s).
        const T n0 = q[3], n1 = q[4], n2 = q[5];
        const T s0 = q[6], s1 = q[7], s2 = q[8];
        const T Q = invr * w;
        const T a = dot3(n0, n1, n2, D0, D1, D2);
        const T b = dot3(m[0], m[1], m[2], D0, D1, D2);
        const T c = dot3(s0, s1, s2, D0, D1, D2);
        const T sm = dot3(s0, s1, s2, m[0], m[1], m[2]);
        const T mn = dot3(m[0], m[1], m[2], n0, n1, n2);
        const T wa = w * a;
        const T wbc = (w * b) * c;
        // m term Q (e0 + c2 wa c): two lane scalars, folded per tile
        am[0] = fma_t(Q, q[9], am[0]);
        am[1] = fma_t(Q, wa * c, am[1]);
        const T NT = Q * fma_t(k.c0, sm, k.c1 * wbc);
        // 3 (c3 sm - 5 wbc) with k.c4 = 3 c3
        const T inner = fma_t(k.c4, sm, T(-15) * wbc);
        const T DT = (Q * w) * fma_t(a, inner, fma_t(c, k.c1 * mn, q[10] * b));
        const T ST = Q * fma_t(k.c4 * wa, b, k.c0 * mn);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const T sd = d == 0 ? s0 : (d == 1 ? s1 : s2);
            const T nd = d == 0 ? n0 : (d == 1 ? n1 : n2);
            const T Dd = d == 0 ? D0 : (d == 1 ? D1 : D2);
            acc[d] = fma_t(nd, NT, acc[d]);
            acc[d] = fma_t(Dd, DT, acc[d]);
            acc[d] = fma_t(sd, ST, acc[d]);
        }
    } else if constexpr (KID == TCT_RT) {
        // RT: invr e + (CsRT0/CsRT1) invr^3 (D.e) D + invr^3 (n.D) s
        const T i3 = invr * w;
        const T De = dot3(D0, D1, D2, q[9], q[10], q[11]);
        const T nD = dot3(D0, D1, D2, q[3], q[4], q[5]);
        const T cd = (k.c0 * i3) * De;
        const T cs = i3 * nD;
        const T Dk[3] = {D0, D1, D2};
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            acc[d] = fma_t(invr, q[9 + d], acc[d]);
            acc[d] = fma_t(cd, Dk[d], acc[d]);
            acc[d] = fma_t(cs, q[6 + d], acc[d]);
        }
    } else if constexpr (KID == TCT_RA) {
        // RA moments: v1 += s / r, v2 += (D.s) D / r^3, v3 += (m.D) s / r^3;
        // the constants and the observation curls are applied by the scatter.
        const T s0 = q[6], s1 = q[7], s2 = q[8];
        const T i3 = invr * w;
        const T Ds = i3 * dot3(D0, D1, D2, s0, s1, s2);
        const T mD = i3 * dot3(D0, D1, D2, m[0], m[1], m[2]);
        acc[0] = fma_t(invr, s0, acc[0]);
        acc[1] = fma_t(invr, s1, acc[1]);
        acc[2] = fma_t(invr, s2, acc[2]);
        acc[3] = fma_t(Ds, D0, acc[3]);
        acc[4] = fma_t(Ds, D1, acc[4]);
        acc[5] = fma_t(Ds, D2, acc[5]);
        acc[6] = fma_t(mD, s0, acc[6]);
        acc[7] = fma_t(mD, s1, acc[7]);
        acc[8] = fma_t(mD, s2, acc[8]);
    } else {
        // RH: M[k][m] += P(G)[k][m] / r - 2 D_k (D^T G)_m / r^3, the first
        // term as Phi[k][m] += G[k][m] / r (am), mapped by P once per tile
        const T Dk[3] = {D0, D1, D2};
        const T qd = (T(-2) * invr) * w;
        T u[3];
#pragma unroll
        for (int mm = 0; mm < 3; ++mm) u[mm] = dot3(D0, D1, D2, q[3 + mm], q[6 + mm], q[9 + mm]);
#pragma unroll
        for (int kk = 0; kk < 3; ++kk) {
            const T dq = qd * Dk[kk];
#pragma unroll
            for (int mm = 0; mm < 3; ++mm) {
                am[kk * 3 + mm] = fma_t(invr, q[3 + kk * 3 + mm], am[kk * 3 + mm]);
                acc[kk * 3 + mm] = fma_t(dq, u[mm], acc[kk * 3 + mm]);
            }
        }
    }
}

// acc += m * (the m-term scalars), am reset; a no-op for kernels without.
template <int KID, typename T>
__device__ __forceinline__ void fold_mterm(const T m[3], const KConst<T> &k, T *acc, T *am) {
#pragma clang fp contract(off)
    if constexpr (KID == TCT_H) {
        const T f = fma_t(k.c2, am[1], am[0]);
        acc[0] = fma_t(m[0], f, acc[0]);
        acc[1] = fma_t(m[1], f, acc[1]);
        acc[2] = fma_t(m[2], f, acc[2]);
        am[0] = am[1] = T(0);
    } else if constexpr (KID == TCT_A) {
        const T f = -k.c0 * am[0];
        acc[0] = fma_t(m[0], f, acc[0]);
        acc[1] = fma_t(m[1], f, acc[1]);
        acc[2] = fma_t(m[2], f, acc[2]);
        am[0] = T(0);
    } else if constexpr (KID == TCT_RH) {
        // acc += P(Phi) = CsRH1 tr(Phi) I + CsRH2 Phi^T - 2 Phi
        // (k.c0 = CsRH1, k.c1 = CsRH2)
        const T tr = (am[0] + am[4]) + am[8];
#pragma unroll
        for (int kk = 0; kk < 3; ++kk)
#pragma unroll
            for (int mm = 0; mm < 3; ++mm) {
                T a = fma_t(k.c1, am[mm * 3 + kk], acc[kk * 3 + mm]);
                a = fma_t(T(-2), am[kk * 3 + mm], a);
                if (kk == mm) a = fma_t(k.c0, tr, a);
                acc[kk * 3 + mm] = a;
            }
#pragma unroll
        for (int q = 0; q < 9; ++q) am[q] = T(0);
    }
}

template <typename T>
inline KConst<T> kernel_consts(int kid, const double *params, double eps2) {
    ElasticConsts c = elastic_consts(params);
    KConst<T> k{};
    k.eps2 = (T)eps2;
    switch (kid) {
    case TCT_U: k.c0 = (T)(c.CsU0 / c.CsU1); break;
    case TCT_T:
    case TCT_A: k.c0 = (T)(c.CsT0 / c.CsT1); break;
    case TCT_H:
        k.c0 = (T)c.CsH1;
        k.c1 = (T)c.CsH3;
        k.c2 = (T)(3.0 * c.CsH1);
        k.c3 = (T)c.nu;
        k.c4 = (T)(3.0 * c.nu);
        break;
    case TCT_RT: k.c0 = (T)(c.CsRT0 / c.CsRT1); break;
    case TCT_RH:
        // the map P(G) = CsRH1 tr(G) I + CsRH2 G^T - 2 G, applied per tile
        // (point form) and per node (treecode expansion)
        k.c0 = (T)c.CsRH1;
        k.c1 = (T)c.CsRH2;
        break;
    default: break;
    }
    return k;
}

}  // namespace tct
