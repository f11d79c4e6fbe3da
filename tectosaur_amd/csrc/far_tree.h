// Treecode templates shared by far_tree.hip (p2m, m2m, ABI) and the
// per-kernel evaluation units far_tree_eval_{u,rt,ra,rh}.hip.
#pragma once
// Treecode farfield for the elastic kernels U, RT, RA, RH on gfx950.
//
// Reference: the treecode of tectosaur/fmm/tsfmm.py (p2p, p2m, m2m, m2p
// passes over an octree; acceptance test of tectosaur/fmm/traversal.cpp:124-
// 160) with the kernels of tectosaur/fmm/ts_kernels.cl (elasticU3, RT3, RA3,
// RH3).  The result approximates the direct sum of far_p2p.hip; its error
// falls with the expansion order and the acceptance parameter.
//
// Formulation (own design): every supported kernel is a combination of
// Laplace potentials and their gradients at the observation point, so one
// Cartesian Taylor expansion of 1/r serves all of them.  With D = y - x,
// 1/r = 1/|D| and d/dx_k (1/r) = D_k / r^3:
//   sum_j (D.f_j) D_k / r^3 = d_k Psi - x'_b d_k Phi_b,
//   Phi_b = sum_j f_jb / r,   Psi = sum_j (y'_j . f_j) / r,
// with y' = y - c and x' = x - c relative to the expansion centre c.
// Channels (Laplace charges per source point, from the source record):
//   U, RA: s (3), y'.s (1);
//   RT:    e (3), y'.e (1), s_a n_j (9);
//   RH:    G_jm (9), y'.G_:m (3).
// Outputs are assembled per kernel from Phi, Psi and their gradients (see
// m2p_eval).  Moments of a node with centre c and radius rho are stored
// scaled, M_a = sum q ((y - c)/rho)^a, and evaluated with the derivatives
// a_a(u) = (1/a!) d^a (1/|u|) at u = (c - x)/rho, so every stored and
// computed number is O(1) in FP32 whatever the mesh units.
//
// Layout: the tree is a binary tree over the 64-point Morton tiles of the
// packed source set (far_p2p.hip's layout; a node covers a contiguous range
// of tiles).  Interaction lists are per observation tile (one wave): entry
// e >= 0 is an m2p node, e < 0 is a p2p source tile -(e + 1).  Each list
// is split into segments of a few entries, one wave per segment, which
// walks it with wave-uniform reads (FP32 source tiles and node moments
// staged in LDS, FP64 source records through the scalar cache);
// tree_reduce_kernel adds a tile's segment sums in segment order.
#include <type_traits>
#include <utility>

#include "far_pair.h"

namespace tct {

// ---------------------------------------------------------------------------
// Multi-indices in graded order: degree n ascending; within a degree, the
// x exponent descending, then the y exponent descending.
__host__ __device__ constexpr int n_terms(int p) {
    return p < 0 ? 0 : (p + 1) * (p + 2) * (p + 3) / 6;
}
__host__ __device__ constexpr int mi_pos(int a, int b, int c) {
    return n_terms(a + b + c - 1) + (b + c) * (b + c + 1) / 2 + c;
}
struct Mi {
    int a, b, c;
};
__host__ __device__ constexpr Mi mi_at(int i) {
    int n = 0;
    while (n_terms(n) <= i) ++n;
    int r = i - n_terms(n - 1);
    int k = 0;  // k = n - a
    while ((k + 1) * (k + 2) / 2 <= r) ++k;
    const int c = r - k * (k + 1) / 2;
    return Mi{n - k, k - c, c};
}

template <int I, int N, typename F>
__device__ __forceinline__ void static_for(F &&f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, N>(f);
    }
}

// a_a(u) = (1/a!) d^a (1/|u|) for |a| <= P, by the recurrence
// n |u|^2 a_a = -(2n - 1) sum_k u_k a_{a-e_k} - (n - 1) sum_k a_{a-2e_k}.
template <int P, typename T>
__device__ __forceinline__ void taylor_derivs(const T u[3], T *a) {
    const T u2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    const T inv_u2 = T(1) / u2;
    a[0] = sqrt(inv_u2);
    const T w[3] = {u[0] * inv_u2, u[1] * inv_u2, u[2] * inv_u2};
    static_for<1, n_terms(P)>([&](auto ic) {
        constexpr int I = decltype(ic)::value;
        constexpr Mi m = mi_at(I);
        constexpr int n = m.a + m.b + m.c;
        constexpr T c1 = T(2 * n - 1) / T(n);
        constexpr T c2 = T(n - 1) / T(n);
        T s1 = T(0), s2 = T(0);
        if constexpr (m.a >= 1) s1 += w[0] * a[mi_pos(m.a - 1, m.b, m.c)];
        if constexpr (m.b >= 1) s1 += w[1] * a[mi_pos(m.a, m.b - 1, m.c)];
        if constexpr (m.c >= 1) s1 += w[2] * a[mi_pos(m.a, m.b, m.c - 1)];
        if constexpr (m.a >= 2) s2 += a[mi_pos(m.a - 2, m.b, m.c)];
        if constexpr (m.b >= 2) s2 += a[mi_pos(m.a, m.b - 2, m.c)];
        if constexpr (m.c >= 2) s2 += a[mi_pos(m.a, m.b, m.c - 2)];
        if constexpr (n >= 2)
            a[I] = -(c1 * s1 + c2 * inv_u2 * s2);
        else
            a[I] = -(c1 * s1);
    });
}

// Channel layout per kernel.
template <int KID>
struct TreeK;
template <>
struct TreeK<TCT_U> {
    static constexpr int C = 4;
};
template <>
struct TreeK<TCT_RA> {
    static constexpr int C = 4;
};
template <>
struct TreeK<TCT_RT> {
    static constexpr int C = 13;
};
template <>
struct TreeK<TCT_RH> {
    static constexpr int C = 12;
};

// Laplace charge of channel ch for one source record q (yp = y - c, physical).
template <int KID, typename T>
__device__ __forceinline__ double charge(int ch, const T *q, const double yp[3]) {
    if constexpr (KID == TCT_U || KID == TCT_RA) {
        if (ch < 3) return (double)q[6 + ch];
        return yp[0] * q[6] + yp[1] * q[7] + yp[2] * q[8];
    } else if constexpr (KID == TCT_RT) {
        if (ch < 3) return (double)q[9 + ch];
        if (ch == 3) return yp[0] * q[9] + yp[1] * q[10] + yp[2] * q[11];
        const int a = (ch - 4) / 3, j = (ch - 4) % 3;
        return (double)q[6 + a] * (double)q[3 + j];
    } else {
        if (ch < 9) return (double)q[3 + ch];
        const int m = ch - 9;
        return yp[0] * q[3 + m] + yp[1] * q[6 + m] + yp[2] * q[9 + m];
    }
}

// The y'-dependent channels: Psi channel index and its three source
// channels (Psi = sum_b y'_b q_b), for the shift of centre in m2m.
template <int KID>
__device__ __forceinline__ int psi_count() {
    return KID == TCT_RH ? 3 : 1;
}
template <int KID>
__device__ __forceinline__ int psi_channel(int i) {
    return KID == TCT_RH ? 9 + i : 3;
}
template <int KID>
__device__ __forceinline__ int psi_source(int i, int b) {
    return KID == TCT_RH ? 3 * b + i : b;
}

// ---------------------------------------------------------------------------
// Tile moments (tree level 0).  grid = n_tiles, block = 256.  The tile's 64
// points first put their Laplace charges and the powers of their scaled
// offsets ((y - c)/rho)^e, e <= order, in LDS; then thread o computes output
// (channel, term) o as a 64-term sum of table products.
// node_cr[4 * node] = centre (3) and radius; tile t is node t.
template <int KID, typename T, bool ABS>
__global__ __launch_bounds__(256) void tree_p2m_kernel(int order, int64_t n_src,
                                                       const T *__restrict__ rec,
                                                       const double *__restrict__ src_org,
                                                       const double *__restrict__ node_cr,
                                                       T *__restrict__ moments) {
    constexpr int C = TreeK<KID>::C;
    constexpr int RW = rec_words<KID>();
    constexpr int PM = TCT_TREE_MAX_ORDER + 1;
    __shared__ T pw[3][PM][G];
    __shared__ T chg[C][G];
    const int t = blockIdx.x;
    const int TT = n_terms(order);
    const double *cr = node_cr + 4 * (size_t)t;
    const int tid = threadIdx.x;
    if (tid < G) {
        // padding records (index >= n_src) carry no charge
        const bool real = (int64_t)t * G + tid < n_src;
        const T *q = rec + ((size_t)t * G + tid) * RW;
        double yp[3];
        for (int k = 0; k < 3; ++k)
            yp[k] = real ? (double)q[k] + ((ABS ? 0.0 : src_org[3 * t + k]) - cr[k]) : 0.0;
        const double inv_rho = 1.0 / cr[3];
        for (int k = 0; k < 3; ++k) {
            const T u = (T)(yp[k] * inv_rho);
            T v = T(1);
            for (int e = 0; e <= order; ++e) {
                pw[k][e][tid] = v;
                v *= u;
            }
        }
        for (int ch = 0; ch < C; ++ch) chg[ch][tid] = real ? (T)charge<KID, T>(ch, q, yp) : T(0);
    }
    __syncthreads();
    for (int o = tid; o < C * TT; o += blockDim.x) {
        const int ch = o / TT;
        const Mi m = mi_at(o % TT);
        T acc = T(0);
        for (int j = 0; j < G; ++j) acc += chg[ch][j] * pw[0][m.a][j] * pw[1][m.b][j] * pw[2][m.c][j];
        moments[(size_t)t * C * TT + o] = acc;
    }
}

// Moments of node pn from its nc children c0 (and c0 + 1):
// M_p[a] = sum_child sum_{b <= a} C(a, b) (delta/rho_p)^(a-b) (rho_c/rho_p)^|b| M'_c[b]
// with delta = c_child - c_parent and M'_c the child moments after the Psi
// channels absorb the centre shift: Psi' = Psi + delta . q_src.  Called by
// a whole workgroup: the children's shifted moments, the binomials and the
// powers of delta/rho_p and rho_c/rho_p are staged in LDS first.  Ends with
// the node's moments written (no trailing barrier).
template <int KID, typename T>
__device__ __forceinline__ void m2m_node(int order, int pn, int c0, int nc,
                                         const double *__restrict__ node_cr,
                                         T *__restrict__ moments) {
    constexpr int C = TreeK<KID>::C;
    constexpr int PM = TCT_TREE_MAX_ORDER + 1;
    constexpr int TMAX = n_terms(TCT_TREE_MAX_ORDER);
    __shared__ double mc[2][C * TMAX];
    __shared__ double binom_t[PM][PM];
    __shared__ double pw_d[2][3][PM];
    __shared__ double pw_s[2][PM];
    __shared__ double delta[2][3];
    const int TT = n_terms(order);
    const double *cp = node_cr + 4 * (size_t)pn;
    const int tid = threadIdx.x;
    if (tid < PM) {
        double v = 1.0;  // row tid of Pascal's triangle
        for (int kk = 0; kk < PM; ++kk) {
            binom_t[tid][kk] = kk <= tid ? v : 0.0;
            v = v * (tid - kk) / (kk + 1);
        }
    } else if (tid >= 32 && tid < 32 + 2 * 4) {
        const int cc = (tid - 32) / 4, k = (tid - 32) % 4;
        if (cc < nc) {
            const double *ccr = node_cr + 4 * (size_t)(c0 + cc);
            const double base = k < 3 ? (ccr[k] - cp[k]) / cp[3] : ccr[3] / cp[3];
            double v = 1.0;
            for (int e = 0; e < PM; ++e) {
                if (k < 3)
                    pw_d[cc][k][e] = v;
                else
                    pw_s[cc][e] = v;
                v *= base;
            }
            if (k < 3) delta[cc][k] = ccr[k] - cp[k];
        }
    }
    for (int o = tid; o < nc * C * TT; o += blockDim.x) {
        const int cc = o / (C * TT), r = o % (C * TT);
        mc[cc][r] = (double)moments[(size_t)(c0 + cc) * C * TT + r];
    }
    __syncthreads();
    // Psi channels absorb the shift of centre (read the unshifted sources)
    for (int o = tid; o < nc * psi_count<KID>() * TT; o += blockDim.x) {
        const int cc = o / (psi_count<KID>() * TT), r = o % (psi_count<KID>() * TT);
        const int ps = r / TT, bi = r % TT;
        double add = 0.0;
        for (int b = 0; b < 3; ++b) add += delta[cc][b] * mc[cc][psi_source<KID>(ps, b) * TT + bi];
        mc[cc][psi_channel<KID>(ps) * TT + bi] += add;
    }
    __syncthreads();
    for (int o = tid; o < C * TT; o += blockDim.x) {
        const int ch = o / TT;
        const Mi m = mi_at(o % TT);
        double acc = 0.0;
        for (int cc = 0; cc < nc; ++cc) {
            const double *M = mc[cc] + ch * TT;
            for (int ba = 0; ba <= m.a; ++ba) {
                const double ca = binom_t[m.a][ba] * pw_d[cc][0][m.a - ba];
                for (int bb = 0; bb <= m.b; ++bb) {
                    const double cb = ca * binom_t[m.b][bb] * pw_d[cc][1][m.b - bb];
                    for (int bc = 0; bc <= m.c; ++bc)
                        acc += cb * binom_t[m.c][bc] * pw_d[cc][2][m.c - bc] *
                               pw_s[cc][ba + bb + bc] * M[mi_pos(ba, bb, bc)];
                }
            }
        }
        moments[(size_t)pn * C * TT + o] = (T)acc;
    }
}


// One level of the upward pass: parents p_off.. from children c_off + 2i.
template <int KID, typename T>
__global__ __launch_bounds__(256) void tree_m2m_kernel(int order, int p_off, int n_par, int c_off,
                                                       int n_child,
                                                       const double *__restrict__ node_cr,
                                                       T *__restrict__ moments) {
    const int i = blockIdx.x;
    if (i >= n_par) return;
    m2m_node<KID, T>(order, p_off + i, c_off + 2 * i, (2 * i + 1 < n_child) ? 2 : 1, node_cr,
                     moments);
}

// ---------------------------------------------------------------------------
// One multipole node evaluated at one observation point.  a: derivatives up
// to order P + 1 at u = (c - x)/rho; Rt = c - x (physical); M: the node's
// scaled moments [C][n_terms(P)] (wave-uniform); out accumulates NO sums.
//
// The moments are wave-uniform: tree_eval_kernel stages the node's C x TT
// values in LDS and they are read back as broadcasts into VGPRs.  They are
// consumed in groups of a few channels x CH terms, fenced by scheduling
// barriers, so only a few dozen of them are live in registers at a time
// (without the fences the scheduler hoists every load of the node and the
// registers spill).
template <int KID>
struct TreeChunk {
    static constexpr int CH = KID == TCT_RT ? 4 : 8;
};

template <int KID, int P, typename T>
__device__ __forceinline__ void m2p_eval(const T *a, const T Rt[3], const T *__restrict__ M,
                                         T inv_rho, const T mn[3], const KConst<T> &k, T *out) {
    constexpr int TT = n_terms(P);
    constexpr int CH = TreeChunk<KID>::CH;
    constexpr int NCH = (TT + CH - 1) / CH;
    // t_k(I) = (a_k + 1) a_{I + e_k}: the x_k-gradient weight of term I
    auto tk = [&](auto ic, int kk) -> T {
        constexpr int I = decltype(ic)::value;
        constexpr Mi m = mi_at(I);
        if (kk == 0) return T(m.a + 1) * a[mi_pos(m.a + 1, m.b, m.c)];
        if (kk == 1) return T(m.b + 1) * a[mi_pos(m.a, m.b + 1, m.c)];
        return T(m.c + 1) * a[mi_pos(m.a, m.b, m.c + 1)];
    };
    const T ir2 = inv_rho * inv_rho;
    if constexpr (KID == TCT_U || KID == TCT_RA) {
        T phi[3] = {0, 0, 0}, g[3] = {0, 0, 0}, h[3] = {0, 0, 0};
        static_for<0, NCH>([&](auto cc) {
            constexpr int I0 = decltype(cc)::value * CH;
            constexpr int I1 = I0 + CH < TT ? I0 + CH : TT;
            __builtin_amdgcn_sched_barrier(0);
            static_for<I0, I1>([&](auto ic) {
                constexpr int I = decltype(ic)::value;
                const T t0 = tk(ic, 0), t1 = tk(ic, 1), t2 = tk(ic, 2);
                const T m0 = M[0 * TT + I], m1 = M[1 * TT + I], m2 = M[2 * TT + I];
                phi[0] += m0 * a[I];
                phi[1] += m1 * a[I];
                phi[2] += m2 * a[I];
                const T nI = M[3 * TT + I] + Rt[0] * m0 + Rt[1] * m1 + Rt[2] * m2;
                g[0] += nI * t0;
                g[1] += nI * t1;
                g[2] += nI * t2;
                if constexpr (KID == TCT_RA) {
                    const T mt = mn[0] * t0 + mn[1] * t1 + mn[2] * t2;
                    h[0] += m0 * mt;
                    h[1] += m1 * mt;
                    h[2] += m2 * mt;
                }
            });
        });
        __builtin_amdgcn_sched_barrier(0);
        if constexpr (KID == TCT_U) {
#pragma unroll
            for (int kk = 0; kk < 3; ++kk) out[kk] += inv_rho * k.c0 * phi[kk] - ir2 * g[kk];
        } else {
#pragma unroll
            for (int kk = 0; kk < 3; ++kk) {
                out[kk] += inv_rho * phi[kk];
                out[3 + kk] -= ir2 * g[kk];
                out[6 + kk] -= ir2 * h[kk];
            }
        }
    } else if constexpr (KID == TCT_RT) {
        T phi[3] = {0, 0, 0}, g[3] = {0, 0, 0}, h[3] = {0, 0, 0};
        static_for<0, NCH>([&](auto cc) {
            constexpr int I0 = decltype(cc)::value * CH;
            constexpr int I1 = I0 + CH < TT ? I0 + CH : TT;
            // e, Psi_e channels
            __builtin_amdgcn_sched_barrier(0);
            static_for<I0, I1>([&](auto ic) {
                constexpr int I = decltype(ic)::value;
                const T m0 = M[0 * TT + I], m1 = M[1 * TT + I], m2 = M[2 * TT + I];
                phi[0] += m0 * a[I];
                phi[1] += m1 * a[I];
                phi[2] += m2 * a[I];
                const T nI = M[3 * TT + I] + Rt[0] * m0 + Rt[1] * m1 + Rt[2] * m2;
                g[0] += nI * tk(ic, 0);
                g[1] += nI * tk(ic, 1);
                g[2] += nI * tk(ic, 2);
            });
            // s_a n_j channels, one a at a time
            static_for<0, 3>([&](auto ac) {
                constexpr int aa = decltype(ac)::value;
                __builtin_amdgcn_sched_barrier(0);
                static_for<I0, I1>([&](auto ic) {
                    constexpr int I = decltype(ic)::value;
                    // one FMA per term (a sum added to h costs two more ops)
                    h[aa] += M[(4 + 3 * aa + 0) * TT + I] * tk(ic, 0);
                    h[aa] += M[(4 + 3 * aa + 1) * TT + I] * tk(ic, 1);
                    h[aa] += M[(4 + 3 * aa + 2) * TT + I] * tk(ic, 2);
                });
            });
        });
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int kk = 0; kk < 3; ++kk)
            out[kk] += inv_rho * phi[kk] - ir2 * (k.c0 * g[kk] + h[kk]);
    } else {
        // RH: M_km += A_km(Phi) - 2 (d_k Psi_m - x'_j d_k Phi_jm),
        // A = CsRH1 tr(Phi) I + CsRH2 Phi^T - 2 Phi (k.c0 = CsRH1, k.c1 = CsRH2).
        T phi[9], b[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) phi[q] = b[q] = T(0);
        static_for<0, NCH>([&](auto cc) {
            constexpr int I0 = decltype(cc)::value * CH;
            constexpr int I1 = I0 + CH < TT ? I0 + CH : TT;
            static_for<0, 3>([&](auto mc) {
                constexpr int mm = decltype(mc)::value;
                __builtin_amdgcn_sched_barrier(0);
                static_for<I0, I1>([&](auto ic) {
                    constexpr int I = decltype(ic)::value;
                    const T g0 = M[(0 + mm) * TT + I], g1 = M[(3 + mm) * TT + I],
                            g2 = M[(6 + mm) * TT + I];
                    phi[0 + mm] += g0 * a[I];
                    phi[3 + mm] += g1 * a[I];
                    phi[6 + mm] += g2 * a[I];
                    const T nI = M[(9 + mm) * TT + I] + Rt[0] * g0 + Rt[1] * g1 + Rt[2] * g2;
                    b[0 * 3 + mm] += nI * tk(ic, 0);
                    b[1 * 3 + mm] += nI * tk(ic, 1);
                    b[2 * 3 + mm] += nI * tk(ic, 2);
                });
            });
        });
        __builtin_amdgcn_sched_barrier(0);
        const T tr = phi[0] + phi[4] + phi[8];
#pragma unroll
        for (int kk = 0; kk < 3; ++kk)
#pragma unroll
            for (int mm = 0; mm < 3; ++mm) {
                T A = k.c1 * phi[mm * 3 + kk] - T(2) * phi[kk * 3 + mm];
                if (kk == mm) A += k.c0 * tr;
                out[kk * 3 + mm] += inv_rho * A + T(2) * ir2 * b[kk * 3 + mm];
            }
    }
}

// Evaluation: one wave (block of 64) per list segment.  seg [n_seg][3] int32
// = (observation tile, first entry, end entry) of `list`; each wave sums its
// segment for the tile's 64 points into seg_partial [n_seg][NO][64] (f64),
// and tree_reduce_kernel adds a tile's segments in order.
// FP32 (ABS = false): obs = loc32 [6][n_obs_pad] local to obs_org per tile,
// records local to src_org per tile.  FP64 (ABS = true): obs = soa64
// absolute, records absolute.  A p2p tile's source records are copied by
// the wave into LDS and read back as broadcasts (FP32; FP64 reads them with
// wave-uniform scalar loads); an m2p node's moments (C x TT values, too many
// for SGPRs) are copied into LDS the same way.
template <int KID, int P, typename T, bool ABS>
__global__ __launch_bounds__(64) void tree_eval_kernel(
    int n_obs_pad, const T *__restrict__ obs, const double *__restrict__ obs_org,
    const T *__restrict__ rec, const double *__restrict__ src_org,
    const double *__restrict__ node_cr, const T *__restrict__ moments,
    const int *__restrict__ seg, const int *__restrict__ list, KConst<T> k,
    double *__restrict__ seg_partial) {
    constexpr int RW = rec_words<KID>();
    constexpr int NO = n_out<KID>();
    constexpr int C = TreeK<KID>::C;
    constexpr int TT = n_terms(P);
    constexpr int NM = C * TT;
    // FP32: a p2p tile's 64 source records are staged in LDS (lane l copies
    // 16-byte words l, l + 64, ... of the tile) and read back with broadcast
    // LDS reads, so the pair body's operands are VGPRs, not the SGPRs of
    // scalar loads (an instruction that reads an SGPR is never dual-issued);
    // the kernel constants are kept in VGPRs for the same reason
    constexpr bool STAGE = !ABS && RW % 4 == 0;
    typedef float tf4 __attribute__((ext_vector_type(4)));
    __shared__ T smom[NM];
    __shared__ tf4 srec[STAGE ? G * RW / 4 : 1];
    KConst<T> kv = k;
    if constexpr (STAGE) {
        asm volatile("" : "+v"(kv.c0));
        asm volatile("" : "+v"(kv.c1));
        asm volatile("" : "+v"(kv.c2));
        asm volatile("" : "+v"(kv.c3));
        asm volatile("" : "+v"(kv.c4));
        asm volatile("" : "+v"(kv.eps2));
    }
    const int sg = blockIdx.x;
    const int tile = seg[3 * sg];
    const int lane = (int)threadIdx.x;
    const int i = tile * G + lane;

    T x[3], mn[3];
    double oo[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        x[d] = obs[(size_t)d * n_obs_pad + i];
        mn[d] = (needs_obs_normal<KID>() || KID == TCT_RA) ? obs[(size_t)(3 + d) * n_obs_pad + i]
                                                          : T(0);
        oo[d] = ABS ? 0.0 : obs_org[3 * tile + d];
    }
    double acc64[NO];
#pragma unroll
    for (int q = 0; q < NO; ++q) acc64[q] = 0.0;

    const int b0 = seg[3 * sg + 1], b1 = seg[3 * sg + 2];
    for (int e = b0; e < b1; ++e) {
        const int code = list[e];
        T acc[NO];
#pragma unroll
        for (int q = 0; q < NO; ++q) acc[q] = T(0);
        if (code < 0) {
            const int t = -code - 1;
            T xs[3];
#pragma unroll
            for (int d = 0; d < 3; ++d)
                xs[d] = ABS ? x[d] : x[d] - (T)(src_org[3 * t + d] - oo[d]);
            const T *r = rec + (size_t)t * G * RW;
            T am[n_mterm_arr<KID>()] = {};
            if constexpr (STAGE) {
                const tf4 *src = (const tf4 *)r;
#pragma unroll
                for (int v = 0; v < RW / 4; ++v) srec[v * G + lane] = src[v * G + lane];
                __syncthreads();
#pragma unroll 2
                for (int j = 0; j < G; ++j) {
                    T q[RW];
#pragma unroll
                    for (int v = 0; v < RW / 4; ++v) {
                        const tf4 f = srec[j * (RW / 4) + v];
                        q[4 * v] = f.x;
                        q[4 * v + 1] = f.y;
                        q[4 * v + 2] = f.z;
                        q[4 * v + 3] = f.w;
                    }
                    pair_eval<KID, T>(xs, mn, q, kv, acc, am);
                }
                __syncthreads();
            } else {
#pragma unroll 2
                for (int j = 0; j < G; ++j) pair_eval<KID, T>(xs, mn, r + j * RW, k, acc, am);
            }
            fold_mterm<KID, T>(mn, kv, acc, am);
        } else {
            const T *Mg = moments + (size_t)code * NM;
            constexpr int NL = (NM + G - 1) / G;
            T mv[NL];
#pragma unroll
            for (int l = 0; l < NL; ++l) {
                const int idx = l * G + lane;
                mv[l] = idx < NM ? Mg[idx] : T(0);
            }
            const double *cr = node_cr + 4 * (size_t)code;
            T Rt[3], u[3];
            const T inv_rho = (T)(1.0 / cr[3]);
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                Rt[d] = (T)(cr[d] - oo[d]) - x[d];
                u[d] = Rt[d] * inv_rho;
            }
            T a[n_terms(P + 1)];
            taylor_derivs<P + 1, T>(u, a);
#pragma unroll
            for (int l = 0; l < NL; ++l) {
                const int idx = l * G + lane;
                if (idx < NM) smom[idx] = mv[l];
            }
            __syncthreads();
            m2p_eval<KID, P, T>(a, Rt, smom, inv_rho, mn, k, acc);
            __syncthreads();
        }
#pragma unroll
        for (int q = 0; q < NO; ++q) acc64[q] += (double)acc[q];
    }
    double *out = seg_partial + (size_t)sg * NO * G;
#pragma unroll
    for (int q = 0; q < NO; ++q) out[q * G + lane] = acc64[q];
}


template <int KID, int P>
inline void launch_eval(int prec, dim3 grid, hipStream_t st, int n_obs_pad, const void *obs,
                        const double *obs_org, const void *rec, const double *src_org,
                        const double *node_cr, const void *moments, const int *seg,
                        const int *list, const double *params, double eps2, double *partial) {
    if (prec == TCT_F32)
        hipLaunchKernelGGL((tree_eval_kernel<KID, P, float, false>), grid, dim3(G), 0, st,
                           n_obs_pad, (const float *)obs, obs_org, (const float *)rec, src_org,
                           node_cr, (const float *)moments, seg, list,
                           kernel_consts<float>(KID, params, eps2), partial);
    else
        hipLaunchKernelGGL((tree_eval_kernel<KID, P, double, true>), grid, dim3(G), 0, st,
                           n_obs_pad, (const double *)obs, obs_org, (const double *)rec, src_org,
                           node_cr, (const double *)moments, seg, list,
                           kernel_consts<double>(KID, params, 0.0), partial);
}

template <int KID>
inline int eval_order(int order, int prec, dim3 grid, hipStream_t st, int n_obs_pad,
                      const void *obs, const double *obs_org, const void *rec,
                      const double *src_org, const double *node_cr, const void *moments,
                      const int *seg, const int *list, const double *params, double eps2,
                      double *partial) {
#define TCT_EV(P)                                                                              \
    launch_eval<KID, P>(prec, grid, st, n_obs_pad, obs, obs_org, rec, src_org, node_cr,        \
                        moments, seg, list, params, eps2, partial)
    switch (order) {
    case 2: TCT_EV(2); break;
    case 3: TCT_EV(3); break;
    case 4: TCT_EV(4); break;
    case 5: TCT_EV(5); break;
    case 6: TCT_EV(6); break;
    case 7: TCT_EV(7); break;
    case 8: TCT_EV(8); break;
    default:
        return set_error(TCT_EUNSUPPORTED, "tree_eval: order %d not built (2-8)", order);
    }
#undef TCT_EV
    return TCT_OK;
}


// Evaluation entry of one kernel (defined in far_tree_eval_<k>.hip).
template <int KID>
int tree_eval_kernel_entry(int order, int prec, dim3 grid, hipStream_t st, int n_obs_pad,
                           const void *obs, const double *obs_org, const void *rec,
                           const double *src_org, const double *node_cr,
                           const void *moments, const int *seg, const int *list,
                           const double *params, double eps2, double *partial);
template <>
int tree_eval_kernel_entry<TCT_U>(int order, int prec, dim3 grid, hipStream_t st,
                                int n_obs_pad, const void *obs, const double *obs_org,
                                const void *rec, const double *src_org,
                                const double *node_cr, const void *moments,
                                const int *seg, const int *list,
                                const double *params, double eps2, double *partial);
template <>
int tree_eval_kernel_entry<TCT_RT>(int order, int prec, dim3 grid, hipStream_t st,
                                int n_obs_pad, const void *obs, const double *obs_org,
                                const void *rec, const double *src_org,
                                const double *node_cr, const void *moments,
                                const int *seg, const int *list,
                                const double *params, double eps2, double *partial);
template <>
int tree_eval_kernel_entry<TCT_RA>(int order, int prec, dim3 grid, hipStream_t st,
                                int n_obs_pad, const void *obs, const double *obs_org,
                                const void *rec, const double *src_org,
                                const double *node_cr, const void *moments,
                                const int *seg, const int *list,
                                const double *params, double eps2, double *partial);
template <>
int tree_eval_kernel_entry<TCT_RH>(int order, int prec, dim3 grid, hipStream_t st,
                                int n_obs_pad, const void *obs, const double *obs_org,
                                const void *rec, const double *src_org,
                                const double *node_cr, const void *moments,
                                const int *seg, const int *list,
                                const double *params, double eps2, double *partial);

}  // namespace tct
