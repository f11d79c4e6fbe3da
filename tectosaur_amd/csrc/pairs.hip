// Nearfield pair integrator, FP64, gfx950.
//
// Reference semantics: single_pairs{Z,N} and single_pairsN_adj
// (tectosaur/assemble.cl:24-103) through integrate_pair
// (tectosaur/integral_primitives.cl:174-213), driven by PairsIntegrator
// (tectosaur/nearfield/pairs_integrator.py:27-93).  For each triangle pair
//   out[b_obs, d_obs, b_src, d_src] =
//       J_obs J_src sum_q w_q phi_b_obs(x_q) phi_b_src(y_q) K_{d_obs d_src}(y_q - x_q)
// (regularized kernels: the surface-curl forms of tectosaur/kernels/
// kernels.cl:29-98), optionally skipping r2 == 0, with the rotation,
// flip and de-rotation of adjacent pairs.
//
// Design:
//  * G lanes share one pair (G = 1, 4, 16 or 64); each lane takes every
//    G-th quadrature point, so a large singular rule (thousands of points)
//    is spread over a whole wavefront ("one workgroup per pair") while a
//    small correction rule keeps one pair per lane.
//  * Accumulators hold only what varies with the quadrature point:
//      U/T/A/H : the 81 basis-weighted kernel entries;
//      RT/RA   : 30 moments (w phi Q1, w phi Q2 D_p D_t, w phi phi Q3 n.D);
//      RH      : 7 moments (w/r and w D_i D_j / r^3) -- the surface curls
//                are constant per triangle and are applied once at the end.
//  * The G partial sums are combined with an xor-butterfly inside the
//    group, then one lane forms and writes the 81 outputs (de-rotated).
#include "pair_common.h"

namespace tct {

// Per-pair geometry of D = y(sx, sy) - x(ox, oy) for the P1 maps
// x = sum_b phi_b(ox, oy) v_b (pts_from_basis, integral_primitives.cl:121-128)
// written in edge form: D = b + sx se0 + sy se1 + ox oe0 + oy oe1 with
// b = s_v0 - o_v0, se = s_v{1,2} - s_v0, oe = -(o_v{1,2} - o_v0): four FMAs
// per component and quadrature point instead of two basis-weighted sums.
struct PairGeom {
    double b[3], se[2][3], oe[2][3];
};

__device__ __forceinline__ void pair_geom(const Tri &ot, const Tri &st, PairGeom &g) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        g.b[d] = st.v[0][d] - ot.v[0][d];
        g.se[0][d] = st.v[1][d] - st.v[0][d];
        g.se[1][d] = st.v[2][d] - st.v[0][d];
        g.oe[0][d] = ot.v[0][d] - ot.v[1][d];
        g.oe[1][d] = ot.v[0][d] - ot.v[2][d];
    }
}

// The r = 0 test of a pair class (CK): CK_NONE for pairs that never touch;
// CK_FULL skips r2 == 0 and the points whose obs and src parameters are
// equal on pairs that are one triangle in one frame; CK_R2 skips r2 == 0
// only, for rules with no point of equal obs and src parameters.
constexpr int CK_NONE = 0, CK_FULL = 1, CK_R2 = 2;

// Normals a kernel body reads: obs (m) and src (n).
struct PairNormals {
    double m[3], n[3];
};

// One rule point of one pair.  The r = 0 test skips the points at r = 0
// (the "Z" kernels, assemble.cl:24-47, and the adjacent kernels, :49-103,
// which test r2 == 0 on the basis-weighted positions).  The edge form
// leaves a rounding residue where the basis-weighted positions coincide
// exactly, so the skip is decided from the rule instead: a pair whose two
// local triangles are the same triangle in the same frame (`same`) has
// exactly coincident points where the obs and src parameters are equal,
// and those are the only exact zeros of a rule (the rules used here have
// none: checked on the host, which then asks for CK_R2); r2 == 0 is skipped
// as well.  A skipped point gets 1/r = 0, which zeroes every term it feeds.
template <int KID, int CK>
__device__ __forceinline__ void accumulate_point(const PairNormals &nm, const PairGeom &g,
                                                 const double *q, double w,
                                                 const ElasticConsts &c, double *acc,
                                                 bool same) {
    const double ox = q[0], oy = q[1], sx = q[2], sy = q[3];
    double D[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double t = fma(sx, g.se[0][d], g.b[d]);
        t = fma(sy, g.se[1][d], t);
        t = fma(ox, g.oe[0][d], t);
        D[d] = fma(oy, g.oe[1][d], t);
    }
    const double r2 = D[0] * D[0] + D[1] * D[1] + D[2] * D[2];
    double invr = rsq_pos(r2);
    if constexpr (CK == CK_FULL) {
        const bool skip = (same && ox == sx && oy == sy) || r2 == 0.0;
        invr = skip ? 0.0 : invr;
    } else if constexpr (CK == CK_R2) {
        invr = r2 == 0.0 ? 0.0 : invr;
    }
    if constexpr (KID <= TCT_H) {
        const double po[3] = {1.0 - ox - oy, ox, oy};
        const double ps[3] = {1.0 - sx - sy, sx, sy};
        double K[9];
        tensor_invr<KID>(D, invr, nm.m, nm.n, c, K);
#pragma unroll
        for (int bo = 0; bo < 3; ++bo)
#pragma unroll
            for (int bs = 0; bs < 3; ++bs) {
                const double f = w * po[bo] * ps[bs];
#pragma unroll
                for (int dob = 0; dob < 3; ++dob)
#pragma unroll
                    for (int ds = 0; ds < 3; ++ds)
                        acc[bo * 27 + dob * 9 + bs * 3 + ds] += f * K[dob * 3 + ds];
            }
    } else if constexpr (KID == TCT_RH) {
        const double wi = w * invr;
        const double wi3 = wi * invr * invr;
        acc[0] += wi;
        const double e[3] = {wi3 * D[0], wi3 * D[1], wi3 * D[2]};
#pragma unroll
        for (int p = 0; p < 3; ++p)
#pragma unroll
            for (int t = p; t < 3; ++t) acc[1 + sym6(p, t)] += e[p] * D[t];
    } else {
        // RT (weights follow the obs basis, normal = src) or RA (src basis, obs normal).
        const double po[3] = {1.0 - ox - oy, ox, oy};
        const double ps[3] = {1.0 - sx - sy, sx, sy};
        const double invr3 = invr * invr * invr;
        const double *nn = KID == TCT_RT ? nm.n : nm.m;
        const double *pL = KID == TCT_RT ? po : ps;
        const double Q1 = c.CsRT1 * invr;
        const double Q2 = c.CsRT0 * invr3;
        const double Q3nD = c.CsRT2 * invr3 * (nn[0] * D[0] + nn[1] * D[1] + nn[2] * D[2]);
        double DD[6];
#pragma unroll
        for (int p = 0; p < 3; ++p)
#pragma unroll
            for (int t = p; t < 3; ++t) DD[sym6(p, t)] = D[p] * D[t];
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const double wl = w * pL[b];
            acc[b] += wl * Q1;
            const double wq = wl * Q2;
#pragma unroll
            for (int k = 0; k < 6; ++k) acc[3 + 6 * b + k] += wq * DD[k];
        }
#pragma unroll
        for (int bo = 0; bo < 3; ++bo)
#pragma unroll
            for (int bs = 0; bs < 3; ++bs) acc[21 + 3 * bo + bs] += w * po[bo] * ps[bs] * Q3nD;
    }
}


// Output value (rotated-frame indices bo, dob, bs, ds) from the reduced
// accumulators, before the J_obs J_src factor.
template <int KID>
__device__ __forceinline__ double final_value(const double *acc, const Tri &ot, const Tri &st,
                                              const ElasticConsts &c, int bo, int dob, int bs,
                                              int ds) {
    if constexpr (KID <= TCT_H) {
        return acc[bo * 27 + dob * 9 + bs * 3 + ds];
    } else if constexpr (KID == TCT_RH) {
        const double *co = ot.curl[bo];
        const double *cs = st.curl[bs];
        const double cc = co[0] * cs[0] + co[1] * cs[1] + co[2] * cs[2];
        const double S0 = acc[0];
        const double S2 = acc[1 + sym6(dob, ds)];
        return c.CsRH0 * (-2.0 * S2 * cc +
                          S0 * (c.CsRH1 * co[dob] * cs[ds] + c.CsRH2 * co[ds] * cs[dob] -
                                (dob == ds ? 2.0 * cc : 0.0)));
    } else if constexpr (KID == TCT_RT) {
        // sum_Ij curl_s[bs][Ij] (e[Ij][do][ds] A1[bo] + sum_t e[Ij][t][ds] A2[bo][do,t])
        //   + delta(do, ds) A3[bo][bs]
        double v = (dob == ds) ? acc[21 + 3 * bo + bs] : 0.0;
        for (int Ij = 0; Ij < 3; ++Ij) {
            double g = levi(Ij, dob, ds) * acc[bo];
            for (int t = 0; t < 3; ++t) g += levi(Ij, t, ds) * acc[3 + 6 * bo + sym6(dob, t)];
            v += st.curl[bs][Ij] * g;
        }
        return v;
    } else {
        // RA: sum_Ij curl_o[bo][Ij] (e[Ij][ds][do] B1[bs] + sum_t e[Ij][t][do] B2[bs][ds,t])
        //   - delta(do, ds) B3[bo][bs]
        double v = (dob == ds) ? -acc[21 + 3 * bo + bs] : 0.0;
        for (int Ij = 0; Ij < 3; ++Ij) {
            double g = levi(Ij, ds, dob) * acc[bs];
            for (int t = 0; t < 3; ++t) g += levi(Ij, t, dob) * acc[3 + 6 * bs + sym6(ds, t)];
            v += ot.curl[bo][Ij] * g;
        }
        return v;
    }
}

// Rule point iq: (obs xi, obs eta, src xi, src eta, weight).
__device__ __forceinline__ void load_point(double (&q)[5], const double *__restrict__ qx,
                                           const double *__restrict__ qw, int iq) {
#pragma unroll
    for (int k = 0; k < 4; ++k) q[k] = qx[4 * (size_t)iq + k];
    q[4] = qw[iq];
}
// The rule point at pq (4 parameters) and pw (weight).
__device__ __forceinline__ void load_point_at(double (&q)[5], const double *pq, const double *pw) {
#pragma unroll
    for (int k = 0; k < 4; ++k) q[k] = pq[k];
    q[4] = pw[0];
}
// An empty statement that reads the last loaded word of the point (so the
// scalar loads are waited for here), keeps memory accesses on their side of
// it, and passes one accumulator and the pair's base offset b through: the
// previous point's update of that accumulator comes before it and the next
// point's arithmetic, which starts from b, after it.
__device__ __forceinline__ void fence_point(double last, double (&b)[3], double &dep) {
    asm volatile("" : "+v"(b[0]), "+v"(b[1]), "+v"(b[2]), "+v"(dep) : "s"(last) : "memory");
}

// The 81 outputs of one pair from its reduced accumulators: the triangle
// records are formed again, from indices the compiler cannot match with the
// first loads (it would otherwise keep those records live through the rule
// loop), and the de-rotation of assemble.cl:77-101 is applied: out[b1] =
// res[od[b1]] with od[i] = (i - clicks) mod 3, entries 0 and 1 of the source
// map swapped when the source was flipped -- written as the inverse map on
// the output address, so every register index stays static.
template <int KID>
__device__ __forceinline__ void write_pair_output(const double *acc,
                                                  const double *__restrict__ pts,
                                                  const int32_t *__restrict__ tris,
                                                  const int32_t *pp, int oc, int sc, int flip,
                                                  const ElasticConsts &c, double *o) {
    int oi = pp[0], si = pp[1];
    asm volatile("" : "+v"(oi), "+v"(si));
    Tri ot, st;
    load_tri(ot, pts, tris, oi, oc, 0);
    load_tri(st, pts, tris, si, sc, flip);
    int sd[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) sd[i] = pmod3(i - sc);
    if (flip) {
        const int tmp = sd[0];
        sd[0] = sd[1];
        sd[1] = tmp;
    }
    int oinv[3], sinv[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        oinv[i] = pmod3(i + oc);
        sinv[i] = (sd[0] == i) ? 0 : ((sd[1] == i) ? 1 : 2);
    }
    const double J = ot.jac * st.jac;
#pragma unroll
    for (int i1 = 0; i1 < 3; ++i1)
#pragma unroll
        for (int d1 = 0; d1 < 3; ++d1)
#pragma unroll
            for (int i2 = 0; i2 < 3; ++i2)
#pragma unroll
                for (int d2 = 0; d2 < 3; ++d2)
                    o[oinv[i1] * 27 + d1 * 9 + sinv[i2] * 3 + d2] =
                        J * final_value<KID>(acc, ot, st, c, i1, d1, i2, d2);
}

// One lane walks rule points [q0, q1) of its pair; every lane of the wave
// walks the same points, read through wave-uniform scalar loads.  Those
// return out of order, so a wait for one waits for all: the next point is
// loaded while the current one computes, and fence_point puts the one wait
// per point after the previous point's arithmetic and before the next load
// is issued and the arithmetic that uses the point.
template <int KID, int CK>
__device__ __forceinline__ void rule_loop_uniform(int q0, int q1, const double *__restrict__ qx,
                                                  const double *__restrict__ qw,
                                                  const PairNormals &nm, PairGeom &g,
                                                  const ElasticConsts &c, double *acc,
                                                  bool same) {
    constexpr int NA = n_acc<KID>();
    double qa[5], qb[5];
    load_point(qa, qx, qw, q0);
    int iq = q0;
    for (; iq + 1 < q1; iq += 2) {
        fence_point(qa[4], g.b, acc[NA - 1]);
        load_point(qb, qx, qw, iq + 1);
        accumulate_point<KID, CK>(nm, g, qa, qa[4], c, acc, same);
        fence_point(qb[4], g.b, acc[NA - 1]);
        load_point(qa, qx, qw, iq + 2 < q1 ? iq + 2 : q1 - 1);
        accumulate_point<KID, CK>(nm, g, qb, qb[4], c, acc, same);
    }
    if (iq < q1) {   // the last point, loaded by the final clamp
        fence_point(qa[4], g.b, acc[NA - 1]);
        accumulate_point<KID, CK>(nm, g, qa, qa[4], c, acc, same);
    }
}

template <int KID, int GL, int CK>
__global__ __launch_bounds__(256) void pairs_kernel(int nq, const double *__restrict__ qx,
                                                    const double *__restrict__ qw,
                                                    const double *__restrict__ pts,
                                                    const int32_t *__restrict__ tris, int n_pairs,
                                                    const int32_t *__restrict__ pairs,
                                                    int pair_width, ElasticConsts c,
                                                    double *__restrict__ out) {
    constexpr int NA = n_acc<KID>();
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int pair = gid / GL;
    const int sub = gid % GL;
    // Lanes of a group past the end still take part in the butterfly.
    const bool live = pair < n_pairs;
    const int pr = live ? pair : n_pairs - 1;
    const int32_t *pp = pairs + (size_t)pr * pair_width;
    int oc = 0, sc = 0, flip = 0;
    if (pair_width == 5) {
        oc = pp[2];
        sc = pp[3];
        flip = pp[4] != 0;
    }
    // The rule loop needs only the pair geometry and the normals; the whole
    // triangle records (curls, Jacobians) are formed again after it, so
    // they hold no registers during the loop.
    PairGeom g;
    PairNormals nm;
    {
        Tri ot, st;
        load_tri(ot, pts, tris, pp[0], oc, 0);
        load_tri(st, pts, tris, pp[1], sc, flip);
        pair_geom(ot, st, g);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            nm.m[d] = ot.n[d];
            nm.n[d] = st.n[d];
        }
    }
    // the two local triangles are one triangle in one frame
    const bool same = pp[0] == pp[1] && oc == sc && !flip;

    double acc[NA];
#pragma unroll
    for (int k = 0; k < NA; ++k) acc[k] = 0.0;
    if constexpr (GL == 1) {
        // one pair per lane: every lane walks the same rule
        if (live && nq > 0) rule_loop_uniform<KID, CK>(0, nq, qx, qw, nm, g, c, acc, same);
    } else if (live) {
        // lanes of a group take every GL-th point (consecutive lanes read
        // consecutive points: coalesced vector loads).  The first nfull
        // double strides (2 GL points) are taken by every lane with one
        // trip count and no index clamps, two points in flight in two
        // register sets; each lane's at most two remaining points follow.
        const int nfull = nq / (2 * GL);
        const double *pq = qx + 4 * (size_t)sub;
        const double *pw = qw + sub;
        if (nfull > 0) {
            double qa[5], qb[5];
            load_point_at(qa, pq, pw);
            for (int k = 0; k < nfull; ++k) {
                load_point_at(qb, pq + 4 * GL, pw + GL);
                accumulate_point<KID, CK>(nm, g, qa, qa[4], c, acc, same);
                pq += 8 * GL;
                pw += 2 * GL;
                if (k + 1 < nfull) load_point_at(qa, pq, pw);
                accumulate_point<KID, CK>(nm, g, qb, qb[4], c, acc, same);
            }
        }
        for (int iq = 2 * GL * nfull + sub; iq < nq; iq += GL) {
            double qa[5];
            load_point(qa, qx, qw, iq);
            accumulate_point<KID, CK>(nm, g, qa, qa[4], c, acc, same);
        }
    }
    if constexpr (GL > 1) {
#pragma unroll
        for (int k = 0; k < NA; ++k) {
            double v = acc[k];
#pragma unroll
            for (int off = GL / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
            acc[k] = v;
        }
    }
    if (!live || sub != 0) return;
    write_pair_output<KID>(acc, pts, tris, pp, oc, sc, flip, c, out + (size_t)pair * 81);
}

// RH3 pairs, one lane per (pair, rule chunk): grid (pair blocks, C slots).
// With S = 1 a workgroup holds 256 pairs and slot ch is rule chunk ch of C,
// points [nq ch / C, nq (ch + 1) / C).  With S = 4 or 8 a workgroup holds
// 64 pairs in S waves, wave w takes chunk S ch + w of S C, and the waves'
// sums are added in wave order through LDS, so a slot holds S chunks.  The rule is
// read through wave-uniform scalar loads as with one lane per pair, so pairs
// too few to fill the device one lane each still do.  The 7 accumulators of
// slot ch go to the pair's own output row, out[pair][7 ch .. 7 ch + 7)
// (C <= 11), and pairs_rh_finish_kernel sums the slots in order and writes
// the 81 values.
constexpr int RH_MAX_CHUNKS = 81 / 7;   // chunk accumulators that fit in an output row
constexpr int RH_WAVE_SPLIT = 4;        // chunks per slot when the waves split one
constexpr int RH_WAVE_SPLIT_MAX = 8;    // ... and with eight waves per workgroup

// threads per workgroup of the chunk kernel
constexpr int rh_chunk_block(int S) { return S > 4 ? 64 * S : 256; }

template <int CK, int S>
__global__ __launch_bounds__(rh_chunk_block(S)) void pairs_rh_chunk_kernel(
    int nq, const double *__restrict__ qx, const double *__restrict__ qw,
    const double *__restrict__ pts, const int32_t *__restrict__ tris, int n_pairs,
    const int32_t *__restrict__ pairs, int pair_width, ElasticConsts c, double *__restrict__ out) {
    constexpr int NA = n_acc<TCT_RH>();
    constexpr int PB = rh_chunk_block(S) / S;   // pairs per workgroup, a multiple of 64
    static_assert(PB % 64 == 0, "a wave must not straddle two chunks");
    const int lane = (int)threadIdx.x % PB;
    const int w = S > 1 ? __builtin_amdgcn_readfirstlane((int)threadIdx.x / PB) : 0;
    const int pair = (int)blockIdx.x * PB + lane;
    // lanes past the end work on the last pair (the waves of a workgroup
    // all reach the LDS barrier) and write nothing
    const bool live = pair < n_pairs;
    const int32_t *pp = pairs + (size_t)(live ? pair : n_pairs - 1) * pair_width;
    const int ch = (int)blockIdx.y, nch = (int)gridDim.y;
    int oc = 0, sc = 0, flip = 0;
    if (pair_width == 5) {
        oc = pp[2];
        sc = pp[3];
        flip = pp[4] != 0;
    }
    PairGeom g;
    PairNormals nm;
    {
        Tri ot, st;
        load_tri(ot, pts, tris, pp[0], oc, 0);
        load_tri(st, pts, tris, pp[1], sc, flip);
        pair_geom(ot, st, g);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            nm.m[d] = ot.n[d];
            nm.n[d] = st.n[d];
        }
    }
    const bool same = pp[0] == pp[1] && oc == sc && !flip;
    double acc[NA];
#pragma unroll
    for (int k = 0; k < NA; ++k) acc[k] = 0.0;
    const int64_t cs = (int64_t)ch * S + w, ncs = (int64_t)nch * S;
    const int q0 = (int)(nq * cs / ncs), q1 = (int)(nq * (cs + 1) / ncs);
    if (q1 > q0) rule_loop_uniform<TCT_RH, CK>(q0, q1, qx, qw, nm, g, c, acc, same);
    if constexpr (S > 1) {
        __shared__ double part[S - 1][NA][PB];
        if (w > 0)
#pragma unroll
            for (int k = 0; k < NA; ++k) part[w - 1][k][lane] = acc[k];
        __syncthreads();
        if (w > 0) return;
#pragma unroll
        for (int s = 0; s < S - 1; ++s)
#pragma unroll
            for (int k = 0; k < NA; ++k) acc[k] += part[s][k][lane];
    }
    if (!live) return;
    double *o = out + (size_t)pair * 81 + NA * ch;
#pragma unroll
    for (int k = 0; k < NA; ++k) o[k] = acc[k];
}

__global__ __launch_bounds__(256) void pairs_rh_finish_kernel(
    int n_pairs, const double *__restrict__ pts, const int32_t *__restrict__ tris,
    const int32_t *__restrict__ pairs, int pair_width, int nch, ElasticConsts c,
    double *__restrict__ out) {
    constexpr int NA = n_acc<TCT_RH>();
    const int pair = blockIdx.x * blockDim.x + threadIdx.x;
    if (pair >= n_pairs) return;
    const int32_t *pp = pairs + (size_t)pair * pair_width;
    int oc = 0, sc = 0, flip = 0;
    if (pair_width == 5) {
        oc = pp[2];
        sc = pp[3];
        flip = pp[4] != 0;
    }
    double *o = out + (size_t)pair * 81;
    // every chunk's accumulators are read (in chunk order) before the row
    // is overwritten with the outputs
    double acc[NA];
#pragma unroll
    for (int k = 0; k < NA; ++k) acc[k] = 0.0;
    for (int ch = 0; ch < nch; ++ch)
#pragma unroll
        for (int k = 0; k < NA; ++k) acc[k] += o[NA * ch + k];
    write_pair_output<TCT_RH>(acc, pts, tris, pp, oc, sc, flip, c, o);
}

__global__ void sub_f64_kernel(int64_t n, const double *__restrict__ a,
                               const double *__restrict__ b, double *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = a[i] - b[i];
}

template <int KID, int GL>
static void launch_pairs(hipStream_t st, int nq, const double *qx, const double *qw,
                         const double *pts, const int32_t *tris, int n_pairs,
                         const int32_t *pairs, int pair_width, int check0,
                         const ElasticConsts &c, double *out) {
    const int64_t threads = (int64_t)n_pairs * GL;
    const unsigned nb = (unsigned)ceil_div(threads, 256);
    if (check0 == CK_FULL)
        hipLaunchKernelGGL((pairs_kernel<KID, GL, CK_FULL>), dim3(nb), dim3(256), 0, st, nq, qx,
                           qw, pts, tris, n_pairs, pairs, pair_width, c, out);
    else if (check0 == CK_R2)
        hipLaunchKernelGGL((pairs_kernel<KID, GL, CK_R2>), dim3(nb), dim3(256), 0, st, nq, qx,
                           qw, pts, tris, n_pairs, pairs, pair_width, c, out);
    else
        hipLaunchKernelGGL((pairs_kernel<KID, GL, CK_NONE>), dim3(nb), dim3(256), 0, st, nq, qx,
                           qw, pts, tris, n_pairs, pairs, pair_width, c, out);
}

template <int KID>
static void launch_pairs_g(int gl, hipStream_t st, int nq, const double *qx, const double *qw,
                           const double *pts, const int32_t *tris, int n_pairs,
                           const int32_t *pairs, int pair_width, int check0,
                           const ElasticConsts &c, double *out) {
    switch (gl) {
    case 1: launch_pairs<KID, 1>(st, nq, qx, qw, pts, tris, n_pairs, pairs, pair_width, check0, c, out); break;
    case 4: launch_pairs<KID, 4>(st, nq, qx, qw, pts, tris, n_pairs, pairs, pair_width, check0, c, out); break;
    case 16: launch_pairs<KID, 16>(st, nq, qx, qw, pts, tris, n_pairs, pairs, pair_width, check0, c, out); break;
    default: launch_pairs<KID, 64>(st, nq, qx, qw, pts, tris, n_pairs, pairs, pair_width, check0, c, out); break;
    }
}

}  // namespace tct

using namespace tct;

extern "C" int tct_pairs_integrate(int kernel, int64_t nq, const double *qx, const double *qw,
                                   const double *pts, const int32_t *tris, int64_t n_pairs,
                                   const int32_t *pairs, int pair_width, int check0,
                                   const double *params_host, int lanes_per_pair, double *out,
                                   void *stream) {
    TCT_CHECK_ARG(kernel >= TCT_U && kernel <= TCT_RH, "pairs: kernel id %d", kernel);
    TCT_CHECK_ARG(nq > 0 && nq < (1LL << 31), "pairs: nq %lld", (long long)nq);
    TCT_CHECK_ARG(n_pairs >= 0 && n_pairs < (1LL << 26), "pairs: n_pairs %lld (max 2^26 per call)",
                  (long long)n_pairs);
    TCT_CHECK_ARG(pair_width == 2 || pair_width == 5, "pairs: pair_width %d", pair_width);
    TCT_CHECK_ARG(check0 == CK_NONE || check0 == CK_FULL || check0 == CK_R2, "pairs: check0 %d",
                  check0);
    TCT_CHECK_ARG(qx && qw && pts && tris && params_host && out && (pairs || n_pairs == 0),
                  "pairs: null pointer");
    if (n_pairs == 0) return TCT_OK;
    int gl = lanes_per_pair;
    int chunks = 0;   // RH3 rule chunks (one lane per pair and chunk)
    if (gl < 0) {
        const int64_t cmax = (int64_t)RH_WAVE_SPLIT * RH_MAX_CHUNKS;
        const int64_t cmax8 = (int64_t)RH_WAVE_SPLIT_MAX * RH_MAX_CHUNKS;
        TCT_CHECK_ARG(kernel == TCT_RH && -(int64_t)gl <= nq &&
                          (-gl <= RH_MAX_CHUNKS || (-gl <= cmax && -gl % RH_WAVE_SPLIT == 0) ||
                           (-gl <= cmax8 && -gl % RH_WAVE_SPLIT_MAX == 0)),
                      "pairs: %d rule chunks (RH3 only, at most nq, and at most %d, a "
                      "multiple of %d up to %d or of %d up to %d)",
                      -gl, RH_MAX_CHUNKS, RH_WAVE_SPLIT, (int)cmax, RH_WAVE_SPLIT_MAX,
                      (int)cmax8);
        chunks = -gl;
        gl = 1;
    } else if (gl == 0) {
        // Fewest lanes per pair that still give enough wavefronts, at most
        // what the rule size can use.  One lane per pair reads the rule
        // through wave-uniform scalar loads: on 1.7M near pairs (4,096-point
        // rule, no r = 0 test) it measured 1.6x faster than 64 lanes, and
        // it stays ahead down to ~10k wavefronts; the classes with the
        // r = 0 test (coincident, adjacent) keep gaining from more lanes up
        // to ~16k-46k wavefronts (64-180 per CU).
        const int gmax = nq >= 1024 ? 64 : (nq >= 128 ? 16 : (nq >= 24 ? 4 : 1));
        const int64_t min_waves = check0 ? 16384 : 4096;
        gl = gmax;
        for (int g = 1; g < gmax; g *= 4)
            if (n_pairs * g >= 64LL * min_waves) {
                gl = g;
                break;
            }
        // RH3 with a large rule: one lane per pair and rule chunk instead of
        // several lanes per pair, enough chunks for the wavefronts the lane
        // choice above asks for: up to RH_MAX_CHUNKS one per slot, then
        // RH_WAVE_SPLIT per slot, then RH_WAVE_SPLIT_MAX per slot (at most
        // RH_MAX_CHUNKS slots; nq >= 512 exceeds every count)
        if (kernel == TCT_RH && gl > 1 && nq >= 512) {
            const int64_t want = ceil_div(64LL * min_waves, n_pairs);
            if (want <= RH_MAX_CHUNKS) {
                chunks = (int)(want < 2 ? 2 : want);
            } else if (want <= (int64_t)RH_WAVE_SPLIT * RH_MAX_CHUNKS) {
                chunks = RH_WAVE_SPLIT * (int)ceil_div(want, (int64_t)RH_WAVE_SPLIT);
            } else {
                // at least 8 slots: on the icosphere's 20,480 coincident
                // pairs 64 chunks measured 0.585 of FP64 peak against 0.57
                // with 56 and 0.576 with 88
                int64_t slots = ceil_div(want, (int64_t)RH_WAVE_SPLIT_MAX);
                slots = slots < 8 ? 8 : (slots > RH_MAX_CHUNKS ? RH_MAX_CHUNKS : slots);
                chunks = RH_WAVE_SPLIT_MAX * (int)slots;
            }
            gl = 1;
        }
    }
    TCT_CHECK_ARG(gl == 1 || gl == 4 || gl == 16 || gl == 64, "pairs: lanes_per_pair %d", gl);
    TCT_CHECK_ARG(n_pairs * gl < (1LL << 31), "pairs: n_pairs * lanes_per_pair >= 2^31");
    const ElasticConsts c = elastic_consts(params_host);
    hipStream_t st = as_stream(stream);
    const int np = (int)n_pairs, q = (int)nq;
    if (chunks > 0) {
        // waves per slot: 1 up to 11 chunks, then 4 (up to 44), then 8
        const int S = chunks <= RH_MAX_CHUNKS ? 1
                      : (chunks <= RH_WAVE_SPLIT * RH_MAX_CHUNKS && chunks % RH_WAVE_SPLIT == 0)
                          ? RH_WAVE_SPLIT
                          : RH_WAVE_SPLIT_MAX;
        const int slots = chunks / S;
        const int pb = S == 1 ? 256 : 64;   // pairs per workgroup
        const dim3 grid((unsigned)ceil_div(n_pairs, (int64_t)pb), (unsigned)slots);
#define TCT_RH_CHUNKS(CK)                                                                      \
    if (S == RH_WAVE_SPLIT_MAX)                                                                \
        hipLaunchKernelGGL((pairs_rh_chunk_kernel<CK, RH_WAVE_SPLIT_MAX>), grid,               \
                           dim3(rh_chunk_block(RH_WAVE_SPLIT_MAX)), 0, st, q, qx, qw, pts,     \
                           tris, np, pairs, pair_width, c, out);                               \
    else if (S == RH_WAVE_SPLIT)                                                               \
        hipLaunchKernelGGL((pairs_rh_chunk_kernel<CK, RH_WAVE_SPLIT>), grid,                   \
                           dim3(rh_chunk_block(RH_WAVE_SPLIT)), 0, st, q, qx, qw, pts, tris,   \
                           np, pairs, pair_width, c, out);                                     \
    else                                                                                       \
        hipLaunchKernelGGL((pairs_rh_chunk_kernel<CK, 1>), grid, dim3(rh_chunk_block(1)), 0,   \
                           st, q, qx, qw, pts, tris, np, pairs, pair_width, c, out)
        if (check0 == CK_FULL)
            TCT_RH_CHUNKS(CK_FULL);
        else if (check0 == CK_R2)
            TCT_RH_CHUNKS(CK_R2);
        else
            TCT_RH_CHUNKS(CK_NONE);
#undef TCT_RH_CHUNKS
        hipLaunchKernelGGL(pairs_rh_finish_kernel, dim3((unsigned)ceil_div(n_pairs, (int64_t)256)),
                           dim3(256), 0, st, np, pts, tris, pairs, pair_width, slots, c, out);
        TCT_CHECK_LAUNCH("pairs_integrate");
        return TCT_OK;
    }
    switch (kernel) {
    case TCT_U: launch_pairs_g<TCT_U>(gl, st, q, qx, qw, pts, tris, np, pairs, pair_width, check0, c, out); break;
    case TCT_T: launch_pairs_g<TCT_T>(gl, st, q, qx, qw, pts, tris, np, pairs, pair_width, check0, c, out); break;
    case TCT_A: launch_pairs_g<TCT_A>(gl, st, q, qx, qw, pts, tris, np, pairs, pair_width, check0, c, out); break;
    case TCT_H: launch_pairs_g<TCT_H>(gl, st, q, qx, qw, pts, tris, np, pairs, pair_width, check0, c, out); break;
    case TCT_RT: launch_pairs_g<TCT_RT>(gl, st, q, qx, qw, pts, tris, np, pairs, pair_width, check0, c, out); break;
    case TCT_RA: launch_pairs_g<TCT_RA>(gl, st, q, qx, qw, pts, tris, np, pairs, pair_width, check0, c, out); break;
    default: launch_pairs_g<TCT_RH>(gl, st, q, qx, qw, pts, tris, np, pairs, pair_width, check0, c, out); break;
    }
    TCT_CHECK_LAUNCH("pairs_integrate");
    return TCT_OK;
}

extern "C" int tct_sub_f64(int64_t n, const double *a, const double *b, double *out,
                           void *stream) {
    TCT_CHECK_ARG(n >= 0 && a && b && out, "sub_f64: bad arguments");
    if (n == 0) return TCT_OK;
    hipLaunchKernelGGL(sub_f64_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0,
                       as_stream(stream), n, a, b, out);
    TCT_CHECK_LAUNCH("sub_f64");
    return TCT_OK;
}
