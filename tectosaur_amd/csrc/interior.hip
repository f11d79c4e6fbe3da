// Interior evaluation, FP64, gfx950: point observers against source
// triangles.
//
// Reference semantics:
//  * interior_pairs (tectosaur/assemble.cl:140-186) through
//    integrate_pt_tri (tectosaur/integral_primitives.cl:216-254):
//      out[d_obs, b_src, d_src] = J_src sum_q w_q phi_b_src(y_q) K_{d_obs d_src}(y_q - x)
//    for U, T, A, H, and for RT the regularized tensor with one observation
//    basis function (tectosaur/kernels/kernels.cl:29-62); de-rotated by the
//    source rotation clicks; finite_part keeps only column b_src = clicks.
//  * interior_corners (tectosaur/interior_corners.cl:23-175): the
//    observation point at vertex `clicks` of the source triangle, a 1-D rule
//    in the polar angle and the closed-form radial integrals of that file.
//  * The 3 x 9 blocks of both are applied as one block-sparse matrix
//    (make_pairs_mat + dot, tectosaur/interior.py:14-26,108-114).
//
// Design:
//  * G lanes share one pair, each taking every G-th quadrature point, as in
//    the triangle-pair integrator; the partial sums meet in an xor-butterfly.
//  * U/T/A/H keep the 27 basis-weighted kernel entries; RT keeps 10
//    moments (w Q1, w Q2 D_p D_t, w phi Q3 n.D) and contracts them with the
//    source surface curls once at the end.
#include "pair_common.h"

namespace tct {

template <int KID>
constexpr int n_acc_pt() {
    return KID <= TCT_H ? 27 : 10;
}

template <int KID>
__device__ __forceinline__ void accumulate_pt(const double x[3], const double m[3], const Tri &st,
                                              double qx, double qy, double w, int check0,
                                              const ElasticConsts &c, double *acc) {
    const double ps[3] = {1.0 - qx - qy, qx, qy};
    double D[3];
#pragma unroll
    for (int d = 0; d < 3; ++d)
        D[d] = ps[0] * st.v[0][d] + ps[1] * st.v[1][d] + ps[2] * st.v[2][d] - x[d];
    const double r2 = D[0] * D[0] + D[1] * D[1] + D[2] * D[2];
    if (check0 && r2 == 0.0) return;
    if constexpr (KID <= TCT_H) {
        double K[9];
        tensor<KID>(D, r2, m, st.n, c, K);
#pragma unroll
        for (int bs = 0; bs < 3; ++bs) {
            const double f = w * ps[bs];
#pragma unroll
            for (int dob = 0; dob < 3; ++dob)
#pragma unroll
                for (int ds = 0; ds < 3; ++ds) acc[dob * 9 + bs * 3 + ds] += f * K[dob * 3 + ds];
        }
    } else {
        const double invr = rsqrt(r2);
        const double invr3 = invr * invr * invr;
        const double wq2 = w * c.CsRT0 * invr3;
        const double wq3nD =
            w * c.CsRT2 * invr3 * (st.n[0] * D[0] + st.n[1] * D[1] + st.n[2] * D[2]);
        acc[0] += w * c.CsRT1 * invr;
#pragma unroll
        for (int p = 0; p < 3; ++p)
#pragma unroll
            for (int t = p; t < 3; ++t) acc[1 + sym6(p, t)] += wq2 * D[p] * D[t];
#pragma unroll
        for (int bs = 0; bs < 3; ++bs) acc[7 + bs] += ps[bs] * wq3nD;
    }
}

// Value at (d_obs, b_src, d_src) of the rotated frame, before J_src.
template <int KID>
__device__ __forceinline__ double final_pt(const double *acc, const Tri &st, int dob, int bs,
                                           int ds) {
    if constexpr (KID <= TCT_H) {
        return acc[dob * 9 + bs * 3 + ds];
    } else {
        // sum_Ij curl[bs][Ij] (e[Ij][dob][ds] A1 + sum_t e[Ij][t][ds] A2[dob,t])
        //   + delta(dob, ds) A3[bs]
        double v = (dob == ds) ? acc[7 + bs] : 0.0;
        for (int Ij = 0; Ij < 3; ++Ij) {
            double g = levi(Ij, dob, ds) * acc[0];
            for (int t = 0; t < 3; ++t) g += levi(Ij, t, ds) * acc[1 + sym6(dob, t)];
            v += st.curl[bs][Ij] * g;
        }
        return v;
    }
}

// Writes the 27 outputs of one pair: out[d1*9 + b2*3 + d2] =
// J res[d1*9 + derot(b2)*3 + d2] with derot(b) = (b - clicks) mod 3
// (assemble.cl:157-175); with finite_part only b2 == clicks is nonzero.
// Written as the inverse map on the output address, res basis b going to
// b2 = (b + clicks) mod 3, so the accumulators are read at static indices
// (read at the rotated index they would be held in scratch memory); with
// finite_part only b = 0 is kept.
template <typename F>
__device__ __forceinline__ void write_derotated(double *o, int clicks, int finite_part, double J,
                                                F value) {
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        const int b2 = pmod3(b + clicks);
        const bool keep = !finite_part || b == 0;
#pragma unroll
        for (int d1 = 0; d1 < 3; ++d1)
#pragma unroll
            for (int d2 = 0; d2 < 3; ++d2)
                o[d1 * 9 + b2 * 3 + d2] = keep ? J * value(d1, b, d2) : 0.0;
    }
}

template <int KID, int GL>
__global__ __launch_bounds__(256) void interior_pairs_kernel(
    int nq, const double *__restrict__ qx, const double *__restrict__ qw,
    const double *__restrict__ obs_pts, const double *__restrict__ obs_ns,
    const double *__restrict__ src_pts, const int32_t *__restrict__ src_tris, int n_pairs,
    const int32_t *__restrict__ pairs, int check0, int finite_part, ElasticConsts c,
    double *__restrict__ out) {
    constexpr int NA = n_acc_pt<KID>();
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int pair = gid / GL;
    const int sub = gid % GL;
    const bool live = pair < n_pairs;
    const int pr = live ? pair : n_pairs - 1;
    const int32_t *pp = pairs + (size_t)pr * 3;
    const int clicks = pp[2];
    Tri st;
    load_tri(st, src_pts, src_tris, pp[1], clicks, 0);
    double x[3], m[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        x[d] = obs_pts[3 * (size_t)pp[0] + d];
        m[d] = obs_ns[3 * (size_t)pp[0] + d];
    }
    double acc[NA];
#pragma unroll
    for (int k = 0; k < NA; ++k) acc[k] = 0.0;
    if (live)
        for (int iq = sub; iq < nq; iq += GL)
            accumulate_pt<KID>(x, m, st, qx[2 * iq], qx[2 * iq + 1], qw[iq], check0, c, acc);
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
    write_derotated(out + (size_t)pair * 27, clicks, finite_part, st.jac,
                    [&](int d1, int b, int d2) { return final_pt<KID>(acc, st, d1, b, d2); });
}

// interior_corners.cl: t in [0, pi/2] parameterises the ray from vertex 0
// of the rotated triangle, a = v1 cos t + v2 sin t, rho_max = 1/(cos t +
// sin t).  The radial integrals (file lines 69-73):
//   I1 = -1/(|a|^3 rho_max),  I2 = (log 2 + log|a| + log rho_max - 1)/|a|^3,
//   I3_ij = -a_i a_j/(|a|^5 rho_max);
// the I4 entries there are `1/3*(...)` in C, i.e. zero.  The summed terms
// are CsH0 CsH3 mn T1[b] + delta CsH0 CsH1 mn T0[b] with elasticH3's
// constants, T0 = (-I2 cos - I2 sin + I1, I2 cos, I2 sin), T1 = (I3, 0, 0).
__global__ __launch_bounds__(256) void interior_corners_kernel(
    int nq, const double *__restrict__ qt, const double *__restrict__ qw,
    const double *__restrict__ obs_ns, const double *__restrict__ src_pts,
    const int32_t *__restrict__ src_tris, int n_pairs, const int32_t *__restrict__ pairs,
    ElasticConsts c, double *__restrict__ out) {
    const int pair = blockIdx.x * blockDim.x + threadIdx.x;
    if (pair >= n_pairs) return;
    const int32_t *pp = pairs + (size_t)pair * 3;
    const int clicks = pp[2];
    Tri st;
    load_tri(st, src_pts, src_tris, pp[1], clicks, 0);
    const double *mo = obs_ns + 3 * (size_t)pp[0];
    const double mn = mo[0] * st.n[0] + mo[1] * st.n[1] + mo[2] * st.n[2];
    double v1[3], v2[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        v1[d] = st.v[1][d] - st.v[0][d];
        v2[d] = st.v[2][d] - st.v[0][d];
    }
    // S0[b] = sum w T0[b]; S1[i,j] = sum w I3_ij (symmetric)
    double S0[3] = {0.0, 0.0, 0.0}, S1[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const double log2 = log(2.0);
    for (int iq = 0; iq < nq; ++iq) {
        const double t = qt[iq], w = qw[iq];
        double sn, cs;
        sincos(t, &sn, &cs);
        const double rhomax = 1.0 / (cs + sn);
        double a[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) a[d] = v1[d] * cs + v2[d] * sn;
        const double A2 = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
        const double A = sqrt(A2);
        const double A3 = A2 * A;
        const double I1 = -1.0 / (A3 * rhomax);
        const double I2 = (log2 + log(A) + log(rhomax) - 1.0) / A3;
        S0[0] += w * (-I2 * cs - I2 * sn + I1);
        S0[1] += w * I2 * cs;
        S0[2] += w * I2 * sn;
        const double f = -w / (A3 * A2 * rhomax);
#pragma unroll
        for (int p = 0; p < 3; ++p)
#pragma unroll
            for (int q = p; q < 3; ++q) S1[sym6(p, q)] += f * a[p] * a[q];
    }
    const double k1 = c.CsH0 * c.CsH3 * mn, k0 = c.CsH0 * c.CsH1 * mn;
    write_derotated(out + (size_t)pair * 27, clicks, 0, st.jac, [&](int d1, int b, int d2) {
        double v = (b == 0) ? k1 * S1[sym6(d1, d2)] : 0.0;
        if (d1 == d2) v += k0 * S0[b];
        return v;
    });
}

// y[3 i + d] (+)= sum_{k in row i} sum_j data[k][d][j] x[9 indices[k] + j].
// One thread per output row (point, component).
__global__ __launch_bounds__(256) void bsr39_kernel(int64_t n_rows,
                                                    const int32_t *__restrict__ indptr,
                                                    const int32_t *__restrict__ indices,
                                                    const double *__restrict__ data,
                                                    const double *__restrict__ x,
                                                    double *__restrict__ y, int accumulate) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= 3 * n_rows) return;
    const int64_t i = r / 3;
    const int d = (int)(r % 3);
    double s = 0.0;
    for (int k = indptr[i]; k < indptr[i + 1]; ++k) {
        const double *b = data + (size_t)k * 27 + d * 9;
        const double *xv = x + 9 * (size_t)indices[k];
#pragma unroll
        for (int j = 0; j < 9; ++j) s += b[j] * xv[j];
    }
    y[r] = accumulate ? y[r] + s : s;
}

__device__ __forceinline__ double dot3(const double *a, const double *b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

__device__ __forceinline__ void cross3(const double *a, const double *b, double *c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

// Near-singular pairs by the polar sinh rule (no reference counterpart):
// the source triangle is split at p, the point of the triangle closest to
// the projection of x, into sub-triangles (p, v_k, v_k+1), and each of
// those at the foot of the perpendicular from p into two right triangles
// (signed when the foot falls outside the edge).  A right triangle with
// legs h and L is integrated in polar coordinates about p: the angle
// through tan(phi) = sinh(s), s in [0, asinh(L/h)] (rho_max = h cosh(s),
// smooth to the thin end), the radius through rho = d sinh(u), d = |x - p|
// (the sinh transformation of Johnston and Elliott, Int. J. Numer. Meth.
// Engng 62 (2005) 564-578).  tq/tw and rq/rw are Gauss rules on [0, 1] for
// s and u.  One wavefront per pair: each right triangle's n_t n_r points
// are dealt to the lanes; weights are physical (no J factor).
template <int KID>
__global__ __launch_bounds__(256) void interior_pairs_sinh_kernel(
    int n_t, const double *__restrict__ tq, const double *__restrict__ tw, int n_r,
    const double *__restrict__ rq, const double *__restrict__ rw,
    const double *__restrict__ obs_pts, const double *__restrict__ obs_ns,
    const double *__restrict__ src_pts, const int32_t *__restrict__ src_tris, int n_pairs,
    const int32_t *__restrict__ pairs, ElasticConsts c, double *__restrict__ out) {
    constexpr int NA = n_acc_pt<KID>();
    const int pair = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    const int lane = threadIdx.x % 64;
    const bool live = pair < n_pairs;
    const int pr = live ? pair : n_pairs - 1;
    const int32_t *pp = pairs + (size_t)pr * 3;
    const int clicks = pp[2];
    Tri st;
    load_tri(st, src_pts, src_tris, pp[1], clicks, 0);
    double x[3], m[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        x[d] = obs_pts[3 * (size_t)pp[0] + d];
        m[d] = obs_ns[3 * (size_t)pp[0] + d];
    }
    const double *nn = st.n;
    double e1[3], e2[3], pv[3];
    const double ds = (x[0] - st.v[0][0]) * nn[0] + (x[1] - st.v[0][1]) * nn[1] +
                      (x[2] - st.v[0][2]) * nn[2];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        e1[k] = st.v[1][k] - st.v[0][k];
        e2[k] = st.v[2][k] - st.v[0][k];
        pv[k] = x[k] - ds * nn[k];
    }
    const double g11 = dot3(e1, e1), g12 = dot3(e1, e2), g22 = dot3(e2, e2);
    const double det = g11 * g22 - g12 * g12;
    // split point: the projection if inside, else the closest boundary point
    double p[3] = {pv[0], pv[1], pv[2]};
    {
        double w[3] = {pv[0] - st.v[0][0], pv[1] - st.v[0][1], pv[2] - st.v[0][2]};
        const double h1 = dot3(w, e1), h2 = dot3(w, e2);
        const double xi = (g22 * h1 - g12 * h2) / det, eta = (g11 * h2 - g12 * h1) / det;
        if (!(xi >= 0 && eta >= 0 && xi + eta <= 1)) {
            double best = INFINITY;
            for (int ed = 0; ed < 3; ++ed) {
                const double *va = st.v[ed], *vb = st.v[(ed + 1) % 3];
                double ev[3], wa[3];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    ev[k] = vb[k] - va[k];
                    wa[k] = pv[k] - va[k];
                }
                double t = dot3(wa, ev) / dot3(ev, ev);
                t = t < 0 ? 0 : (t > 1 ? 1 : t);
                double dq[3], q[3];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    q[k] = va[k] + t * ev[k];
                    dq[k] = pv[k] - q[k];
                }
                const double dd = dot3(dq, dq);
                if (dd < best) {
                    best = dd;
#pragma unroll
                    for (int k = 0; k < 3; ++k) p[k] = q[k];
                }
            }
        }
    }
    double d;
    {
        double xp[3] = {x[0] - p[0], x[1] - p[1], x[2] - p[2]};
        d = sqrt(dot3(xp, xp));
        const double dmin = 1e-12 * sqrt(st.jac);
        d = d < dmin ? dmin : d;
    }
    double acc[NA];
#pragma unroll
    for (int k = 0; k < NA; ++k) acc[k] = 0.0;
    const int per_rt = n_t * n_r;
    if (live) {
#pragma unroll
        for (int ed = 0; ed < 3; ++ed) {
            const double *va = st.v[ed], *vb = st.v[(ed + 1) % 3];
            double a[3], b[3], cr3[3], ev[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                a[k] = va[k] - p[k];
                b[k] = vb[k] - p[k];
                ev[k] = vb[k] - va[k];
            }
            cross3(a, b, cr3);
            const double cr = dot3(cr3, nn);
            if (fabs(cr) <= 1e-14 * st.jac) continue;
            const double sgn = cr > 0 ? 1.0 : -1.0;
            // foot of the perpendicular from p on the edge line
            const double le = sqrt(dot3(ev, ev));
            const double t = -dot3(a, ev) / (le * le);
            double ef[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) ef[k] = a[k] + t * ev[k];
            const double hp = sqrt(dot3(ef, ef));
#pragma unroll
            for (int k = 0; k < 3; ++k) ef[k] /= hp;
            // right triangles (p, foot, v) for v = va (t_v = 0) and vb (t_v = 1)
#pragma unroll
            for (int side = 0; side < 2; ++side) {
                const double Lv = ((double)side - t) * le;
                if (Lv == 0.0) continue;
                const double sig = (side ? (1.0 - t) : t) > 0 ? 1.0 : -1.0;
                const double es = Lv > 0 ? 1.0 / le : -1.0 / le;
                double et[3];
#pragma unroll
                for (int k = 0; k < 3; ++k) et[k] = es * ev[k];
                const double S = asinh(fabs(Lv) / hp);
                for (int item = lane; item < per_rt; item += 64) {
                    const int it = item / n_r, ir = item % n_r;
                    // tan(phi) = sinh(s): rho_max = hp cosh(s), dphi = ds / cosh(s)
                    const double sv = S * tq[it];
                    const double chs = cosh(sv), shs = sinh(sv);
                    const double cphi = 1.0 / chs, sphi = shs / chs;
                    const double U = asinh(hp * chs / d);
                    const double u = U * rq[ir];
                    const double rho = d * sinh(u);
                    const double w = sgn * sig * (S * tw[it] / chs) * U * rw[ir] * d * cosh(u) * rho;
                    double yv[3];
#pragma unroll
                    for (int k = 0; k < 3; ++k)
                        yv[k] = p[k] + rho * (cphi * ef[k] + sphi * et[k]) - st.v[0][k];
                    const double h1 = dot3(yv, e1), h2 = dot3(yv, e2);
                    const double xi = (g22 * h1 - g12 * h2) / det;
                    const double eta = (g11 * h2 - g12 * h1) / det;
                    accumulate_pt<KID>(x, m, st, xi, eta, w, 0, c, acc);
                }
            }
        }
    }
#pragma unroll
    for (int k = 0; k < NA; ++k) {
        double v = acc[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        acc[k] = v;
    }
    if (!live || lane != 0) return;
    write_derotated(out + (size_t)pair * 27, clicks, 0, 1.0,
                    [&](int d1, int b, int d2) { return final_pt<KID>(acc, st, d1, b, d2); });
}

// out[i][c] = sum_k w[i][k] vals[k][i][c]: the limit (or log coefficient)
// of a field sampled at n_eps offsets, with per-point fit weights.
__global__ __launch_bounds__(256) void limit_combine_kernel(int n_eps, int64_t n_pts, int comps,
                                                            const double *__restrict__ w,
                                                            const double *__restrict__ vals,
                                                            double *__restrict__ out) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_pts * comps) return;
    const int64_t i = r / comps;
    double s = 0.0;
    for (int k = 0; k < n_eps; ++k) s += w[i * n_eps + k] * vals[(size_t)k * n_pts * comps + r];
    out[r] = s;
}

template <int KID, int GL>
static void launch_ip(hipStream_t st, int nq, const double *qx, const double *qw,
                      const double *obs_pts, const double *obs_ns, const double *src_pts,
                      const int32_t *src_tris, int n_pairs, const int32_t *pairs, int check0,
                      int finite_part, const ElasticConsts &c, double *out) {
    const int64_t threads = (int64_t)n_pairs * GL;
    hipLaunchKernelGGL((interior_pairs_kernel<KID, GL>), dim3((unsigned)ceil_div(threads, 256)),
                       dim3(256), 0, st, nq, qx, qw, obs_pts, obs_ns, src_pts, src_tris, n_pairs,
                       pairs, check0, finite_part, c, out);
}

template <int KID>
static void launch_ip_g(int gl, hipStream_t st, int nq, const double *qx, const double *qw,
                        const double *obs_pts, const double *obs_ns, const double *src_pts,
                        const int32_t *src_tris, int n_pairs, const int32_t *pairs, int check0,
                        int finite_part, const ElasticConsts &c, double *out) {
#define TCT_IP(GLV)                                                                           \
    launch_ip<KID, GLV>(st, nq, qx, qw, obs_pts, obs_ns, src_pts, src_tris, n_pairs, pairs,  \
                        check0, finite_part, c, out)
    switch (gl) {
    case 1: TCT_IP(1); break;
    case 4: TCT_IP(4); break;
    case 16: TCT_IP(16); break;
    default: TCT_IP(64); break;
    }
#undef TCT_IP
}

}  // namespace tct

using namespace tct;

extern "C" int tct_interior_pairs(int kernel, int64_t nq, const double *qx, const double *qw,
                                  const double *obs_pts, const double *obs_ns,
                                  const double *src_pts, const int32_t *src_tris, int64_t n_pairs,
                                  const int32_t *pairs, int check0, int finite_part,
                                  const double *params_host, int lanes_per_pair, double *out,
                                  void *stream) {
    TCT_CHECK_ARG(kernel >= TCT_U && kernel <= TCT_RT,
                  "interior_pairs: kernel %d has no point-observer form", kernel);
    TCT_CHECK_ARG(nq > 0 && nq < (1LL << 31), "interior_pairs: nq %lld", (long long)nq);
    TCT_CHECK_ARG(n_pairs >= 0 && n_pairs < (1LL << 26),
                  "interior_pairs: n_pairs %lld (max 2^26 per call)", (long long)n_pairs);
    TCT_CHECK_ARG(qx && qw && obs_pts && obs_ns && src_pts && src_tris && params_host && out &&
                      (pairs || n_pairs == 0),
                  "interior_pairs: null pointer");
    if (n_pairs == 0) return TCT_OK;
    int gl = lanes_per_pair;
    if (gl == 0) gl = nq >= 256 ? 64 : (nq >= 48 ? 16 : (nq >= 12 ? 4 : 1));
    TCT_CHECK_ARG(gl == 1 || gl == 4 || gl == 16 || gl == 64,
                  "interior_pairs: lanes_per_pair %d", gl);
    const ElasticConsts c = elastic_consts(params_host);
    hipStream_t st = as_stream(stream);
    const int np = (int)n_pairs, q = (int)nq;
#define TCT_IPK(KID)                                                                          \
    launch_ip_g<KID>(gl, st, q, qx, qw, obs_pts, obs_ns, src_pts, src_tris, np, pairs, check0, \
                     finite_part != 0, c, out)
    switch (kernel) {
    case TCT_U: TCT_IPK(TCT_U); break;
    case TCT_T: TCT_IPK(TCT_T); break;
    case TCT_A: TCT_IPK(TCT_A); break;
    case TCT_H: TCT_IPK(TCT_H); break;
    default: TCT_IPK(TCT_RT); break;
    }
#undef TCT_IPK
    TCT_CHECK_LAUNCH("interior_pairs");
    return TCT_OK;
}

extern "C" int tct_interior_corners(int64_t nq, const double *qt, const double *qw,
                                    const double *obs_ns, const double *src_pts,
                                    const int32_t *src_tris, int64_t n_pairs,
                                    const int32_t *pairs, const double *params_host, double *out,
                                    void *stream) {
    TCT_CHECK_ARG(nq > 0 && nq < (1LL << 31), "interior_corners: nq %lld", (long long)nq);
    TCT_CHECK_ARG(n_pairs >= 0 && n_pairs < (1LL << 31), "interior_corners: n_pairs %lld",
                  (long long)n_pairs);
    TCT_CHECK_ARG(qt && qw && obs_ns && src_pts && src_tris && params_host && out &&
                      (pairs || n_pairs == 0),
                  "interior_corners: null pointer");
    if (n_pairs == 0) return TCT_OK;
    hipLaunchKernelGGL(interior_corners_kernel, dim3((unsigned)ceil_div(n_pairs, 256)), dim3(256),
                       0, as_stream(stream), (int)nq, qt, qw, obs_ns, src_pts, src_tris,
                       (int)n_pairs, pairs, elastic_consts(params_host), out);
    TCT_CHECK_LAUNCH("interior_corners");
    return TCT_OK;
}

extern "C" int tct_bsr39_spmv(int64_t n_block_rows, const int32_t *indptr, const int32_t *indices,
                              const double *data, const double *x, double *y, int accumulate,
                              void *stream) {
    TCT_CHECK_ARG(n_block_rows >= 0 && n_block_rows < (1LL << 31), "bsr39: n_block_rows %lld",
                  (long long)n_block_rows);
    TCT_CHECK_ARG(indptr && x && y && (data || n_block_rows == 0), "bsr39: null pointer");
    if (n_block_rows == 0) return TCT_OK;
    hipLaunchKernelGGL(bsr39_kernel, dim3((unsigned)ceil_div(3 * n_block_rows, 256)), dim3(256),
                       0, as_stream(stream), n_block_rows, indptr, indices, data, x, y,
                       accumulate);
    TCT_CHECK_LAUNCH("bsr39_spmv");
    return TCT_OK;
}

extern "C" int tct_limit_combine(int n_eps, int64_t n_pts, int comps, const double *w,
                                 const double *vals, double *out, void *stream) {
    TCT_CHECK_ARG(n_eps > 0 && n_eps <= 64 && n_pts >= 0 && comps > 0 && comps <= 81,
                  "limit_combine: n_eps %d n_pts %lld comps %d", n_eps, (long long)n_pts, comps);
    TCT_CHECK_ARG(n_pts == 0 || (w && vals && out), "limit_combine: null pointer");
    if (n_pts == 0) return TCT_OK;
    hipLaunchKernelGGL(limit_combine_kernel, dim3((unsigned)ceil_div(n_pts * comps, 256)),
                       dim3(256), 0, as_stream(stream), n_eps, n_pts, comps, w, vals, out);
    TCT_CHECK_LAUNCH("limit_combine");
    return TCT_OK;
}

extern "C" int tct_interior_pairs_sinh(int kernel, int n_t, const double *tq, const double *tw,
                                       int n_r, const double *rq, const double *rw,
                                       const double *obs_pts, const double *obs_ns,
                                       const double *src_pts, const int32_t *src_tris,
                                       int64_t n_pairs, const int32_t *pairs,
                                       const double *params_host, double *out, void *stream) {
    TCT_CHECK_ARG(kernel >= TCT_U && kernel <= TCT_RT,
                  "interior_pairs_sinh: kernel %d has no point-observer form", kernel);
    TCT_CHECK_ARG(n_t > 0 && n_r > 0 && (int64_t)n_t * n_r < (1 << 24),
                  "interior_pairs_sinh: n_t %d n_r %d", n_t, n_r);
    TCT_CHECK_ARG(n_pairs >= 0 && n_pairs < (1LL << 26), "interior_pairs_sinh: n_pairs %lld",
                  (long long)n_pairs);
    TCT_CHECK_ARG(tq && tw && rq && rw && obs_pts && obs_ns && src_pts && src_tris &&
                      params_host && out && (pairs || n_pairs == 0),
                  "interior_pairs_sinh: null pointer");
    if (n_pairs == 0) return TCT_OK;
    const ElasticConsts c = elastic_consts(params_host);
    const dim3 grid((unsigned)ceil_div(n_pairs, 4)), block(256);
    hipStream_t st = as_stream(stream);
#define TCT_SK(KID)                                                                            \
    hipLaunchKernelGGL((interior_pairs_sinh_kernel<KID>), grid, block, 0, st, n_t, tq, tw, n_r, \
                       rq, rw, obs_pts, obs_ns, src_pts, src_tris, (int)n_pairs, pairs, c, out)
    switch (kernel) {
    case TCT_U: TCT_SK(TCT_U); break;
    case TCT_T: TCT_SK(TCT_T); break;
    case TCT_A: TCT_SK(TCT_A); break;
    case TCT_H: TCT_SK(TCT_H); break;
    default: TCT_SK(TCT_RT); break;
    }
#undef TCT_SK
    TCT_CHECK_LAUNCH("interior_pairs_sinh");
    return TCT_OK;
}
