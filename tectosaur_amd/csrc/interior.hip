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
template <typename F>
__device__ __forceinline__ void write_derotated(double *o, int clicks, int finite_part, double J,
                                                F value) {
#pragma unroll
    for (int b2 = 0; b2 < 3; ++b2) {
        const int src_b = pmod3(b2 - clicks);
        const bool keep = !finite_part || b2 == clicks;
#pragma unroll
        for (int d1 = 0; d1 < 3; ++d1)
#pragma unroll
            for (int d2 = 0; d2 < 3; ++d2)
                o[d1 * 9 + b2 * 3 + d2] = keep ? J * value(d1, src_b, d2) : 0.0;
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
