/*
 * tct_oracle.c -- CPU oracle for the farfield / nearfield hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP product
 * path is compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  Nothing in tectosaur_amd/
 * links, imports or calls it.
 *
 * It is a plain restatement, in FP64 with long-double accumulation, of the
 * reference algorithms.  Each function cites the reference file:line whose
 * semantics it follows.  It deliberately follows the reference's own
 * formulation (tensor code, vector code, loop order, r2 == 0 skips,
 * rotation / de-rotation) rather than the factored forms the HIP kernels
 * use, so that agreement between the two is a real check.
 *
 * Kernel ids: 0 elasticU3, 1 elasticT3, 2 elasticA3, 3 elasticH3,
 *             4 elasticRT3, 5 elasticRA3, 6 elasticRH3.
 *
 * Pinning status of each function is recorded in the header of
 * oracle/oracle.py, which is the only loader of this library.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

typedef long double acc_t;

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

enum { K_U = 0, K_T = 1, K_A = 2, K_H = 3, K_RT = 4, K_RA = 5, K_RH = 6 };

/* ------------------------------------------------------------------ */
/* Kernel tensors.  tectosaur/kernels/__init__.py:37-50 (U tensor),   */
/* :101-125 (T/A tensor), :174-214 (H tensor).  Karr[d_obs*3+d_src].  */
/* ------------------------------------------------------------------ */
static void tensor_U(const double *D, double r2, const double *params, double *K)
{
    double G = params[0], nu = params[1];
    double CsU0 = (3.0 - 4.0 * nu) / (G * 16.0 * M_PI * (1.0 - nu));
    double CsU1 = 1.0 / (G * 16.0 * M_PI * (1.0 - nu));
    double invr = 1.0 / sqrt(r2);
    double Q1 = CsU0 * invr;
    double Q2 = CsU1 * invr / r2;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            K[i * 3 + j] = Q2 * D[i] * D[j] + (i == j ? Q1 : 0.0);
}

/* isA = 0: elasticT3 (n = src normal); isA = 1: elasticA3 (n = obs normal) */
static void tensor_TA(int isA, const double *D, double r2, const double *n,
                      const double *params, double *K)
{
    double nu = params[1];
    double CsT0 = (1 - 2.0 * nu) / (8.0 * M_PI * (1.0 - nu));
    double CsT1 = 3.0 / (8.0 * M_PI * (1.0 - nu));
    double invr = 1.0 / sqrt(r2);
    double invr2 = invr * invr;
    double invr3 = invr2 * invr;
    double rn = n[0] * D[0] + n[1] * D[1] + n[2] * D[2];
    /* T: A = +CsT0 invr3, C = -CsT1 invr5;  A: the signs flip (:60-65). */
    double A = (isA ? -CsT0 : CsT0) * invr3;
    double C = (isA ? CsT1 : -CsT1) * invr3 * invr2;
    double nxdy = n[0] * D[1] - n[1] * D[0];
    double nzdx = n[2] * D[0] - n[0] * D[2];
    double nzdy = n[2] * D[1] - n[1] * D[2];
    double s = isA ? -1.0 : 1.0; /* minus_or_plus = -s, plus_or_minus = +s */
    K[0] = A * -rn + C * D[0] * rn * D[0];
    K[1] = A * (-s * nxdy) + C * D[0] * rn * D[1];
    K[2] = A * (s * nzdx) + C * D[0] * rn * D[2];
    K[3] = A * (s * nxdy) + C * D[1] * rn * D[0];
    K[4] = A * -rn + C * D[1] * rn * D[1];
    K[5] = A * (s * nzdy) + C * D[1] * rn * D[2];
    K[6] = A * (-s * nzdx) + C * D[2] * rn * D[0];
    K[7] = A * (-s * nzdy) + C * D[2] * rn * D[1];
    K[8] = A * -rn + C * D[2] * rn * D[2];
}

static void tensor_H(const double *D, double r2, const double *m, const double *n,
                     const double *params, double *K)
{
    double G = params[0], nu = params[1];
    double CsH0 = G / (4 * M_PI * (1 - nu));
    double CsH1 = 1 - 2 * nu;
    double CsH2 = -1 + 4 * nu;
    double CsH3 = 3 * nu;
    double invr = 1.0 / sqrt(r2);
    double invr2 = invr * invr;
    double invr3 = invr2 * invr;
    double Dor[3] = {invr * D[0], invr * D[1], invr * D[2]};
    double rn = n[0] * Dor[0] + n[1] * Dor[1] + n[2] * Dor[2];
    double rm = m[0] * Dor[0] + m[1] * Dor[1] + m[2] * Dor[2];
    double mn = m[0] * n[0] + m[1] * n[1] + m[2] * n[2];
    double Q = CsH0 * invr3;
    double A = Q * 3 * rn;
    double B = Q * CsH1;
    double C = Q * CsH3;
    double MT[3], NT[3], DT[3];
    for (int k = 0; k < 3; k++) {
        MT[k] = Q * CsH2 * n[k] + A * CsH1 * Dor[k];
        NT[k] = B * m[k] + C * Dor[k] * rm;
        DT[k] = B * 3 * n[k] * rm + C * Dor[k] * mn + A * (nu * m[k] - 5 * Dor[k] * rm);
    }
    double ST = A * nu * rm + B * mn;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            K[i * 3 + j] = n[i] * NT[j] + m[i] * MT[j] + Dor[i] * DT[j] + (i == j ? ST : 0.0);
}

/* Dispatch: obs normal m, src normal n. */
static void kernel_tensor(int kid, const double *D, double r2, const double *m,
                          const double *n, const double *params, double *K)
{
    switch (kid) {
    case K_U: tensor_U(D, r2, params, K); break;
    case K_T: tensor_TA(0, D, r2, n, params, K); break;
    case K_A: tensor_TA(1, D, r2, m, params, K); break;
    case K_H: tensor_H(D, r2, m, n, params, K); break;
    default: memset(K, 0, 9 * sizeof(double));
    }
}

/* ------------------------------------------------------------------ */
/* Vector code: out += K(D) . in, tectosaur/kernels/__init__.py:28-36  */
/* (U), :67-99 (T/A), :147-173 (H).                                    */
/* ------------------------------------------------------------------ */
static void vector_code(int kid, const double *D, double r2, const double *m,
                        const double *n, const double *in, const double *params,
                        double *sum)
{
    double Dx = D[0], Dy = D[1], Dz = D[2];
    double inx = in[0], iny = in[1], inz = in[2];
    if (kid == K_U) {
        double G = params[0], nu = params[1];
        double CsU0 = (3.0 - 4.0 * nu) / (G * 16.0 * M_PI * (1.0 - nu));
        double CsU1 = 1.0 / (G * 16.0 * M_PI * (1.0 - nu));
        double invr = 1.0 / sqrt(r2);
        double Q1 = CsU0 * invr;
        double Q2 = CsU1 * invr / r2;
        double ddi = Dx * inx + Dy * iny + Dz * inz;
        sum[0] += Q1 * inx + Q2 * Dx * ddi;
        sum[1] += Q1 * iny + Q2 * Dy * ddi;
        sum[2] += Q1 * inz + Q2 * Dz * ddi;
    } else if (kid == K_T || kid == K_A) {
        int isA = kid == K_A;
        const double *N = isA ? m : n;
        double nu = params[1];
        double CsT0 = (1 - 2.0 * nu) / (8.0 * M_PI * (1.0 - nu));
        double CsT1 = 3.0 / (8.0 * M_PI * (1.0 - nu));
        double invr = 1.0 / sqrt(r2);
        double invr2 = invr * invr;
        double invr3 = invr2 * invr;
        double rn = N[0] * Dx + N[1] * Dy + N[2] * Dz;
        double s = isA ? -1.0 : 1.0;
        double A = s * CsT0 * invr3;
        double C = -s * CsT1 * invr3 * invr2;
        double rnddi = C * rn * (Dx * inx + Dy * iny + Dz * inz);
        double nxdy = N[0] * Dy - N[1] * Dx;
        double nzdx = N[2] * Dx - N[0] * Dz;
        double nzdy = N[2] * Dy - N[1] * Dz;
        sum[0] += A * (-rn * inx - s * nxdy * iny + s * nzdx * inz) + Dx * rnddi;
        sum[1] += A * (s * nxdy * inx - rn * iny + s * nzdy * inz) + Dy * rnddi;
        sum[2] += A * (-s * nzdx * inx - s * nzdy * iny - rn * inz) + Dz * rnddi;
    } else if (kid == K_H) {
        double G = params[0], nu = params[1];
        double CsH0 = G / (4 * M_PI * (1 - nu));
        double CsH1 = 1 - 2 * nu;
        double CsH2 = -1 + 4 * nu;
        double CsH3 = 3 * nu;
        double invr = 1.0 / sqrt(r2);
        double invr2 = invr * invr;
        double invr3 = invr2 * invr;
        double rn = invr * (n[0] * Dx + n[1] * Dy + n[2] * Dz);
        double rm = invr * (m[0] * Dx + m[1] * Dy + m[2] * Dz);
        double mn = m[0] * n[0] + m[1] * n[1] + m[2] * n[2];
        double sn = inx * n[0] + iny * n[1] + inz * n[2];
        double sd = invr * (inx * Dx + iny * Dy + inz * Dz);
        double sm = inx * m[0] + iny * m[1] + inz * m[2];
        double Q = CsH0 * invr3;
        double A = Q * 3 * rn;
        double B = Q * CsH1;
        double C = Q * CsH3;
        double MT = Q * CsH2 * sn + A * CsH1 * sd;
        double NT = B * sm + C * sd * rm;
        double DT = invr * (B * 3 * sn * rm + C * sd * mn + A * (nu * sm - 5 * sd * rm));
        double ST = A * nu * rm + B * mn;
        sum[0] += n[0] * NT + m[0] * MT + Dx * DT + ST * inx;
        sum[1] += n[1] * NT + m[1] * MT + Dy * DT + ST * iny;
        sum[2] += n[2] * NT + m[2] * MT + Dz * DT + ST * inz;
    }
}

/* Public single-point kernel evaluation (used to pin against the sympy
 * derivation): K[9] for displacement D = src - obs. */
void or_kernel_tensor(int kid, const double *D, const double *m, const double *n,
                      const double *params, double *K)
{
    double r2 = D[0] * D[0] + D[1] * D[1] + D[2] * D[2];
    kernel_tensor(kid, D, r2, m, n, params, K);
}

void or_kernel_vector(int kid, const double *D, const double *m, const double *n,
                      const double *in, const double *params, double *out)
{
    double r2 = D[0] * D[0] + D[1] * D[1] + D[2] * D[2];
    out[0] = out[1] = out[2] = 0.0;
    vector_code(kid, D, r2, m, n, in, params, out);
}

/* ------------------------------------------------------------------ */
/* a1: point-to-point N-body, tectosaur/farfield_direct.cl:15-113.    */
/* out[i*3+d] = sum_j K(src_j - obs_i) . in_j, skipping r2 == 0.       */
/* rows: optional list of obs indices to evaluate (n_rows of them);     */
/* out then holds n_rows*3 values.  rows == NULL evaluates all obs.    */
/* ------------------------------------------------------------------ */
void or_farfield_pts(int kid, int64_t n_obs, const double *obs_pts, const double *obs_ns,
                     int64_t n_src, const double *src_pts, const double *src_ns,
                     const double *in, const double *params,
                     const int64_t *rows, int64_t n_rows, double *out)
{
    int64_t n_eval = rows ? n_rows : n_obs;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t r = 0; r < n_eval; r++) {
        int64_t i = rows ? rows[r] : r;
        const double *x = obs_pts + 3 * i;
        const double *m = obs_ns + 3 * i;
        acc_t acc[3] = {0, 0, 0};
        for (int64_t j = 0; j < n_src; j++) {
            const double *y = src_pts + 3 * j;
            double D[3] = {y[0] - x[0], y[1] - x[1], y[2] - x[2]};
            double r2 = D[0] * D[0];
            r2 += D[1] * D[1];
            r2 += D[2] * D[2];
            if (r2 == 0.0)
                continue;
            double s[3] = {0, 0, 0};
            vector_code(kid, D, r2, m, src_ns + 3 * j, in + 3 * j, params, s);
            acc[0] += s[0];
            acc[1] += s[1];
            acc[2] += s[2];
        }
        out[3 * r + 0] = (double)acc[0];
        out[3 * r + 1] = (double)acc[1];
        out[3 * r + 2] = (double)acc[2];
    }
}

/* ------------------------------------------------------------------ */
/* Triangle geometry, tectosaur/integral_primitives.cl:56-118.         */
/* ------------------------------------------------------------------ */
typedef struct {
    double v[3][3];
    double unscaled_n[3];
    double jac;
    double n[3];
    double curl[3][3]; /* curl[basis][component] */
} tri_t;

static const double basis_gradient[3][2] = {{-1.0, -1.0}, {1.0, 0.0}, {0.0, 1.0}};

static int pmod3(int i) { return ((i % 3) + 3) % 3; }

static void tri_info(tri_t *t, const double *pts, const int32_t *tris, int64_t idx,
                     int clicks, int flip)
{
    for (int c = 0; c < 3; c++) {
        int fc = c;
        if (flip) {
            if (c == 0) fc = 1;
            else if (c == 1) fc = 0;
        }
        int64_t p = tris[3 * idx + pmod3(fc + clicks)];
        for (int d = 0; d < 3; d++)
            t->v[c][d] = pts[3 * p + d];
    }
    double s20[3], s21[3];
    for (int d = 0; d < 3; d++) {
        s20[d] = t->v[2][d] - t->v[0][d];
        s21[d] = t->v[2][d] - t->v[1][d];
    }
    t->unscaled_n[0] = s20[1] * s21[2] - s20[2] * s21[1];
    t->unscaled_n[1] = s20[2] * s21[0] - s20[0] * s21[2];
    t->unscaled_n[2] = s20[0] * s21[1] - s20[1] * s21[0];
    if (flip)
        for (int d = 0; d < 3; d++)
            t->unscaled_n[d] *= -1;
    t->jac = sqrt(t->unscaled_n[0] * t->unscaled_n[0] + t->unscaled_n[1] * t->unscaled_n[1] +
                  t->unscaled_n[2] * t->unscaled_n[2]);
    for (int d = 0; d < 3; d++)
        t->n[d] = t->unscaled_n[d] / t->jac;
    /* surf_curl_basis, integral_primitives.cl:95-112 */
    double g1[3], g2[3];
    for (int d = 0; d < 3; d++) {
        g1[d] = t->v[1][d] - t->v[0][d];
        g2[d] = t->v[2][d] - t->v[0][d];
    }
    for (int b = 0; b < 3; b++)
        for (int s = 0; s < 3; s++)
            t->curl[b][s] = (basis_gradient[b][0] * g2[s] - basis_gradient[b][1] * g1[s]) / t->jac;
}

static void basis3(double xh, double yh, double *b)
{
    b[0] = 1 - xh - yh;
    b[1] = xh;
    b[2] = yh;
}

static void interp(const tri_t *t, const double *b, double *x)
{
    for (int d = 0; d < 3; d++)
        x[d] = b[0] * t->v[0][d] + b[1] * t->v[1][d] + b[2] * t->v[2][d];
}

/* Levi-Civita as generated at tectosaur/kernels/kernels.cl:4-5 */
static int levi(int i, int j, int k) { return (i - j) * (j - k) * (k - i) / 2; }

/* ------------------------------------------------------------------ */
/* Regularized (surface-curl) kernel constants, kernels/__init__.py   */
/* :217-258.                                                           */
/* ------------------------------------------------------------------ */
typedef struct {
    double CsRT0, CsRT1, CsRT2, CsRH0, CsRH1, CsRH2;
} rconst_t;

static rconst_t rconsts(const double *params)
{
    rconst_t c;
    double G = params[0], nu = params[1];
    c.CsRT0 = -1.0 / (8.0 * M_PI * (1.0 - nu));
    c.CsRT1 = -(1 - 2 * nu) / (8.0 * M_PI * (1.0 - nu));
    c.CsRT2 = -1.0 / (4 * M_PI);
    c.CsRH0 = G / (8.0 * M_PI * (1 - nu));
    c.CsRH1 = 2 * (1 - nu);
    c.CsRH2 = 4 * nu;
    return c;
}

/* Gimk, tectosaur/kernels/kernels.cl:20-27 */
static double Gimk(int p, int mm, int j, double Q1, double Q2, const double *D)
{
    double A = Q1 * levi(mm, p, j);
    double B = 0.0;
    for (int It = 0; It < 3; It++)
        B += levi(mm, It, j) * D[It];
    return A + Q2 * D[p] * B;
}

/* ------------------------------------------------------------------ */
/* a5 inner integrand: integral_primitives.cl:146-213 (call_tensor_code */
/* and integrate_pair) plus the regularized tensors kernels.cl:29-98.   */
/* Accumulates into res[81] (layout b_obs*27 + d_obs*9 + b_src*3 + d_src). */
/* ------------------------------------------------------------------ */
static void pair_integrand(int kid, const tri_t *ot, const tri_t *st, const double *q, double w,
                           int check0, const double *params, acc_t *res)
{
    double obsb[3], srcb[3], x[3], y[3];
    basis3(q[0], q[1], obsb);
    basis3(q[2], q[3], srcb);
    interp(ot, obsb, x);
    interp(st, srcb, y);
    double D[3] = {y[0] - x[0], y[1] - x[1], y[2] - x[2]};
    double r2 = D[0] * D[0] + D[1] * D[1] + D[2] * D[2];
    if (check0 && r2 == 0.0)
        return;
    if (kid <= K_H) {
        double K[9];
        kernel_tensor(kid, D, r2, ot->n, st->n, params, K);
        for (int b = 0; b < 3; b++)
            obsb[b] *= w;
        int idx = 0;
        for (int bo = 0; bo < 3; bo++)
            for (int dob = 0; dob < 3; dob++)
                for (int bs = 0; bs < 3; bs++)
                    for (int ds = 0; ds < 3; ds++, idx++)
                        res[idx] += obsb[bo] * srcb[bs] * K[dob * 3 + ds];
        return;
    }
    rconst_t c = rconsts(params);
    if (kid == K_RT || kid == K_RA) {
        /* elasticRTA3_setup / elasticRTA3_tensor, kernels.cl:11-68 */
        int isRA = kid == K_RA;
        const double *nn = isRA ? ot->n : st->n;
        double invr = 1.0 / sqrt(r2);
        double invr3 = invr * invr * invr;
        double Q1 = c.CsRT1 * invr;
        double Q2 = c.CsRT0 * invr3;
        double Q3 = c.CsRT2 * invr3;
        double Q3nD = Q3 * (nn[0] * D[0] + nn[1] * D[1] + nn[2] * D[2]);
        /* L = obs, R = src for RT; L = src, R = obs for RA */
        for (int dob = 0; dob < 3; dob++) {
            for (int ds = 0; ds < 3; ds++) {
                int L_d = isRA ? ds : dob;
                int R_d = isRA ? dob : ds;
                for (int bR = 0; bR < 3; bR++) {
                    double Kval = 0.0;
                    for (int Ij = 0; Ij < 3; Ij++) {
                        double g = Gimk(L_d, Ij, R_d, Q1, Q2, D);
                        const double *curlR = isRA ? ot->curl[bR] : st->curl[bR];
                        Kval += g * curlR[Ij];
                    }
                    if (dob == ds) {
                        double Rb = isRA ? obsb[bR] : srcb[bR];
                        Kval += Q3nD * Rb * (isRA ? -1 : 1);
                    }
                    for (int bL = 0; bL < 3; bL++) {
                        double Lb = isRA ? srcb[bL] : obsb[bL];
                        double val = w * Lb * Kval;
                        int bo = isRA ? bR : bL;
                        int bs = isRA ? bL : bR;
                        res[bo * 27 + dob * 9 + bs * 3 + ds] += val;
                    }
                }
            }
        }
        return;
    }
    /* kid == K_RH: elasticRH3_tensor, kernels.cl:70-98 */
    {
        double invr2 = 1.0 / r2;
        double invr = sqrt(invr2);
        for (int dob = 0; dob < 3; dob++)
            for (int ds = 0; ds < 3; ds++)
                for (int It = 0; It < 3; It++)
                    for (int Im = 0; Im < 3; Im++) {
                        double Kval = w * c.CsRH0 * invr *
                                      (-2 * D[dob] * D[ds] * invr2 * (It == Im) +
                                       c.CsRH1 * (dob == It) * (ds == Im) +
                                       c.CsRH2 * (dob == Im) * (ds == It) -
                                       2 * (dob == ds) * (It == Im));
                        for (int bs = 0; bs < 3; bs++)
                            for (int bo = 0; bo < 3; bo++)
                                res[bo * 27 + dob * 9 + bs * 3 + ds] +=
                                    Kval * ot->curl[bo][It] * st->curl[bs][Im];
                    }
    }
}

/* ------------------------------------------------------------------ */
/* a5: single_pairs{Z,N} and single_pairsN_adj, tectosaur/assemble.cl  */
/* :24-103.  pairs has pair_width columns: 2 (obs, src) or 5           */
/* (obs, src, obs_clicks, src_clicks, src_flip).  For width 5 the       */
/* result is de-rotated exactly as assemble.cl:77-101 does.            */
/* out[n_pairs*81].                                                     */
/* ------------------------------------------------------------------ */
void or_pairs_integrate(int kid, int64_t n_q, const double *qx, const double *qw,
                        const double *pts, const int32_t *tris, int64_t n_pairs,
                        const int64_t *pairs, int pair_width, int check0,
                        const double *params, double *out)
{
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t p = 0; p < n_pairs; p++) {
        const int64_t *pr = pairs + p * pair_width;
        int oc = 0, sc = 0, flip = 0;
        if (pair_width == 5) {
            oc = (int)pr[2];
            sc = (int)pr[3];
            flip = pr[4] != 0;
        }
        tri_t ot, st;
        tri_info(&ot, pts, tris, pr[0], oc, 0);
        tri_info(&st, pts, tris, pr[1], sc, flip);
        acc_t res[81];
        for (int k = 0; k < 81; k++)
            res[k] = 0;
        for (int64_t iq = 0; iq < n_q; iq++)
            pair_integrand(kid, &ot, &st, qx + 4 * iq, qw[iq], check0, params, res);
        double J = ot.jac * st.jac;
        double *o = out + 81 * p;
        if (pair_width != 5) {
            for (int k = 0; k < 81; k++)
                o[k] = J * (double)res[k];
            continue;
        }
        int od[3], sd[3];
        for (int i = 0; i < 3; i++) {
            od[i] = pmod3(-oc + i);
            sd[i] = pmod3(-sc + i);
        }
        if (flip) {
            int tmp = sd[0];
            sd[0] = sd[1];
            sd[1] = tmp;
        }
        for (int b1 = 0; b1 < 3; b1++)
            for (int d1 = 0; d1 < 3; d1++)
                for (int b2 = 0; b2 < 3; b2++)
                    for (int d2 = 0; d2 < 3; d2++)
                        o[b1 * 27 + d1 * 9 + b2 * 3 + d2] =
                            J * (double)res[od[b1] * 27 + d1 * 9 + sd[b2] * 3 + d2];
    }
}

/* ------------------------------------------------------------------ */
/* a2: farfield_tris_to_tris{K}, tectosaur/matrix_free.cl:19-107, with  */
/* the surface-curl vector codes of tectosaur/kernels/kernels.cl:100-206. */
/* qx[n_q*2], qw[n_q]: the 2-D triangle rule.  in[n_src*9] ->            */
/* out[n_obs_eval*9].  rows: optional obs-triangle subset.              */
/* ------------------------------------------------------------------ */
void or_farfield_tris(int kid, int64_t n_q, const double *qx, const double *qw,
                      const double *pts, int64_t n_obs, const int32_t *obs_tris,
                      int64_t n_src, const int32_t *src_tris, const double *in,
                      const double *params, const int64_t *rows, int64_t n_rows, double *out)
{
    int64_t n_eval = rows ? n_rows : n_obs;
    rconst_t c = rconsts(params);
    (void)n_obs;
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t r = 0; r < n_eval; r++) {
        int64_t oi = rows ? rows[r] : r;
        tri_t ot;
        tri_info(&ot, pts, obs_tris, oi, 0, 0);
        acc_t sum[9];
        for (int k = 0; k < 9; k++)
            sum[k] = 0;
        for (int64_t sj = 0; sj < n_src; sj++) {
            tri_t st;
            tri_info(&st, pts, src_tris, sj, 0, 0);
            const double *inp = in + 9 * sj;
            for (int64_t q1 = 0; q1 < n_q; q1++) {
                double obsb[3], x[3];
                basis3(qx[2 * q1], qx[2 * q1 + 1], obsb);
                interp(&ot, obsb, x);
                for (int64_t q2 = 0; q2 < n_q; q2++) {
                    double srcb[3], y[3];
                    basis3(qx[2 * q2], qx[2 * q2 + 1], srcb);
                    interp(&st, srcb, y);
                    double quadw = qw[q1] * qw[q2];
                    double D[3] = {y[0] - x[0], y[1] - x[1], y[2] - x[2]};
                    double r2 = D[0] * D[0] + D[1] * D[1] + D[2] * D[2];
                    if (r2 == 0.0)
                        continue;
                    double factor = ot.jac * st.jac * quadw;
                    double lsum[3] = {0, 0, 0}, inq[3] = {0, 0, 0};
                    for (int d = 0; d < 3; d++)
                        for (int b = 0; b < 3; b++)
                            inq[d] += inp[b * 3 + d] * srcb[b];
                    if (kid <= K_H) {
                        vector_code(kid, D, r2, ot.n, st.n, inq, params, lsum);
                    } else if (kid == K_RT) {
                        /* elasticRT3_vector, kernels.cl:100-130 */
                        double ssc[3][3];
                        for (int Ij = 0; Ij < 3; Ij++)
                            for (int d = 0; d < 3; d++) {
                                ssc[d][Ij] = 0.0;
                                for (int b = 0; b < 3; b++)
                                    ssc[d][Ij] += st.curl[b][Ij] * inp[b * 3 + d];
                            }
                        double invr = 1.0 / sqrt(r2);
                        double invr3 = invr * invr * invr;
                        double Q1 = c.CsRT1 * invr, Q2 = c.CsRT0 * invr3, Q3 = c.CsRT2 * invr3;
                        double Q3nD = Q3 * (st.n[0] * D[0] + st.n[1] * D[1] + st.n[2] * D[2]);
                        for (int dob = 0; dob < 3; dob++)
                            for (int ds = 0; ds < 3; ds++)
                                for (int Ij = 0; Ij < 3; Ij++)
                                    lsum[dob] += Gimk(dob, Ij, ds, Q1, Q2, D) * ssc[ds][Ij];
                        for (int d = 0; d < 3; d++)
                            lsum[d] += Q3nD * inq[d];
                    } else if (kid == K_RA) {
                        /* elasticRA3_vector, kernels.cl:132-152 */
                        double invr = 1.0 / sqrt(r2);
                        double invr3 = invr * invr * invr;
                        double Q1 = c.CsRT1 * invr, Q2 = c.CsRT0 * invr3, Q3 = c.CsRT2 * invr3;
                        double Q3nD = Q3 * (ot.n[0] * D[0] + ot.n[1] * D[1] + ot.n[2] * D[2]);
                        for (int dob = 0; dob < 3; dob++)
                            for (int ds = 0; ds < 3; ds++)
                                for (int Ij = 0; Ij < 3; Ij++) {
                                    double Kval = factor * Gimk(ds, Ij, dob, Q1, Q2, D) * inq[ds];
                                    for (int bo = 0; bo < 3; bo++)
                                        sum[bo * 3 + dob] += Kval * ot.curl[bo][Ij];
                                }
                        for (int d = 0; d < 3; d++)
                            lsum[d] -= Q3nD * inq[d];
                    } else if (kid == K_RH) {
                        /* elasticRH3_vector, kernels.cl:154-206 */
                        double invr2 = 1.0 / r2;
                        double invr = sqrt(invr2);
                        double F = factor * c.CsRH0 * invr;
                        double ssc[3][3];
                        for (int Ij = 0; Ij < 3; Ij++)
                            for (int d = 0; d < 3; d++) {
                                ssc[d][Ij] = 0.0;
                                for (int b = 0; b < 3; b++)
                                    ssc[d][Ij] += st.curl[b][Ij] * inp[b * 3 + d];
                            }
                        double B = (ssc[0][0] + ssc[1][1] + ssc[2][2]) * c.CsRH1;
                        for (int bo = 0; bo < 3; bo++) {
                            const double *oc = ot.curl[bo];
                            double A = 0.0;
                            for (int ds = 0; ds < 3; ds++)
                                A += D[ds] * (ssc[ds][0] * oc[0] + ssc[ds][1] * oc[1] + ssc[ds][2] * oc[2]);
                            for (int dob = 0; dob < 3; dob++) {
                                double C = ssc[0][dob] * oc[0] + ssc[1][dob] * oc[1] + ssc[2][dob] * oc[2];
                                double Dd = ssc[dob][0] * oc[0] + ssc[dob][1] * oc[1] + ssc[dob][2] * oc[2];
                                sum[bo * 3 + dob] += F * (-2 * D[dob] * invr2 * A + oc[dob] * B +
                                                          c.CsRH2 * C - 2 * Dd);
                            }
                        }
                    }
                    for (int bo = 0; bo < 3; bo++)
                        for (int dob = 0; dob < 3; dob++)
                            sum[bo * 3 + dob] += factor * obsb[bo] * lsum[dob];
                }
            }
        }
        for (int k = 0; k < 9; k++)
            out[9 * r + k] = (double)sum[k];
    }
}

/* ------------------------------------------------------------------ */
/* a7: brute-force near-pair test with the exact inclusion predicate of  */
/* tectosaur/mesh/fast_find_nearfield.cpp:76-83:                          */
/*   include (i, j) iff dist2(c_i, c_j) <= pow((r_i + r_j) * thr, 2).    */
/* Returns the number of pairs; writes up to cap pairs into out[2*k].    */
/* ------------------------------------------------------------------ */
int64_t or_near_pairs_bruteforce(int64_t n_a, const double *ca, const double *ra,
                                 int64_t n_b, const double *cb, const double *rb,
                                 double thr, int64_t *out, int64_t cap)
{
    int64_t k = 0;
    for (int64_t i = 0; i < n_a; i++)
        for (int64_t j = 0; j < n_b; j++) {
            double d2 = 0.0;
            for (int d = 0; d < 3; d++) {
                double t = ca[3 * i + d] - cb[3 * j + d];
                d2 += t * t;
            }
            double lim = pow((ra[i] + rb[j]) * thr, 2);
            if (d2 > lim)
                continue;
            if (k < cap) {
                out[2 * k] = i;
                out[2 * k + 1] = j;
            }
            k++;
        }
    return k;
}

/* ------------------------------------------------------------------ */
/* a8: BCOO 9x9 block matvec, tectosaur/util/fast_sparse.cpp:128-163.   */
/* y[rows[k]*bs + i] += sum_j data[k][i][j] x[cols[k]*bs + j]; y is       */
/* accumulated (caller zeroes it).                                       */
/* ------------------------------------------------------------------ */
void or_bcoo_matvec(int64_t n_blocks, int bs, const int64_t *rows, const int64_t *cols,
                    const double *data, const double *x, double *y)
{
    for (int64_t k = 0; k < n_blocks; k++) {
        const double *A = data + (int64_t)bs * bs * k;
        for (int i = 0; i < bs; i++) {
            acc_t s = 0;
            for (int j = 0; j < bs; j++)
                s += (acc_t)A[i * bs + j] * x[cols[k] * bs + j];
            y[rows[k] * bs + i] += (double)s;
        }
    }
}

/* ------------------------------------------------------------------ */
/* Interior evaluation (point observers).                              */
/* interior_pairs, tectosaur/assemble.cl:140-186, through               */
/* integrate_pt_tri (tectosaur/integral_primitives.cl:216-254): the      */
/* tensor code with one observation "basis" obsb = {1}; for RT the        */
/* regularized tensor elasticRTA3_tensor('obs', 'src', 1)               */
/* (tectosaur/kernels/kernels.cl:29-62).  pairs[n][3] = (obs point,     */
/* src triangle, src rotation clicks).  check0 skips r2 == 0, as the    */
/* point farfield farfield_tris_to_pts does (matrix_free.cl:165-167);   */
/* interior_pairs itself does not.  finite_part writes only column      */
/* b_src = clicks (assemble.cl:164-172), the rest stay zero.            */
/* out[n*27], index d_obs*9 + b_src*3 + d_src, times J_src.              */
/* ------------------------------------------------------------------ */
static void pt_integrand(int kid, const double *x, const double *m, const tri_t *st,
                         const double *q, double w, int check0, const double *params,
                         acc_t *res)
{
    double srcb[3], y[3];
    basis3(q[0], q[1], srcb);
    interp(st, srcb, y);
    double D[3] = {y[0] - x[0], y[1] - x[1], y[2] - x[2]};
    double r2 = D[0] * D[0] + D[1] * D[1] + D[2] * D[2];
    if (check0 && r2 == 0.0)
        return;
    if (kid <= K_H) {
        double K[9];
        kernel_tensor(kid, D, r2, m, st->n, params, K);
        for (int dob = 0; dob < 3; dob++)
            for (int bs = 0; bs < 3; bs++)
                for (int ds = 0; ds < 3; ds++)
                    res[dob * 9 + bs * 3 + ds] += w * srcb[bs] * K[dob * 3 + ds];
        return;
    }
    /* RT */
    rconst_t c = rconsts(params);
    double invr = 1.0 / sqrt(r2);
    double invr3 = invr * invr * invr;
    double Q1 = c.CsRT1 * invr;
    double Q2 = c.CsRT0 * invr3;
    double Q3nD = c.CsRT2 * invr3 * (st->n[0] * D[0] + st->n[1] * D[1] + st->n[2] * D[2]);
    for (int dob = 0; dob < 3; dob++)
        for (int ds = 0; ds < 3; ds++)
            for (int bs = 0; bs < 3; bs++) {
                double Kval = 0.0;
                for (int Ij = 0; Ij < 3; Ij++)
                    Kval += Gimk(dob, Ij, ds, Q1, Q2, D) * st->curl[bs][Ij];
                if (dob == ds)
                    Kval += Q3nD * srcb[bs];
                res[dob * 9 + bs * 3 + ds] += w * Kval;
            }
}

static void pt_derotate(const acc_t *res, double J, int clicks, int finite_part, double *o)
{
    int derot[3];
    for (int i = 0; i < 3; i++)
        derot[i] = pmod3(-clicks + i);
    for (int k = 0; k < 27; k++)
        o[k] = 0.0;
    for (int d1 = 0; d1 < 3; d1++)
        for (int b2 = 0; b2 < 3; b2++) {
            if (finite_part && b2 != clicks)
                continue;
            for (int d2 = 0; d2 < 3; d2++)
                o[d1 * 9 + b2 * 3 + d2] = J * (double)res[d1 * 9 + derot[b2] * 3 + d2];
        }
}

void or_interior_pairs(int kid, int64_t n_q, const double *qx, const double *qw,
                       const double *obs_pts, const double *obs_ns, const double *src_pts,
                       const int32_t *src_tris, int64_t n_pairs, const int64_t *pairs,
                       int check0, int finite_part, const double *params, double *out)
{
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t p = 0; p < n_pairs; p++) {
        const int64_t *pr = pairs + 3 * p;
        int clicks = (int)pr[2];
        tri_t st;
        tri_info(&st, src_pts, src_tris, pr[1], clicks, 0);
        const double *x = obs_pts + 3 * pr[0];
        const double *m = obs_ns + 3 * pr[0];
        acc_t res[27];
        for (int k = 0; k < 27; k++)
            res[k] = 0;
        for (int64_t iq = 0; iq < n_q; iq++)
            pt_integrand(kid, x, m, &st, qx + 2 * iq, qw[iq], check0, params, res);
        pt_derotate(res, st.jac, clicks, finite_part, out + 27 * p);
    }
}

/* ------------------------------------------------------------------ */
/* interior_corners, tectosaur/interior_corners.cl:23-175: observation  */
/* point at vertex `clicks` of the source triangle, a 1-D rule in the   */
/* angle t over [0, pi/2] (interior.py:66-69) and the radial integrals  */
/* I1, I2, I3 of that file in closed form.  The kernel is elasticH3's   */
/* constants whatever operator the caller builds (the .cl file fixes    */
/* K = elasticH3), and only the terms left active there are summed:     */
/* CsH0 CsH3 mn T1 + delta CsH0 CsH1 mn T0.  The I4 entries of that     */
/* file are written as `1/3*(...)` and `-1/3*(...)`; C integer division */
/* makes both factors zero, so I4 = 0 there and here.                   */
/* pairs[n][3] = (obs point, src triangle, vertex index); out[n*27].     */
/* ------------------------------------------------------------------ */

/* The radial integrals of interior_corners.cl:79-83 for one angle: a is */
/* the in-plane direction v1 cos t + v2 sin t, rhomax = 1/(cos t+sin t), */
/* n the source normal.  out = [I1, I2, I3 (3x3 row-major), I4 (3x3)].   */
/* I4 is the reference's value, zero (its `1/3*(...)` factors are C     */
/* integer divisions); n enters only through that zeroed I4.            */
void or_corner_radial(const double *a, double rhomax, const double *n, double *out)
{
    (void)n;
    double A2 = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
    double A = sqrt(A2);
    double A3 = A2 * A, A5 = A3 * A2;
    out[0] = -1.0 / (A3 * rhomax);
    out[1] = (log(2.0) + log(A) + log(rhomax) - 1.0) / A3;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            out[2 + 3 * i + j] = -a[i] * a[j] / (A5 * rhomax);
            out[11 + 3 * i + j] = 0.0;
        }
}

void or_interior_corners(int64_t n_q, const double *qt, const double *qw,
                         const double *obs_ns, const double *src_pts,
                         const int32_t *src_tris, int64_t n_pairs, const int64_t *pairs,
                         const double *params, double *out)
{
    double G = params[0], nu = params[1];
    double CsH0 = G / (4 * M_PI * (1 - nu));
    double CsH1 = 1 - 2 * nu;
    double CsH3 = 3 * nu;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t p = 0; p < n_pairs; p++) {
        const int64_t *pr = pairs + 3 * p;
        int clicks = (int)pr[2];
        tri_t st;
        tri_info(&st, src_pts, src_tris, pr[1], clicks, 0);
        const double *m = obs_ns + 3 * pr[0];
        double v1[3], v2[3];
        for (int d = 0; d < 3; d++) {
            v1[d] = st.v[1][d] - st.v[0][d];
            v2[d] = st.v[2][d] - st.v[0][d];
        }
        double mn = m[0] * st.n[0] + m[1] * st.n[1] + m[2] * st.n[2];
        acc_t res[27];
        for (int k = 0; k < 27; k++)
            res[k] = 0;
        for (int64_t iq = 0; iq < n_q; iq++) {
            double t = qt[iq], w = qw[iq];
            double ct = cos(t), s = sin(t);
            double rhomax = 1.0 / (ct + s);
            double a[3], I[20];
            for (int d = 0; d < 3; d++)
                a[d] = v1[d] * ct + v2[d] * s;
            or_corner_radial(a, rhomax, st.n, I);
            double I1 = I[0], I2 = I[1];
            double T0[3] = {-I2 * ct - I2 * s + I1, I2 * ct, I2 * s};
            for (int dob = 0; dob < 3; dob++)
                for (int b = 0; b < 3; b++)
                    for (int ds = 0; ds < 3; ds++) {
                        double I3 = I[2 + 3 * dob + ds], I4 = I[11 + 3 * dob + ds];
                        double T1[3] = {-I4 * ct - I4 * s + I3, I4 * ct, I4 * s};
                        double K = CsH0 * CsH3 * T1[b] * mn;
                        if (dob == ds)
                            K += CsH0 * CsH1 * mn * T0[b];
                        res[dob * 9 + b * 3 + ds] += w * K;
                    }
        }
        pt_derotate(res, st.jac, clicks, 0, out + 27 * p);
    }
}

/* ------------------------------------------------------------------ */
/* Near-singular point-triangle integrals by a polar sinh rule.         */
/* No reference counterpart: the reference integrates close pairs with  */
/* a fixed triangle Gauss rule.  The source triangle is split at a      */
/* point p (below) into three sub-                                      */
/* triangles (p, v_k, v_k+1), and each of those at the foot of the     */
/* perpendicular from p into two right triangles (signed when the foot  */
/* falls outside the edge).  A right triangle with legs h (p to foot)   */
/* and L is integrated in polar coordinates about p: the angle through  */
/* tan(phi) = sinh(s), s in [0, asinh(L/h)], so rho_max = h cosh(s) and */
/* the thin-triangle end stays smooth; the radius through rho =         */
/* d sinh(u) with d = |x - p|, the sinh transformation of Johnston and  */
/* Elliott (Int. J. Numer. Meth. Engng 62 (2005) 564-578), which        */
/* removes the near-singular peak of width d.  p is the point of the    */
/* triangle closest to the projection of x.  tq/tw, rq/rw: Gauss rules  */
/* on [0, 1] for s and for u.  Weights are physical (no J factor).      */
/* pairs and out as or_interior_pairs.                                  */
/* ------------------------------------------------------------------ */
static void cross3(const double *a, const double *b, double *c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

static double dot3(const double *a, const double *b)
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

void or_interior_pairs_sinh(int kid, int64_t n_t, const double *tq, const double *tw,
                            int64_t n_r, const double *rq, const double *rw,
                            const double *obs_pts, const double *obs_ns, const double *src_pts,
                            const int32_t *src_tris, int64_t n_pairs, const int64_t *pairs,
                            const double *params, double *out)
{
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t p_i = 0; p_i < n_pairs; p_i++) {
        const int64_t *pr = pairs + 3 * p_i;
        int clicks = (int)pr[2];
        tri_t st;
        tri_info(&st, src_pts, src_tris, pr[1], clicks, 0);
        const double *x = obs_pts + 3 * pr[0];
        const double *m = obs_ns + 3 * pr[0];
        const double *nn = st.n;
        double xv0[3] = {x[0] - st.v[0][0], x[1] - st.v[0][1], x[2] - st.v[0][2]};
        double ds = dot3(xv0, nn);
        double pp[3] = {x[0] - ds * nn[0], x[1] - ds * nn[1], x[2] - ds * nn[2]};
        /* barycentric map: y - v0 = xi e1 + eta e2 */
        double e1[3], e2[3];
        for (int k = 0; k < 3; k++) {
            e1[k] = st.v[1][k] - st.v[0][k];
            e2[k] = st.v[2][k] - st.v[0][k];
        }
        double g11 = dot3(e1, e1), g12 = dot3(e1, e2), g22 = dot3(e2, e2);
        double det = g11 * g22 - g12 * g12;
        /* split point p: the point of the triangle closest to the projection */
        double p[3] = {pp[0], pp[1], pp[2]};
        {
            double pv[3] = {pp[0] - st.v[0][0], pp[1] - st.v[0][1], pp[2] - st.v[0][2]};
            double h1 = dot3(pv, e1), h2 = dot3(pv, e2);
            double xi = (g22 * h1 - g12 * h2) / det, eta = (g11 * h2 - g12 * h1) / det;
            if (xi >= 0 && eta >= 0 && xi + eta <= 1) {
                for (int k = 0; k < 3; k++)
                    p[k] = pp[k];
            } else {
                double best = INFINITY;
                for (int ed = 0; ed < 3; ed++) {
                    const double *va = st.v[ed], *vb = st.v[(ed + 1) % 3];
                    double ev[3] = {vb[0] - va[0], vb[1] - va[1], vb[2] - va[2]};
                    double w[3] = {pp[0] - va[0], pp[1] - va[1], pp[2] - va[2]};
                    double t = dot3(w, ev) / dot3(ev, ev);
                    t = t < 0 ? 0 : (t > 1 ? 1 : t);
                    double q[3] = {va[0] + t * ev[0], va[1] + t * ev[1], va[2] + t * ev[2]};
                    double dq[3] = {pp[0] - q[0], pp[1] - q[1], pp[2] - q[2]};
                    double dd = dot3(dq, dq);
                    if (dd < best) {
                        best = dd;
                        for (int k = 0; k < 3; k++)
                            p[k] = q[k];
                    }
                }
            }
        }
        double xp[3] = {x[0] - p[0], x[1] - p[1], x[2] - p[2]};
        double d = sqrt(dot3(xp, xp));
        double dmin = 1e-12 * sqrt(st.jac);
        if (d < dmin)
            d = dmin;
        acc_t res[27];
        for (int k = 0; k < 27; k++)
            res[k] = 0;
        for (int ed = 0; ed < 3; ed++) {
            const double *va = st.v[ed], *vb = st.v[(ed + 1) % 3];
            double a[3], b[3], c[3];
            for (int k = 0; k < 3; k++) {
                a[k] = va[k] - p[k];
                b[k] = vb[k] - p[k];
            }
            cross3(a, b, c);
            double cr = dot3(c, nn);
            if (fabs(cr) <= 1e-14 * st.jac)
                continue;
            double sgn = cr > 0 ? 1.0 : -1.0;
            /* foot of the perpendicular from p on the edge line */
            double ev[3] = {vb[0] - va[0], vb[1] - va[1], vb[2] - va[2]};
            double le = sqrt(dot3(ev, ev));
            double t = -dot3(a, ev) / (le * le);
            double f[3] = {a[0] + t * ev[0], a[1] + t * ev[1], a[2] + t * ev[2]};
            double hp = sqrt(dot3(f, f));
            double ef[3] = {f[0] / hp, f[1] / hp, f[2] / hp};
            /* right triangles (p, foot, v) for v = va (t_v = 0), vb (t_v = 1) */
            for (int side = 0; side < 2; side++) {
                double tv = side ? 1.0 : 0.0;
                double Lv = (tv - t) * le;
                if (Lv == 0.0)
                    continue;
                double sig = (side ? (1.0 - t) : t) > 0 ? 1.0 : -1.0;
                double L = fabs(Lv);
                double et[3];
                for (int k = 0; k < 3; k++)
                    et[k] = (Lv > 0 ? 1.0 : -1.0) * ev[k] / le;
                /* phi from ef: tan(phi) = sinh(s), rho_max = hp cosh(s) */
                double S = asinh(L / hp);
                for (int64_t it = 0; it < n_t; it++) {
                    double sv = S * tq[it];
                    double chs = cosh(sv), shs = sinh(sv);
                    double cphi = 1.0 / chs, sphi = shs / chs;
                    double rmax = hp * chs;
                    double dir[3];
                    for (int k = 0; k < 3; k++)
                        dir[k] = cphi * ef[k] + sphi * et[k];
                    double U = asinh(rmax / d);
                    double wphi = S * tw[it] / chs;
                    for (int64_t ir = 0; ir < n_r; ir++) {
                        double u = U * rq[ir];
                        double rho = d * sinh(u);
                        double w = sgn * sig * wphi * U * rw[ir] * d * cosh(u) * rho;
                        double y[3], yv[3];
                        for (int k = 0; k < 3; k++) {
                            y[k] = p[k] + rho * dir[k];
                            yv[k] = y[k] - st.v[0][k];
                        }
                        double h1 = dot3(yv, e1), h2 = dot3(yv, e2);
                        double xi = (g22 * h1 - g12 * h2) / det;
                        double eta = (g11 * h2 - g12 * h1) / det;
                        double q[2] = {xi, eta};
                        pt_integrand(kid, x, m, &st, q, w, 0, params, res);
                    }
                }
            }
        }
        pt_derotate(res, 1.0, clicks, 0, out + 27 * p_i);
    }
}
