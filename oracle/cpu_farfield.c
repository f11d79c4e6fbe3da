/* CPU baseline of the farfield N-body -- TEST / BENCH INFRASTRUCTURE.
 *
 * Only bench.py's cpu_baseline leg and tests/ use this file; the product
 * never loads it.  It is the reference's point N-body loop
 * (farfield_pts{K}, tectosaur/farfield_direct.cl:15-113) for the H kernel
 * (vector code of tectosaur/kernels/__init__.py:147-173), restated for a
 * many-core CPU: one observation point per iteration of an OpenMP loop,
 * the source loop vectorised (sources in structure-of-arrays form), pairs
 * with r2 == 0 skipped (farfield_direct.cl:89-91; here by a zero weight,
 * which keeps the loop branch-free).  Built with -O3 -march=native
 * -ffast-math, the CPU counterpart of the reference's fast-math OpenCL /
 * CUDA builds (tectosaur/util/opencl.py:76-83, tectosaur/util/cuda.py:70),
 * in FP32 and FP64.  result[i*3+d] = sum_j H(x_i, y_j; m_i, n_j) in_j.
 */
#define _DEFAULT_SOURCE
#include <math.h>
#include <stdint.h>

#define DEFINE_H(NAME, Real, SQRT)                                                        \
    void NAME(int64_t n_obs, const Real *obs_pts, const Real *obs_ns, int64_t n_src,      \
              const Real *sx, const Real *sy, const Real *sz, const Real *nx,             \
              const Real *ny, const Real *nz, const Real *ix, const Real *iy,             \
              const Real *iz, const double *params, Real *result)                         \
    {                                                                                     \
        const Real G = (Real)params[0], nu = (Real)params[1];                             \
        const Real CsH0 = G / (Real)(4.0 * M_PI * (1.0 - params[1]));                     \
        const Real CsH1 = 1 - 2 * nu;                                                     \
        const Real CsH2 = -1 + 4 * nu;                                                    \
        const Real CsH3 = 3 * nu;                                                         \
        _Pragma("omp parallel for schedule(dynamic, 16)")                                  \
        for (int64_t i = 0; i < n_obs; i++) {                                             \
            const Real ox = obs_pts[3 * i], oy = obs_pts[3 * i + 1], oz = obs_pts[3 * i + 2]; \
            const Real mx = obs_ns[3 * i], my = obs_ns[3 * i + 1], mz = obs_ns[3 * i + 2]; \
            Real sumx = 0, sumy = 0, sumz = 0;                                            \
            _Pragma("omp simd reduction(+ : sumx, sumy, sumz)")                            \
            for (int64_t j = 0; j < n_src; j++) {                                         \
                const Real Dx = sx[j] - ox, Dy = sy[j] - oy, Dz = sz[j] - oz;             \
                const Real r2 = Dx * Dx + Dy * Dy + Dz * Dz;                              \
                const Real invr = r2 > 0 ? (Real)1 / SQRT(r2) : (Real)0;                  \
                const Real invr2 = invr * invr;                                           \
                const Real invr3 = invr2 * invr;                                          \
                const Real rn = invr * (nx[j] * Dx + ny[j] * Dy + nz[j] * Dz);            \
                const Real rm = invr * (mx * Dx + my * Dy + mz * Dz);                     \
                const Real mn = mx * nx[j] + my * ny[j] + mz * nz[j];                     \
                const Real sn = ix[j] * nx[j] + iy[j] * ny[j] + iz[j] * nz[j];            \
                const Real sd = invr * (ix[j] * Dx + iy[j] * Dy + iz[j] * Dz);            \
                const Real sm = ix[j] * mx + iy[j] * my + iz[j] * mz;                     \
                const Real Q = CsH0 * invr3;                                              \
                const Real A = Q * 3 * rn;                                                \
                const Real B = Q * CsH1;                                                  \
                const Real C = Q * CsH3;                                                  \
                const Real MT = Q * CsH2 * sn + A * CsH1 * sd;                            \
                const Real NT = B * sm + C * sd * rm;                                     \
                const Real DT = invr * (B * 3 * sn * rm + C * sd * mn +                   \
                                        A * (nu * sm - 5 * sd * rm));                     \
                const Real ST = A * nu * rm + B * mn;                                     \
                sumx += nx[j] * NT + mx * MT + Dx * DT + ST * ix[j];                      \
                sumy += ny[j] * NT + my * MT + Dy * DT + ST * iy[j];                      \
                sumz += nz[j] * NT + mz * MT + Dz * DT + ST * iz[j];                      \
            }                                                                             \
            result[3 * i] = sumx;                                                         \
            result[3 * i + 1] = sumy;                                                     \
            result[3 * i + 2] = sumz;                                                     \
        }                                                                                 \
    }

DEFINE_H(cpu_farfield_pts_H_f32, float, sqrtf)
DEFINE_H(cpu_farfield_pts_H_f64, double, sqrt)
