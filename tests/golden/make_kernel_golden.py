"""Write tests/golden/kernels_golden.npz: Kelvin point kernels U, T, A, H
evaluated symbolically (sympy) at seeded random points.

The derivation is independent of every numeric code path in this project:
  * U is the Kelvin fundamental solution of 3-D isotropic elasticity,
      U_kj = 1/(16 pi mu (1-nu) r) [(3-4nu) delta_kj + d_k d_j / r^2],
    with d = y - x (source minus observation point);
  * the traction of a displacement field u at a point with normal n is
      t_i = sigma_il n_l,  sigma = lambda tr(eps) I + 2 mu eps,
      eps = (grad u + grad u^T) / 2,  lambda = 2 mu nu / (1 - 2 nu);
  * T_kj: traction operator w.r.t. the source point (normal n, index j)
    applied to the field u_dir = U_k,dir;
  * A_kj: traction operator w.r.t. the observation point (normal m, index k)
    applied to u_dir = U_dir,j;
  * H_kj: both operators, source first then observation.
This is the method tectosaur documents in tests/elastic_derive.py:47-147 and
compares its hand-coded kernels against (tests/test_elastic.py:38-67): the
output row index k is the observation component and the column j the source
component, K[k][j] == (kernel applied to e_j)[k].

Run:  python tests/golden/make_kernel_golden.py
"""
import os

import numpy as np
import sympy as sp

x1, y1, z1, x2, y2, z2 = sp.symbols('x1 y1 z1 x2 y2 z2', real=True)
mu, nu = sp.symbols('mu nu', positive=True)
nx, ny, nz, mx, my, mz = sp.symbols('nx ny nz mx my mz', real=True)
SRC = [x1, y1, z1]
OBS = [x2, y2, z2]
NSRC = [nx, ny, nz]
NOBS = [mx, my, mz]


def kelvin(k, j):
    d = [SRC[i] - OBS[i] for i in range(3)]
    r2 = sum(di ** 2 for di in d)
    c = 1 / (16 * sp.pi * mu * (1 - nu))
    return c / sp.sqrt(r2) * ((3 - 4 * nu) * (1 if k == j else 0) + d[k] * d[j] / r2)


def traction(u, pos, index, normal):
    grad = [[sp.diff(u[a], pos[b]) for b in range(3)] for a in range(3)]
    eps = [[(grad[a][b] + grad[b][a]) / 2 for b in range(3)] for a in range(3)]
    tr = sum(eps[a][a] for a in range(3))
    lam = 2 * mu * nu / (1 - 2 * nu)
    sig = [[lam * tr * (1 if a == b else 0) + 2 * mu * eps[a][b] for b in range(3)]
           for a in range(3)]
    return sum(sig[index][b] * normal[b] for b in range(3))


def derive():
    U = [[kelvin(k, j) for j in range(3)] for k in range(3)]
    T = [[traction([kelvin(k, dd) for dd in range(3)], SRC, j, NSRC) for j in range(3)]
         for k in range(3)]
    A = [[traction([kelvin(dd, j) for dd in range(3)], OBS, k, NOBS) for j in range(3)]
         for k in range(3)]
    H = []
    for k in range(3):
        row = []
        for j in range(3):
            inner = [traction([kelvin(od, sd) for sd in range(3)], SRC, j, NSRC)
                     for od in range(3)]
            row.append(traction(inner, OBS, k, NOBS))
        H.append(row)
    return {'elasticU3': U, 'elasticT3': T, 'elasticA3': A, 'elasticH3': H}


def main():
    rng = np.random.default_rng(20240613)
    n = 24
    mus = rng.uniform(0.5, 3.0e10, n)
    mus[: n // 2] = rng.uniform(0.5, 2.0, n // 2)
    nus = rng.uniform(0.05, 0.45, n)
    src = rng.uniform(-2, 2, (n, 3))
    obs = rng.uniform(-2, 2, (n, 3))
    nsrc = rng.normal(size=(n, 3))
    nsrc /= np.linalg.norm(nsrc, axis=1)[:, None]
    nobs = rng.normal(size=(n, 3))
    nobs /= np.linalg.norm(nobs, axis=1)[:, None]
    kernels = derive()
    args = (mu, nu, x1, y1, z1, x2, y2, z2, nx, ny, nz, mx, my, mz)
    out = dict(mu=mus, nu=nus, src=src, obs=obs, nsrc=nsrc, nobs=nobs)
    for name, K in kernels.items():
        vals = np.zeros((n, 3, 3))
        for k in range(3):
            for j in range(3):
                f = sp.lambdify(args, K[k][j], 'mpmath')
                for t in range(n):
                    vals[t, k, j] = float(f(mus[t], nus[t], *src[t], *obs[t], *nsrc[t],
                                            *nobs[t]))
        out[name] = vals
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'kernels_golden.npz')
    np.savez(path, **out)
    print('wrote', path)


if __name__ == '__main__':
    main()
