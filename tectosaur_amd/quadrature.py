"""Host-side quadrature rules (one-time setup, FP64 numpy).

Mirrors tectosaur/util/quadrature.py: ``gaussxw`` (:6-12), ``map_to``
(:15-19), ``gauss2d_tri`` (:33-45) and ``gauss4d_tri`` (:47-60).  The point
and weight conventions are the reference's: triangle rules live on the unit
reference triangle {xhat, yhat >= 0, xhat + yhat <= 1} with weights summing
to 1/2, and 4-D rules list (obs xhat, obs yhat, src xhat, src yhat).
"""
import numpy as np

from .tri_rules_data import TRI_RULES


def gaussxw(n):
    """n-point Gauss-Legendre rule on [-1, 1] (Golub-Welsch eigen-solve of the
    Jacobi matrix, as tectosaur/util/quadrature.py:6-12), nodes ascending."""
    if n < 1:
        raise ValueError("gaussxw needs n >= 1")
    k = np.arange(1.0, n)
    off = k / np.sqrt(4 * k * k - 1)
    J = np.diag(off, 1) + np.diag(off, -1)
    x, V = np.linalg.eigh(J)
    w = 2.0 * V[0, :] ** 2
    return x, w


def map_to(qr, interval):
    """Affine map of a [-1, 1] rule onto ``interval`` (quadrature.py:15-19)."""
    x01 = (qr[0] + 1) / 2
    outx = interval[0] + (interval[1] - interval[0]) * x01
    outw = (qr[1] / 2) * (interval[1] - interval[0])
    return outx, outw


def quadrature(f, qr):
    return np.sum(f(qr[0]) * qr[1])


def gauss2d_rect(n):
    xg, wg = gaussxw(n)
    X, Y = np.meshgrid(xg, xg)
    pts = np.array([X.flatten(), Y.flatten()]).T
    w = np.outer(wg, wg).flatten()
    return pts, w


def _tri_table(order):
    """Smallest published symmetric rule whose degree covers ``order``."""
    if order <= 0:
        raise ValueError("bad triangle rule order")
    for deg in sorted(TRI_RULES):
        if order <= deg:
            p, w = TRI_RULES[deg]
            return np.array(p, dtype=np.float64), np.array(w, dtype=np.float64)
    return None, None


def gauss2d_tri(n):
    """Triangle rule exact to degree 2n-1 (quadrature.py:33-45): a published
    symmetric rule when one exists (n <= 7), else the n*n collapsed
    (Duffy) Gauss product rule."""
    p, w = _tri_table(2 * n - 1)
    if p is not None:
        return p, w * 0.5
    rp, rw = gauss2d_rect(n)
    x01 = (rp[:, 0] + 1) / 2
    y01 = (rp[:, 1] + 1) / 2
    pts = np.empty_like(rp)
    pts[:, 0] = x01 * (1 - y01)
    pts[:, 1] = y01
    w = (rw / 4.0) * (1 - y01)
    return pts, w


def gauss4d_tri(n_outer, n_inner):
    """Tensor product of two triangle rules: obs (outer) x src (inner), the
    ordering of quadrature.py:47-60."""
    po, wo = gauss2d_tri(n_outer)
    pi, wi = gauss2d_tri(n_inner)
    no, ni = po.shape[0], pi.shape[0]
    pts = np.empty((no * ni, 4))
    pts[:, 0] = np.repeat(po[:, 0], ni)
    pts[:, 1] = np.repeat(po[:, 1], ni)
    pts[:, 2] = np.tile(pi[:, 0], no)
    pts[:, 3] = np.tile(pi[:, 1], no)
    w = np.repeat(wo, ni) * np.tile(wi, no)
    return pts, w


def gauss2d_tri_size(n):
    """Number of points gauss2d_tri(n) returns."""
    p, _ = _tri_table(2 * n - 1)
    return p.shape[0] if p is not None else n * n
