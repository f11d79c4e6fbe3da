"""Limit to the boundary: a field evaluated at points pushed off the surface
by a sequence of distances eps_k, extrapolated to eps -> 0.

The model is the polynomial (+ log) fit of tectosaur tests/richardson_error.py
:4-21: with n samples,

    f(eps) = a log(eps) + sum_{i < n-1} c_i eps^i      (include_log)
    f(eps) = sum_{i < n} c_i eps^i                      (otherwise)

solved exactly through the n samples; the limit is c_0 (and, with the log
term, its coefficient a).  The library has no such operator (the fit exists
there only as that test script); here it is an operator over InteriorOp:

    BoundaryLimitOp(obs_pts, obs_ns, src_mesh, K_name, ...).dot(v)

evaluates the interior operator at x_i + side * eps_k h_i n_i for all k in one
InteriorOp (n_eps x n_obs points, one device pass), then combines the samples
with the per-point fit weights on the device (tct_limit_combine).  h_i is a
length scale per point (default: the bounding radius of the closest source
triangle).
"""
import numpy as np
import torch

from . import _native as N
from .interior import InteriorOp


def make_terms(n_terms, include_log):
    """The basis of the fit (tests/richardson_error.py:4-12)."""
    poly_terms = n_terms
    terms = []
    if include_log:
        poly_terms -= 1
        terms = [lambda e: np.log(e)]
    for i in range(poly_terms):
        terms.append(lambda e, i=i: e ** i)
    return terms


def limit(eps_vals, f_vals, include_log):
    """tests/richardson_error.py:14-21: c_0, or (c_0, a) with the log term.
    f_vals may carry trailing dimensions (one fit per trailing index)."""
    terms = make_terms(len(eps_vals), include_log)
    mat = np.array([[t(e) for t in terms] for e in eps_vals])
    coeffs = np.linalg.solve(mat, np.asarray(f_vals))
    if include_log:
        return coeffs[1], coeffs[0]
    return coeffs[0]


def limit_weights(eps, include_log):
    """Weights w (and w_log) with limit = w . f and log coefficient = w_log
    . f, for eps [..., n] (a batch of sample distance sets)."""
    eps = np.asarray(eps, dtype=np.float64)
    n = eps.shape[-1]
    cols = []
    if include_log:
        cols.append(np.log(eps))
    for i in range(n - (1 if include_log else 0)):
        cols.append(eps ** i)
    mat = np.stack(cols, axis=-1)                         # [..., n (samples), n (terms)]
    # row c of inv(mat) maps samples to coefficient c
    inv = np.linalg.inv(mat)
    c0 = 1 if include_log else 0
    return inv[..., c0, :], (inv[..., 0, :] if include_log else None)


def nearest_tri_scale(obs_pts, src_pts, src_tris):
    """Bounding radius of the source triangle whose centroid is closest to
    each point (the default length scale of the offsets)."""
    import scipy.spatial
    v = np.asarray(src_pts)[np.asarray(src_tris)]
    c = v.mean(axis=1)
    r = np.sqrt(np.max(np.sum((v - c[:, None, :]) ** 2, axis=2), axis=1))
    _, idx = scipy.spatial.cKDTree(c).query(np.asarray(obs_pts))
    return r[idx]


class BoundaryLimitOp:
    """lim_{eps -> 0+} InteriorOp(x + side eps n) applied to v, per point,
    from n_eps samples eps_k = eps_start h step^-k.  Same operator
    arguments as InteriorOp (threshold, nq_vertex, nq_far, nq_near, params,
    float_type); close pairs use the polar sinh rule by default, since the
    samples approach the surface.  ``dot`` returns the limits [n_obs * 3];
    ``dot_log`` also the log coefficients (include_log only)."""

    def __init__(self, obs_pts, obs_ns, src_mesh, K_name, threshold, nq_vertex, nq_far, nq_near,
                 params, float_type, eps_start=0.1, eps_step=2.0, n_eps=6, include_log=False,
                 side=1.0, h=None, near_method='sinh', sinh_order=24):
        obs_pts = np.asarray(obs_pts, dtype=np.float64).reshape((-1, 3))
        obs_ns = np.asarray(obs_ns, dtype=np.float64).reshape((-1, 3))
        nrm = np.linalg.norm(obs_ns, axis=1)
        if np.any(nrm == 0):
            raise ValueError('zero observation normal')
        unit = obs_ns / nrm[:, None]
        n = obs_pts.shape[0]
        if h is None:
            h = nearest_tri_scale(obs_pts, src_mesh[0], src_mesh[1])
        h = np.broadcast_to(np.asarray(h, dtype=np.float64), (n,))
        if n_eps < 2 or (include_log and n_eps < 3):
            raise ValueError('n_eps too small for the fit')
        self.n_obs = n
        self.n_eps = n_eps
        self.include_log = include_log
        self.eps = eps_start * h[:, None] * eps_step ** (-np.arange(n_eps))[None, :]   # [n, k]
        pts = (obs_pts[None, :, :] + side * self.eps.T[:, :, None] * unit[None, :, :])
        self.op = InteriorOp(pts.reshape((-1, 3)), np.tile(obs_ns, (n_eps, 1)), src_mesh, K_name,
                             threshold, nq_vertex, nq_far, nq_near, params, float_type,
                             near_method=near_method, sinh_order=sinh_order)
        self.shape = (3 * n, self.op.shape[1])
        self.device = dev = self.op.device
        w, wl = limit_weights(self.eps, include_log)
        self.w = torch.as_tensor(np.ascontiguousarray(w), dtype=torch.float64, device=dev)
        self.w_log = None if wl is None else torch.as_tensor(np.ascontiguousarray(wl),
                                                             dtype=torch.float64, device=dev)
        self._samples = torch.empty(n_eps * n * 3, dtype=torch.float64, device=dev)

    def samples_dev(self, v_dev):
        """Field at every offset: [n_eps, n_obs, 3] (device)."""
        self.op.dot_dev(v_dev, self._samples)
        return self._samples.view(self.n_eps, self.n_obs, 3)

    def _combine(self, w, out):
        N.call('tct_limit_combine', self.n_eps, self.n_obs, 3, N.ptr(w), N.ptr(self._samples),
               N.ptr(out), N.stream_ptr())
        return out

    def dot_dev(self, v_dev, out=None):
        self.samples_dev(v_dev)
        if out is None:
            out = torch.empty(3 * self.n_obs, dtype=torch.float64, device=self.device)
        return self._combine(self.w, out)

    def dot(self, v):
        x = torch.as_tensor(np.ascontiguousarray(v, dtype=np.float64), device=self.device)
        return self.dot_dev(x).cpu().numpy()

    def dot_log(self, v):
        if not self.include_log:
            raise ValueError('built without the log term')
        x = torch.as_tensor(np.ascontiguousarray(v, dtype=np.float64), device=self.device)
        lim = self.dot_dev(x)
        lg = self._combine(self.w_log, torch.empty_like(lim))
        return lim.cpu().numpy(), lg.cpu().numpy()
