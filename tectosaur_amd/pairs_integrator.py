"""``PairsIntegrator``: the nearfield pair operator, drop-in for
tectosaur/nearfield/pairs_integrator.py:27-93.

Every method integrates a 4-D quadrature rule over a list of triangle pairs
on the device (pairs.hip, FP64) and returns [n, 3, 3, 3, 3] blocks laid out
[b_obs, d_obs, b_src, d_src].  The nearfield is computed in FP64 whatever
``float_type`` says (the reference computes it in float_type with Kahan
summation); results are returned as float64.

The ``*_dev`` variants take and return device tensors, for assembling the
nearfield matrix without host round trips.
"""
import threading

import numpy as np
import torch

from . import _native as N
from . import triangle_rules
from .quadrature import gauss4d_tri

# pairs per launch: bounds the output buffer of one call (81 doubles/pair)
CALL_SIZE = 1 << 20


_RULES = {}
_RULES_LOCK = threading.Lock()


class DeviceRule(tuple):
    """A 4-D rule on the device, ``(qx [nq, 4], qw [nq])``.
    ``has_equal_params`` says whether some point has equal observation and
    source parameters: on a pair that is one triangle in one frame those
    points are at r = 0, and the kernel's r = 0 test must look for them; a
    rule without such points needs only the r2 == 0 test."""

    def __new__(cls, qx, qw, has_equal_params):
        t = super().__new__(cls, (qx, qw))
        t.has_equal_params = bool(has_equal_params)
        return t


def rule_has_equal_params(qx):
    """Whether some point of a 4-D rule (qx [nq, 4]: obs xhat, obs yhat,
    src xhat, src yhat) has equal observation and source parameters."""
    qx = np.asarray(qx, dtype=np.float64)
    return bool(qx.size) and bool(np.any((qx[:, 0] == qx[:, 2]) & (qx[:, 1] == qx[:, 3])))


def r0_test_mode(q, check0):
    """The check0 argument of tct_pairs_integrate: 0 without the r = 0
    test, 1 for the full test, 2 for r2 == 0 alone (rules known to have no
    point of equal obs and src parameters)."""
    if not check0:
        return 0
    return 1 if getattr(q, 'has_equal_params', True) else 2


class PairsIntegrator:
    def __init__(self, kernel, params, float_type, nq_far, nq_near, pts, tris):
        self.kernel = kernel
        self.kid = N.KERNEL_IDS[kernel]
        self.float_type = float_type
        self.params = N.params_array(params)
        self.device = N.require_device()
        tris = np.asarray(tris)
        if tris.size and (tris.min() < 0 or tris.max() >= np.asarray(pts).shape[0]):
            raise IndexError('triangle vertex index out of range')
        self.gpu_pts = N.upload(np.asarray(pts, dtype=np.float64), None, self.device)
        self.gpu_tris = N.upload(np.asarray(tris, dtype=np.int32), None, self.device)
        self._rules = {}
        self.near_q = self._rule(('near4d', nq_near), lambda: gauss4d_tri(nq_near, nq_near))
        self.far_q = self._rule(('far4d', nq_far), lambda: gauss4d_tri(nq_far, nq_far))

    def quad_to_gpu(self, q):
        qx = np.asarray(q[0], dtype=np.float64)
        return DeviceRule(N.upload(qx, torch.float64, self.device),
                          N.upload(q[1], torch.float64, self.device), rule_has_equal_params(qx))

    def _rule(self, key, make):
        if key not in self._rules:
            # rules are pure functions of their orders and read-only on the
            # device: one copy per device shared by every integrator
            dkey = (key, str(self.device))
            q = _RULES.get(dkey)
            if q is None:
                with _RULES_LOCK:
                    q = _RULES.get(dkey)
                    if q is None:
                        q = _RULES[dkey] = self.quad_to_gpu(make())
            self._rules[key] = q
        return self._rules[key]

    def pairs_quad_dev(self, q, pairs_dev, check0, out=None, lanes_per_pair=0):
        """pairs_dev: int32 device tensor [n, 2] or [n, 5]; returns [n, 81]."""
        n = pairs_dev.shape[0]
        width = pairs_dev.shape[1] if pairs_dev.dim() == 2 else 2
        if out is None:
            out = torch.empty((n, 81), dtype=torch.float64, device=self.device)
        st = N.stream_ptr()
        for s in range(0, n, CALL_SIZE):
            e = min(n, s + CALL_SIZE)
            N.call('tct_pairs_integrate', self.kid, q[0].shape[0], N.ptr(q[0]), N.ptr(q[1]),
                   N.ptr(self.gpu_pts), N.ptr(self.gpu_tris), e - s, N.ptr(pairs_dev[s:e]),
                   width, r0_test_mode(q, check0), self.params.ctypes.data, lanes_per_pair,
                   N.ptr(out[s:e]), st)
        return out

    def pairs_quad(self, q, pairs_list, check0):
        pairs_list = np.asarray(pairs_list)
        n = pairs_list.shape[0]
        if n == 0:
            return np.empty((0, 3, 3, 3, 3), dtype=np.float64)
        check_pairs(pairs_list, self.gpu_tris.shape[0])
        p = N.upload(pairs_list.astype(np.int32), None, self.device)
        return self.pairs_quad_dev(q, p, check0).cpu().numpy().reshape((n, 3, 3, 3, 3))

    def correction(self, pairs_list, check0):
        return self.pairs_quad(self.far_q, np.asarray(pairs_list)[:, :2], check0)

    def nearfield(self, pairs_list):
        return self.pairs_quad(self.near_q, pairs_list, False)

    def vert_adj(self, nq, pairs_list):
        if isinstance(nq, (int, np.integer)):
            nq = (nq, nq, nq)
        q = self._rule(('va',) + tuple(nq), lambda: triangle_rules.vertex_adj_quad(*nq))
        return self.pairs_quad(q, self._width5(pairs_list), True)

    def coincident(self, nq, pairs_list):
        q = self._rule(('co', nq), lambda: triangle_rules.coincident_quad(nq))
        return self.pairs_quad(q, pairs_list, True)

    def edge_adj(self, nq, pairs_list):
        q = self._rule(('ea', nq), lambda: triangle_rules.edge_adj_quad(nq))
        return self.pairs_quad(q, self._width5(pairs_list), True)

    @staticmethod
    def _width5(pairs_list):
        p = np.asarray(pairs_list).astype(np.int64)
        if p.shape[0] == 0:
            return np.zeros((0, 5), dtype=np.int64)
        if p.shape[1] == 4:  # (obs, src, obs_clicks, src_clicks) with no flip column
            p = np.concatenate([p, np.zeros((p.shape[0], 1), dtype=np.int64)], axis=1)
        return p


def check_pairs(pairs, n_tris):
    """Host-side validation before a launch: the kernel trusts these indices."""
    p = np.asarray(pairs)
    if p.ndim != 2 or p.shape[1] not in (2, 5):
        raise ValueError('pairs must be [n, 2] or [n, 5]')
    if p.shape[0] == 0:
        return
    if p[:, :2].min() < 0 or p[:, :2].max() >= n_tris:
        raise IndexError('pair triangle index out of range [0, %d)' % n_tris)
    if p.shape[1] == 5 and (p[:, 2:4].min() < 0 or p[:, 2:4].max() > 2 or
                            p[:, 4].min() < 0 or p[:, 4].max() > 1):
        raise ValueError('rotation clicks must be 0..2 and flip 0/1')
