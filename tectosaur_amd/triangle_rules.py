"""Singular and adjacent 4-D quadrature rules for triangle pairs.

Mirrors tectosaur/nearfield/triangle_rules.py:
  * ``coincident_quad`` (:78-146) -- Erichsen & Sauter (1998), six
    sub-simplex mappings of the coincident pair,
  * ``edge_adj_quad`` (:148-213) -- Sauter & Schwab, five mappings for a
    shared edge,
  * ``vertex_adj_quad`` (:31-76) -- polar / Duffy split about a shared vertex.

The mappings are the published ones, so they are necessarily the same
formulas; the evaluation here is vectorised numpy and the point ordering
follows the reference loops (mapping-major, then the loop nests in their
order), so a rule built here lists the same points in the same order.
Points are (obs xhat, obs yhat, src xhat, src yhat) in the frames that the
pair integrator sets up (shared vertex at local vertex 0; shared edge is
local edge 0-1 of both triangles, traversed in opposite directions).
"""
import numpy as np

from .quadrature import gaussxw, map_to, gauss2d_tri


def _line01(nq):
    return map_to(gaussxw(nq), [0.0, 1.0])


def coincident_quad(nq):
    qtri = gauss2d_tri(nq)
    lx, lw = _line01(nq)
    nt, nl = qtri[0].shape[0], lx.shape[0]
    # loop order: mapping, triangle point, x3, eta
    w1 = np.repeat(qtri[0][:, 0], nl * nl)
    w2 = np.repeat(qtri[0][:, 1], nl * nl)
    f12 = np.repeat(qtri[1], nl * nl)
    x3 = np.tile(np.repeat(lx, nl), nt)
    f3 = np.tile(np.repeat(lw, nl), nt)
    eta = np.tile(np.tile(lx, nl), nt)
    feta = np.tile(np.tile(lw, nl), nt)
    w3 = x3 * (1 - w1 - w2)
    F = f12 * (f3 * (1 - w1 - w2)) * feta * w1
    maps = [
        (w1 + w2 + w3, w1 + w2, w1 * (1 - eta) + w2 + w3, w2),
        (w1 * (1 - eta) + w2 + w3, w1 * (1 - eta) + w2, w1 + w2 + w3, w2),
        (w1 + w2 + w3, w1 * eta + w2, w2 + w3, w2),
        (w1 * (1 - eta) + w2 + w3, w2, w1 + w2 + w3, w1 + w2),
        (w1 + w2 + w3, w2, w1 * (1 - eta) + w2 + w3, w1 * (1 - eta) + w2),
        (w2 + w3, w2, w1 + w2 + w3, w1 * eta + w2),
    ]
    pts, wts = [], []
    for x1, x2, y1, y2 in maps:
        pts.append(np.stack([1 - x1, x2, 1 - y1, y2], axis=1))
        wts.append(F)
    return np.concatenate(pts), np.concatenate(wts)


def edge_adj_quad(nq):
    lx, lw = _line01(nq)
    n = lx.shape[0]
    # loop order: mapping, e1, e2, e3, P
    idx = np.indices((n, n, n, n)).reshape(4, -1)
    e1, e2, e3, P = (lx[i] for i in idx)
    f = lw[idx[0]] * lw[idx[1]] * lw[idx[2]] * lw[idx[3]]
    maps = [
        (P, P * e1 * e3, P * (1 - e1 * e2), P * e1 * (1 - e2), P ** 3 * e1 ** 2),
        (P, P * e1, P * (1 - e1 * e2 * e3), P * e1 * e2 * (1 - e3), P ** 3 * e1 ** 2 * e2),
        (P * (1 - e1 * e2), P * e1 * (1 - e2), P, P * e1 * e2 * e3, P ** 3 * e1 ** 2 * e2),
        (P * (1 - e1 * e2 * e3), P * e1 * e2 * (1 - e3), P, P * e1, P ** 3 * e1 ** 2 * e2),
        (P * (1 - e1 * e2 * e3), P * e1 * (1 - e2 * e3), P, P * e1 * e2, P ** 3 * e1 ** 2 * e2),
    ]
    pts, wts = [], []
    for x1, x2, y1, y2, jac in maps:
        sy = y2
        sx = (1 - sy) - (1 - y1)
        pts.append(np.stack([1 - x1, x2, sx, sy], axis=1))
        wts.append(f * jac)
    return np.concatenate(pts), np.concatenate(wts)


def vertex_adj_quad(n_theta, n_beta, n_alpha):
    tq = map_to(gaussxw(n_theta), [0, np.pi / 2])
    bq = gaussxw(n_beta)
    aq = gaussxw(n_alpha)

    def rho_max(t):
        return 1.0 / (np.cos(t) + np.sin(t))

    pts, wts = [], []
    # loop order: theta_p, theta_q, (beta lower half, beta upper half), beta, alpha
    for tp, tpw in zip(*tq):
        for tqv, tqw in zip(*tq):
            split = np.arctan(rho_max(tqv) / rho_max(tp))
            for lo, hi, amax in (
                (0.0, split, lambda b: rho_max(tp) / np.cos(b)),
                (split, np.pi / 2, lambda b: rho_max(tqv) / np.sin(b)),
            ):
                bx, bw = map_to(bq, [lo, hi])
                for b, bwv in zip(bx, bw):
                    ax, aw = map_to(aq, [0, amax(b)])
                    jac = ax ** 3 * np.cos(b) * np.sin(b)
                    rp = ax * np.cos(b)
                    rq = ax * np.sin(b)
                    pts.append(np.stack(
                        [rp * np.cos(tp), rp * np.sin(tp), rq * np.cos(tqv), rq * np.sin(tqv)],
                        axis=1))
                    wts.append(jac * aw * (bwv * tpw * tqw))
    return np.concatenate(pts), np.concatenate(wts)
