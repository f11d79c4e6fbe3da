"""The pair integrator skips a rule point of a pair whose local triangles
are one triangle in one frame when its obs and src parameters are equal,
instead of testing r2 == 0 on basis-weighted positions as
tectosaur/assemble.cl:24-103 does.  The two agree when those are the only
points at r = 0 and every other point stays far from r = 0: checked here
for the coincident rules and the product rules (the coincident class and
its farfield-rule correction), and, for the adjacent rules, that no point
comes near r = 0 at all."""
import numpy as np
import pytest

from tectosaur_amd import triangle_rules as T
from tectosaur_amd.quadrature import gauss4d_tri


def _phys(tri, xh, yh):
    return (1 - xh - yh)[:, None] * tri[0] + xh[:, None] * tri[1] + yh[:, None] * tri[2]


TRI = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.2, 0.9, 0.0]])


@pytest.mark.parametrize('make', [lambda n: T.coincident_quad(n), lambda n: gauss4d_tri(n, n)])
@pytest.mark.parametrize('n', range(1, 9))
def test_same_triangle_zero_points_are_equal_parameters(make, n):
    q, _ = make(n)
    same = (q[:, 0] == q[:, 2]) & (q[:, 1] == q[:, 3])
    d = np.linalg.norm(_phys(TRI, q[:, 0], q[:, 1]) - _phys(TRI, q[:, 2], q[:, 3]), axis=1)
    assert np.all(d[same] == 0.0)
    assert d[~same].min() > 1e-5


@pytest.mark.parametrize('n', range(1, 9))
def test_adjacent_rules_stay_away_from_r0(n):
    # edge-adjacent: the source traverses the shared edge opposite to the
    # observer (local edge 0-1 of both); vertex-adjacent: local vertex 0 shared
    ea = np.array([TRI[1], TRI[0], [0.5, -0.8, 0.1]])
    va = np.array([TRI[0], [-0.7, 0.3, 0.2], [-0.5, -0.8, 0.1]])
    for (q, _), S in ((T.edge_adj_quad(n), ea), (T.vertex_adj_quad(n, n, n), va)):
        d = np.linalg.norm(_phys(TRI, q[:, 0], q[:, 1]) - _phys(S, q[:, 2], q[:, 3]), axis=1)
        assert d.min() > 1e-5


def test_r0_test_mode_follows_the_rule():
    # the product rule has points with equal obs and src parameters (on a
    # coincident pair those are at r = 0: the full test); the singular rules
    # have none, so r2 == 0 alone is tested
    from tectosaur_amd.pairs_integrator import r0_test_mode, rule_has_equal_params
    from tectosaur_amd.quadrature import gauss4d_tri

    class Rule(tuple):
        pass

    for q, eq in ((gauss4d_tri(2, 2), True), (T.coincident_quad(3), False),
                  (T.coincident_quad(8), False),
                  (T.edge_adj_quad(8), False),
                  (T.vertex_adj_quad(8, 8, 8), False)):
        assert rule_has_equal_params(q[0]) == eq
        r = Rule((None, None))
        r.has_equal_params = eq
        assert r0_test_mode(r, True) == (1 if eq else 2)
        assert r0_test_mode(r, False) == 0
    # a rule without the attribute gets the full test
    assert r0_test_mode((None, None), True) == 1
