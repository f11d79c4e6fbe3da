"""Write tests/golden/corner_radial_golden.npz: the radial integrals of the
interior-corner rule (observation point at a vertex of the source triangle),
derived here with sympy and checked by mpmath.

Set-up (the regularisation tectosaur documents in
tectosaur/derive_hypersingular_vertex.ipynb, cells 1-7): the source triangle
is parametrised in polar coordinates about the vertex, y = v0 + rho a(t)
with a(t) = v1 cos t + v2 sin t in the triangle plane, |a| = A, and the
observation point is lifted off the vertex along the unit normal n by eps,
so the separation is sep = rho a - eps n and r^2 = eps^2 + A^2 rho^2
(a . n = 0).  The four radial integrals over rho in [0, rho_max] are

    I1    = int rho   / r^3
    I2    = int rho^2 / r^3
    I3_ij = int rho   sep_i sep_j / r^5
    I4_ij = int rho^2 sep_i sep_j / r^5

each taken as its finite part as eps -> 0: the constant term of the
asymptotic expansion, with the eps^-1 and log(eps) terms dropped (the
notebook replaces asinh(A rho_max / eps) by log 2 + log A + log rho_max,
which is the same convention).

Derivation (own): sep_i sep_j = a_i a_j rho^2 - eps rho (a_i n_j + n_i a_j)
+ eps^2 n_i n_j, so every integral is a combination of the scalar integrals
J(p, q) = int_0^R rho^p / r^q, times powers of eps.  sympy integrates
each J in closed form and expands it about eps = 0; the constant terms give

    FP I1 = fp[J(1,3)],  FP I2 = fp[J(2,3)],
    FP I3 = a a^T fp[J(3,5)] - (a n^T + n a^T) fp[eps J(2,5)] + n n^T fp[eps^2 J(1,5)]
    FP I4 = a a^T fp[J(4,5)] - (a n^T + n a^T) fp[eps J(3,5)] + n n^T fp[eps^2 J(2,5)].

Cross-check (independent of the closed forms): each fp[eps^k J(p, q)] is
also computed by mpmath quadrature of the regularised integrand at small
eps, with the divergent terms of the expansion subtracted, and must agree.

The fixture holds seeded random inputs (A-vector a, normal n, rho_max) and
the four integrals; the "true" I4 is kept although tectosaur's kernel code
writes its entries as `1/3*(...)`, which C integer division makes zero.

Run:  python tests/golden/make_corner_golden.py
"""
import os

import mpmath as mp
import numpy as np
import sympy as sp

eps, A, R, rho = sp.symbols('eps A R rho', positive=True)


def J(p, q):
    """Closed form of int_0^R rho^p / (eps^2 + A^2 rho^2)^(q/2) d rho."""
    f = rho ** p / (eps ** 2 + A ** 2 * rho ** 2) ** sp.Rational(q, 2)
    return sp.simplify(sp.integrate(f, (rho, 0, R)))


def finite_part(expr):
    """Constant term of the expansion of expr about eps = 0+, with the
    negative powers of eps and the log(eps) terms dropped."""
    # asinh(A R / eps) = log(A R / eps + sqrt((A R / eps)^2 + 1)): sympy
    # expands the log form about eps = 0
    s = sp.series(expr.rewrite(sp.log), eps, 0, 1).removeO()
    s = sp.expand(s)
    lg = sp.log(eps)
    out = 0
    for term in sp.Add.make_args(s):
        if term.has(lg):
            continue
        if term.has(eps):
            continue
        out += term
    return sp.simplify(out), s


def derive():
    pieces = {}
    for name, (p, q, k) in {'J13': (1, 3, 0), 'J23': (2, 3, 0), 'J35': (3, 5, 0),
                            'eJ25': (2, 5, 1), 'e2J15': (1, 5, 2), 'J45': (4, 5, 0),
                            'eJ35': (3, 5, 1), 'e2J25': (2, 5, 2)}.items():
        full = eps ** k * J(p, q)
        fp, series = finite_part(full)
        pieces[name] = (p, q, k, fp, series)
    return pieces


def mp_check(pieces, A_val, R_val):
    """fp[eps^k J(p,q)] by quadrature at small eps minus the divergent terms
    of the series; returns the largest relative disagreement."""
    mp.mp.dps = 40
    worst = 0.0
    for name, (p, q, k, fp, series) in pieces.items():
        # divergent part: the terms with a negative power of eps or log(eps)
        div = sp.Integer(0)
        for t in sp.Add.make_args(series):
            if t.has(sp.log(eps)):
                div += t
                continue
            pw = sp.Poly(t * eps ** 8, eps).degree() - 8 if t.has(eps) else 0
            if pw < 0:
                div += t
        ests = []
        for e in (mp.mpf('1e-9'), mp.mpf('1e-10')):
            f = lambda r: e ** k * r ** p / (e ** 2 + A_val ** 2 * r ** 2) ** (mp.mpf(q) / 2)
            val = mp.quad(f, [0, e / A_val, 10 * e / A_val, 1000 * e / A_val, R_val])
            dv = sp.N(div.subs({eps: sp.Float(str(e), 40), A: sp.Float(str(A_val), 40),
                                R: sp.Float(str(R_val), 40)}), 40)
            ests.append(val - mp.mpf(str(dv)))
        ref = float(sp.N(fp.subs({A: A_val, R: R_val}), 30))
        scale = max(abs(ref), 1.0 / float(A_val) ** (q - p - 1 + k) / max(float(R_val), 1e-30))
        worst = max(worst, abs(float(ests[-1]) - ref) / scale)
    return worst


def main():
    pieces = derive()
    fns = {name: sp.lambdify((A, R), v[3], 'numpy') for name, v in pieces.items()}
    # the cross-check at a few (A, R)
    for A_val, R_val in ((1.0, 1.0), (0.37, 0.81), (2.5, 0.55)):
        w = mp_check(pieces, mp.mpf(str(A_val)), mp.mpf(str(R_val)))
        assert w < 1e-7, ('mpmath disagrees with the closed forms', A_val, R_val, w)
    rng = np.random.default_rng(20240611)
    n_cases = 24
    a_all = np.zeros((n_cases, 3))
    n_all = np.zeros((n_cases, 3))
    rmax = np.zeros(n_cases)
    I1 = np.zeros(n_cases)
    I2 = np.zeros(n_cases)
    I3 = np.zeros((n_cases, 3, 3))
    I4 = np.zeros((n_cases, 3, 3))
    for c in range(n_cases):
        n = rng.standard_normal(3)
        n /= np.linalg.norm(n)
        a = rng.standard_normal(3)
        a -= n * (a @ n)
        a *= rng.uniform(0.05, 20.0) / np.linalg.norm(a)
        # rho_max = 1 / (cos t + sin t) for t in [0, pi/2] lies in [1/sqrt 2, 1]
        t = rng.uniform(0.0, np.pi / 2)
        R_val = 1.0 / (np.cos(t) + np.sin(t))
        Av = float(np.linalg.norm(a))
        a_all[c], n_all[c], rmax[c] = a, n, R_val
        I1[c] = fns['J13'](Av, R_val)
        I2[c] = fns['J23'](Av, R_val)
        aa = np.outer(a, a)
        an = np.outer(a, n) + np.outer(n, a)
        nn = np.outer(n, n)
        I3[c] = aa * fns['J35'](Av, R_val) - an * fns['eJ25'](Av, R_val) + \
            nn * fns['e2J15'](Av, R_val)
        I4[c] = aa * fns['J45'](Av, R_val) - an * fns['eJ35'](Av, R_val) + \
            nn * fns['e2J25'](Av, R_val)
    forms = {name: str(v[3]) for name, v in pieces.items()}
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'corner_radial_golden.npz')
    np.savez(out, a=a_all, n=n_all, rho_max=rmax, I1=I1, I2=I2, I3=I3, I4_true=I4,
             forms=np.array(sorted(forms.items()), dtype='U200'))
    for k, v in sorted(forms.items()):
        print('%-6s %s' % (k, v))
    print('wrote', out)


if __name__ == '__main__':
    main()
