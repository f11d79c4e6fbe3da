"""GMRES (algorithm on CPU torch tensors; the device path is exercised by
tests/test_solve_gpu.py)."""
import numpy as np
import pytest
import torch

from tectosaur_amd.solver import gmres


def dense_apply(A):
    At = torch.as_tensor(A)

    def apply(x, out, prec):
        torch.mv(At, x, out=out)
        return out
    return apply


@pytest.mark.parametrize('restart', [5, 20, 200])
def test_gmres_matches_direct_solve(restart):
    rng = np.random.default_rng(1)
    n = 60
    A = np.eye(n) * 4 + rng.normal(size=(n, n)) / np.sqrt(n)
    b = rng.normal(size=n)
    res = gmres(dense_apply(A), torch.as_tensor(b), tol=1e-12, restart=restart, maxiter=2000)
    assert res.converged
    x = res.x.numpy()
    assert np.linalg.norm(A @ x - b) <= 1e-11 * np.linalg.norm(b)
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-9, atol=1e-10)


def test_mixed_precision_refinement_reaches_fp64_residual():
    # inner operator perturbed by 1e-6 (as an FP32 farfield is); restarts use
    # the exact operator, so the final residual is the exact operator's.
    rng = np.random.default_rng(2)
    n = 80
    A = np.eye(n) * 3 + rng.normal(size=(n, n)) / np.sqrt(n)
    E = 1e-6 * rng.normal(size=(n, n)) / np.sqrt(n)
    At, Ait = torch.as_tensor(A), torch.as_tensor(A + E)

    def apply(x, out, prec):
        torch.mv(At if prec == 'f64' else Ait, x, out=out)
        return out
    b = torch.as_tensor(rng.normal(size=n))
    res = gmres(apply, b, tol=1e-12, restart=50, inner_prec='f32', outer_prec='f64',
                inner_reduction=1e-4)
    assert res.converged and res.restarts >= 2
    assert np.linalg.norm(A @ res.x.numpy() - b.numpy()) <= 1e-12 * np.linalg.norm(b.numpy())


def test_coarse_residuals_never_decide_convergence():
    # three operators: 'tree' (inner, 1e-4 off), 'f32' (coarse residuals,
    # 1e-7 off) and the exact 'f64'.  Early restarts may use the coarse
    # operator; the last residual, which decides convergence, must be exact.
    from tectosaur_amd.solver import GmresState
    rng = np.random.default_rng(3)
    n = 90
    A = np.eye(n) * 3 + rng.normal(size=(n, n)) / np.sqrt(n)
    ops = {'f64': torch.as_tensor(A),
           'f32': torch.as_tensor(A + 1e-7 * rng.normal(size=(n, n)) / np.sqrt(n)),
           'tree': torch.as_tensor(A + 1e-4 * rng.normal(size=(n, n)) / np.sqrt(n))}
    used = []

    def apply(x, out, prec):
        used.append(prec)
        torch.mv(ops[prec], x, out=out)
        return out
    b = torch.as_tensor(rng.normal(size=n))
    st = GmresState(apply, b, tol=1e-10, restart=60, maxiter=2000, inner_prec='tree',
                    outer_prec='f64', inner_reduction=1e-3, coarse_prec='f32', coarse_above=1e-4)
    res = st.run()
    assert res.converged
    outer = [p for p in used if p != 'tree']
    assert outer[-1] == 'f64' and 'f32' in outer
    assert np.linalg.norm(A @ res.x.numpy() - b.numpy()) <= 1e-10 * np.linalg.norm(b.numpy())


def test_residual_updates_save_exact_residuals_and_never_decide():
    # as above, plus residual updates r <- r - A_f32 d once the exact
    # residual is small: fewer exact ('f64') applications, the same
    # converged accuracy, and the deciding residual is still exact
    from tectosaur_amd.solver import GmresState
    rng = np.random.default_rng(4)
    n = 120
    A = np.eye(n) * 3 + rng.normal(size=(n, n)) / np.sqrt(n)
    ops = {'f64': torch.as_tensor(A),
           'f32': torch.as_tensor(A + 1e-7 * rng.normal(size=(n, n)) / np.sqrt(n)),
           'tree': torch.as_tensor(A + 3e-3 * rng.normal(size=(n, n)) / np.sqrt(n))}
    b = torch.as_tensor(rng.normal(size=n))
    counts = {}
    for upd in (None, 'f32'):
        used = []

        def apply(x, out, prec):
            used.append(prec)
            torch.mv(ops[prec], x, out=out)
            return out
        st = GmresState(apply, b, tol=1e-10, restart=60, maxiter=4000, inner_prec='tree',
                        outer_prec='f64', inner_reduction=3e-3, coarse_prec='f32',
                        coarse_above=3e-5, update_prec=upd, update_below=1e-3)
        res = st.run()
        assert res.converged
        assert used[-1] == 'f64'
        assert np.linalg.norm(A @ res.x.numpy() - b.numpy()) <= 1e-10 * np.linalg.norm(b.numpy())
        counts[upd] = used.count('f64')
    assert counts['f32'] < counts[None]
    # a planted error in the update operator (1e-3 off) would stall a
    # converged-looking update: the exact confirmation then keeps iterating
    # rather than accepting it
    used = []
    bad = {'f64': ops['f64'], 'tree': ops['tree'],
           'f32': torch.as_tensor(A + 1e-3 * rng.normal(size=(n, n)) / np.sqrt(n))}

    def apply_bad(x, out, prec):
        used.append(prec)
        torch.mv(bad[prec], x, out=out)
        return out
    res = GmresState(apply_bad, b, tol=1e-10, restart=60, maxiter=4000, inner_prec='tree',
                     outer_prec='f64', inner_reduction=3e-3, coarse_prec=None,
                     update_prec='f32', update_below=1e-3).run()
    assert used[-1] == 'f64'
    if res.converged:
        assert np.linalg.norm(A @ res.x.numpy() - b.numpy()) <= 1e-10 * np.linalg.norm(b.numpy())


def test_gmres_zero_rhs_and_nan():
    A = np.eye(4)
    r = gmres(dense_apply(A), torch.zeros(4, dtype=torch.float64), tol=1e-8)
    assert r.converged and np.all(r.x.numpy() == 0)
    b = torch.tensor([1.0, np.nan, 0, 0], dtype=torch.float64)
    r = gmres(dense_apply(A), b, tol=1e-8)
    assert not r.converged and r.info != 0


def test_gmres_reports_nonconvergence():
    rng = np.random.default_rng(3)
    n = 50
    A = rng.normal(size=(n, n))
    b = torch.as_tensor(rng.normal(size=n))
    r = gmres(dense_apply(A), b, tol=1e-14, restart=3, maxiter=1)
    assert not r.converged and r.info != 0 and r.iterations == 1


def test_step_callback_sees_every_arnoldi_step():
    # the per-step callback gets the running iteration count and the cycle's
    # residual estimate (monotone within a cycle, relative to ||b||)
    from tectosaur_amd.solver import GmresState
    rng = np.random.default_rng(5)
    n = 50
    A = np.eye(n) * 3 + rng.normal(size=(n, n)) / np.sqrt(n)
    b = rng.normal(size=n)
    seen = []
    st = GmresState(dense_apply(A), torch.as_tensor(b), tol=1e-10, restart=8, maxiter=500,
                    step_callback=lambda it, est: seen.append((it, est)))
    res = st.run()
    assert res.converged
    assert [it for it, _ in seen] == list(range(1, res.iterations + 1))
    for (i0, e0), (i1, e1) in zip(seen, seen[1:]):
        if i1 % 8 != 1:     # same cycle: GMRES residual estimates do not grow
            assert e1 <= e0 * (1 + 1e-12)
    assert seen[-1][1] <= 1e-10 * 1.01
