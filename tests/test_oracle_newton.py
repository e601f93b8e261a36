"""CPU: the oracle's line search (do_newton_with_linesearch, src_par/do_newton.cpp:454-517)
and Gauss-Newton step (do_newton.cpp:264-311) against what the mathematics fixes.

- A linear F takes the full Newton step: one try, lambda = 1, the exact solution.
- atan(u) from u0 = 3 (|u0| > 1.3917: plain Newton diverges, the iterates alternate in
  sign and grow) converges with the line search.  The first backtrack is the minimiser of
  the quadratic through f(0) = f_old, f'(0) = slope = -2 f_old and f(1), which reduces to
  lambda = f_old / (f(1) + f_old) (clamped to [0.1, 0.5]); a slope taken as -f_old (the
  derivative of a halved f) or a halved f with the full slope gives other values.
- The cubic stage: the second backtrack is the minimiser of the cubic through f(0),
  slope, f(lambda_1) and f(lambda_2), here found by solving the interpolation conditions
  as a linear system (not the closed form the search uses).
- Every accepted lambda satisfies the Armijo condition f <= f_old + 1e-4 lambda slope.
- Gauss-Newton: for a square nonsingular J, (J^T J)^-1 J^T F = J^-1 F.
"""
import numpy as np
import pytest

from oracle import newton as ON


def test_linear_problem_takes_the_full_step():
    rng = np.random.default_rng(3)
    A = rng.standard_normal((6, 6)) + 6 * np.eye(6)
    b = rng.standard_normal(6)
    pr = ON.VectorProblem(lambda u: A @ u - b, lambda u: A, np.zeros(6))
    conv, err, lam, tries, p = ON.linesearch_step(pr, 1e-14, stepmax=1e9)
    assert not conv and lam == 1.0 and len(tries) == 1
    np.testing.assert_allclose(pr.u, np.linalg.solve(A, b), rtol=0, atol=1e-13)


def test_stepmax_scales_by_the_signed_maximum():
    A = np.diag([1.0, 2.0])
    b = np.array([-4.0, 1.0])          # Newton step p = -A^-1 b... u0 = 0: p = (4, -0.5)
    pr = ON.VectorProblem(lambda u: A @ u - b, lambda u: A, np.zeros(2))
    conv, err, lam, tries, p = ON.linesearch_step(pr, 1e-14, ntrymax=1, stepmax=2.0)
    # p = J^-1 F(0) = (4, -0.5): max 4 > 2, scaled by 1/2; the absolute maximum would be
    # the same here, so also check a step whose largest magnitude is negative
    np.testing.assert_allclose(p, [2.0, -0.25])
    b2 = np.array([4.0, -1.0])         # p = (-4, 0.5): signed max 0.5 < 2, no scaling
    pr2 = ON.VectorProblem(lambda u: A @ u - b2, lambda u: A, np.zeros(2))
    _, _, _, _, p2 = ON.linesearch_step(pr2, 1e-14, ntrymax=1, stepmax=2.0)
    np.testing.assert_allclose(p2, [-4.0, 0.5])


def _atan():
    return ON.VectorProblem(lambda u: np.arctan(u), lambda u: np.array([[1.0 / (1.0 + u[0] ** 2)]]), [3.0])


def test_plain_newton_diverges_on_atan():
    u = 3.0
    for _ in range(6):
        u = u - np.arctan(u) * (1.0 + u * u)
    assert abs(u) > 1e6


def test_linesearch_converges_on_atan_with_the_quadratic_backtrack():
    pr = _atan()
    u0 = 3.0
    fold = np.arctan(u0) ** 2
    p = np.arctan(u0) * (1.0 + u0 * u0)
    f1 = np.arctan(u0 - p) ** 2
    conv, err, lam, tries, _ = ON.linesearch_step(pr, 1e-12, stepmax=1e9)
    assert tries[0][0] == 1.0 and tries[0][1] == pytest.approx(f1, rel=1e-14)
    expect = min(max(fold / (f1 + fold), 0.1), 0.5)
    assert tries[1][0] == pytest.approx(expect, rel=1e-14)
    for it in range(40):
        conv, err, lam, tries, _ = ON.linesearch_step(pr, 1e-12, stepmax=1e9)
        if conv:
            break
    assert conv and abs(pr.u[0]) < 1e-12


def test_armijo_holds_at_every_accepted_lambda():
    pr = _atan()
    for it in range(40):
        F0 = pr.residual()
        fold = float(np.sum(F0 * F0))
        conv, err, lam, tries, _ = ON.linesearch_step(pr, 1e-12, stepmax=1e9)
        if conv:
            break
        lam_last, f_last = tries[-1]
        assert f_last <= 5e-13 or f_last <= fold - 2e-4 * lam_last * fold, (it, tries)


def test_cubic_backtrack_is_the_cubic_interpolant_minimiser():
    # atan from u0 = 10: the full step lands at -139 and the quadratic backtrack
    # (lambda ~ 0.47) at -60, both with f > f_old, so the cubic stage runs
    pr = _atan()
    pr.u = np.array([10.0])
    F0 = pr.residual()
    fold = float(F0 @ F0)
    slope = -2.0 * fold
    conv, err, lam, tries, p = ON.linesearch_step(pr, 1e-12, stepmax=1e9)
    assert len(tries) >= 3, tries
    (l1, f1), (l2, f2) = tries[0], tries[1]
    # cubic c0 + c1 l + c2 l^2 + c3 l^3 with c0 = fold, c1 = slope, through (l1, f1), (l2, f2)
    M = np.array([[l1 ** 2, l1 ** 3], [l2 ** 2, l2 ** 3]])
    c2, c3 = np.linalg.solve(M, [f1 - fold - slope * l1, f2 - fold - slope * l2])
    roots = np.roots([3 * c3, 2 * c2, slope])
    mins = [r.real for r in roots if abs(r.imag) < 1e-12 and 6 * c3 * r.real + 2 * c2 > 0]
    expect = min(max(mins[0], 0.1 * l2), 0.5 * l2)
    assert tries[2][0] == pytest.approx(expect, rel=1e-10)


def test_gauss_newton_direction_equals_newton_for_square_systems():
    rng = np.random.default_rng(5)
    A = rng.standard_normal((8, 8)) + 4 * np.eye(8)
    F = rng.standard_normal(8)
    pr = ON.VectorProblem(lambda u: A @ u - F, lambda u: A, np.zeros(8))
    p_gn = ON.newton_direction(pr, F, "gauss-newton")
    np.testing.assert_allclose(p_gn, np.linalg.solve(A, F), rtol=1e-11, atol=1e-12)
    # and Gauss-Newton iterations reach the root of a nonlinear square system
    G = lambda u: np.array([u[0] ** 2 + u[1] - 3.0, u[0] - u[1] ** 3 + 1.0])
    dG = lambda u: np.array([[2 * u[0], 1.0], [1.0, -3 * u[1] ** 2]])
    pr = ON.VectorProblem(G, dG, [1.5, 0.5])
    for _ in range(30):
        conv, err, _ = ON.gauss_newton_step(pr, 1e-13)
        if conv:
            break
    assert conv and np.abs(G(pr.u)).max() < 1e-13
