"""ORACLE (test infrastructure only): the other solver modes of System_of_eqs --
do_newton_with_linesearch (src_par/do_newton.cpp:454-517, with compute_f :594-599,
check_positive / check_negative :576-592, update_fields :667-673,
compute_old_and_var :675-686) and the GAUSS_NEWTON branch of do_newton (:264-311) --
over any problem that exposes

    residual() -> F,  jacobian() -> J,  state() -> s,  restore(s),
    step(s, X, lam) -> sets u = u(s) - lam * X (X in unknown order)

`XctsProblem` adapts the XCTS oracle system; `VectorProblem` is a plain F(u) on R^n
(for closed-form checks).

Used only by tests/.
"""
import numpy as np
import scipy.linalg

from .kadath import vsub, vmul_scal


class RoundoffError(RuntimeError):
    """The reference's abort "roundoff error in do_newton_with_linesearch"."""


def newton_direction(prob, F, solver="newton"):
    """p with J p = F (do_newton's pdgesv_), or J^T J p = J^T F (GAUSS_NEWTON)."""
    J = prob.jacobian()
    if solver == "gauss-newton":
        return scipy.linalg.lu_solve(scipy.linalg.lu_factor(J.T @ J), J.T @ F)
    return scipy.linalg.lu_solve(scipy.linalg.lu_factor(J), F)


def linesearch_step(prob, precision, ntrymax=10, stepmax=1.0, solver="newton"):
    """One do_newton_with_linesearch iteration.  Returns (converged at entry, error at
    entry, lambda applied, [(lambda, f) per try], the scaled step p)."""
    F = prob.residual()
    error = float(np.abs(F).max())
    if error < precision:
        return True, error, 0.0, [], None
    fold = float(np.sum(F * F))           # compute_f: the plain sum of squares
    slope = -2.0 * fold
    if slope >= 0.0:
        raise RoundoffError
    p = newton_direction(prob, F, solver)
    pmax = p.max()                        # Array max: the signed maximum
    if pmax > stepmax:
        p = p * (stepmax / pmax)
    old = prob.state()
    alpha, lambdamin = 1.0e-4, 1.0e-10
    lam, lam2, f2 = 1.0, 0.0, 0.0
    tries = []
    for itry in range(1, ntrymax + 1):
        prob.step(old, p, lam)
        Fn = prob.residual()
        f = float(np.sum(Fn * Fn))
        tries.append((lam, f))
        if f <= 5.0e-13:
            break
        if f <= fold + alpha * lam * slope or lam <= lambdamin:
            break
        if lam == 1.0:
            lt = -0.5 * slope / (f - fold - slope)
        else:
            rhs1 = f - fold - lam * slope
            rhs2 = f2 - fold - lam2 * slope
            a = (rhs1 / lam ** 2 - rhs2 / lam2 ** 2) / (lam - lam2)
            b = (-lam2 * rhs1 / lam ** 2 + lam * rhs2 / lam2 ** 2) / (lam - lam2)
            if a == 0.0:
                lt = -0.5 * slope / b
            else:
                delta = b * b - 3.0 * a * slope
                if delta < 0.0:
                    raise RoundoffError
                lt = (-b + np.sqrt(delta)) / (3.0 * a)
            if lt > 0.5 * lam:
                lt = 0.5 * lam
        lam2, f2 = lam, f
        lam = max(lt, 0.1 * lam)
        if itry == ntrymax:
            prob.step(old, p, lam)
    return False, error, lam, tries, p


def gauss_newton_step(prob, precision):
    """One do_newton iteration with solver = GAUSS_NEWTON: u <- u - (J^T J)^-1 J^T F."""
    F = prob.residual()
    error = float(np.abs(F).max())
    if error < precision:
        return True, error, None
    p = newton_direction(prob, F, "gauss-newton")
    prob.step(prob.state(), p, 1.0)
    return False, error, p


class XctsProblem:
    """The XCTS oracle system (oracle.xcts.XCTSSystem) as a line-search problem: the
    state is its fields (and Ome in spin mode); a step subtracts lam * the fields of X
    (xx_to_vars of the scaled step, then var = old - lam p_var)."""

    def __init__(self, sys):
        self.sys = sys

    def residual(self):
        return self.sys.residual()

    def jacobian(self):
        return self.sys.jacobian()

    def state(self):
        return ({n: dict(v) for n, v in self.sys.fields.items()}, self.sys.ome)

    def restore(self, s):
        self.sys.set_fields({n: dict(v) for n, v in s[0].items()})
        self.sys.ome = s[1]

    def step(self, s, X, lam):
        corr, dome = self.sys.xx_to_fields(X)
        new = {}
        for name, byd in s[0].items():
            new[name] = {d: vsub(byd[d], vmul_scal(corr[name][d], lam)) for d in byd}
        self.sys.set_fields(new)
        if self.sys.spin is not None:
            self.sys.ome = s[1] - lam * dome


class VectorProblem:
    """F(u) on R^n with its Jacobian, for closed-form checks."""

    def __init__(self, F, J, u0):
        self.F, self.J, self.u = F, J, np.asarray(u0, dtype=np.float64).copy()

    def residual(self):
        return np.atleast_1d(self.F(self.u)).astype(np.float64)

    def jacobian(self):
        return np.atleast_2d(self.J(self.u)).astype(np.float64)

    def state(self):
        return self.u.copy()

    def restore(self, s):
        self.u = s.copy()

    def step(self, s, X, lam):
        self.u = s - lam * X
