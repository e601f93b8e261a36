"""GPU: the other solver modes of the do_newton boundary on the device, through the
C-ABI (kad_system_do_newton_linesearch; kad_system_do_newton with solver 4).

- do_newton_with_linesearch (src_par/do_newton.cpp:454-517) on the Kerr system (9x7x4,
  3 domains, Ome = 0.05) against the oracle's restatement (oracle/newton.py), started
  from the same fields:
    * the default 1e-3 perturbation: two full steps, then a backtrack to lambda = 0.1
      (plain Newton diverges from there at this resolution);
    * a perturbation of amplitude 1: the step is scaled by stepmax / max(p), then the
      search runs all ten tries through the cubic model and applies the last lambda.
  The residual norm at entry, the number of tries and lambda must agree (lambda to
  1e-5 relative after ten cubic tries, 1e-6 measured), and the fields after each
  search.  The Newton step p itself carries the LU's rounding amplified by
  cond(J) (~1e9 here), so the fields agree to 1e-7 relative, not to 1e-12.
- Gauss-Newton (do_newton.cpp:264-311): J^T J X = J^T F squares the condition number,
  so its steps differ from Newton's at the 1e-4 level (cond(J)^2 u); the check is that
  the device reaches ||F||_inf < 1e-10 and the same solution as the device's Newton
  (1e-10), and that its first step matches the oracle's Gauss-Newton step to 1e-3.
"""
import numpy as np
import pytest

from kadath_frontera_amd import programs

pytestmark = pytest.mark.gpu

RES, BOUNDS = (9, 7, 4), [1.0, 2.0, 4.0]
NAMES = {"P": 0, "N": 1, "bx": 2, "by": 3, "bz": 4}


def _load(s, fields):
    for name, f in NAMES.items():
        for d in range(1, len(BOUNDS) + 1):
            v = fields[name][d]
            v.coef()
            s.set_coef(f, d, v.cf)


def _field_err(s, fields):
    worst = 0.0
    for name, f in NAMES.items():
        for d in range(1, len(BOUNDS) + 1):
            v = fields[name][d]
            v.coef()
            worst = max(worst, np.abs(s.coef(f, d) - v.cf).max() / max(1.0, np.abs(v.cf).max()))
    return worst


@pytest.mark.parametrize("amp,steps", [(1e-3, 3), (1.0, 2)])
def test_linesearch_matches_oracle(ctx, amp, steps):
    from oracle import xcts as OX, newton as ON
    sp = OX.Space(RES, BOUNDS)
    sy = OX.XCTSSystem(sp, 1, 3, n0=0.5, a=1.0, ome=0.05)
    sy.set_fields(OX.initial_fields(sy, 1.0, OX.initial_perturbation(sy, amp=amp)))
    pr = ON.XctsProblem(sy)
    s = programs.kerr(ctx, RES, BOUNDS, n0=0.5, a=1.0, ome=0.05).build(order=0, batch=256)
    _load(s, sy.fields)
    saw_backtrack = False
    for step in range(steps):
        conv_o, err_o, lam_o, tries_o, _ = ON.linesearch_step(pr, 1e-11)
        conv, err, lam, ntry = s.do_newton_with_linesearch(1e-11)
        ferr = _field_err(s, sy.fields)
        print(f"amp {amp} step {step}: |F| {err:.6e} / {err_o:.6e}  lambda {lam:.9e} / {lam_o:.9e}  "
              f"tries {ntry} / {len(tries_o)}  fields {ferr:.2e}", flush=True)
        assert not conv and not conv_o
        assert abs(err - err_o) <= 1e-7 * err_o
        assert ntry == len(tries_o)
        # the cubic model divides differences of f by lambda^2 (lambda down to 6e-4 after
        # ten tries, with f flat to 1e-9): the ~1e-9 relative differences of f between
        # the two LU solves reach lambda at about 1e-6
        assert abs(lam - lam_o) <= 1e-5 * lam_o
        assert ferr <= 1e-7
        saw_backtrack |= lam < 1.0
    assert saw_backtrack
    s.close()


def test_gauss_newton_reaches_the_newton_solution(ctx):
    from oracle import xcts as OX, newton as ON
    sp = OX.Space(RES, BOUNDS)
    sy = OX.XCTSSystem(sp, 1, 3, n0=0.5, a=1.0, ome=0.0)
    sy.set_fields(OX.initial_fields(sy, 1.0, OX.initial_perturbation(sy)))
    start = {n: dict(v) for n, v in sy.fields.items()}
    gn, nr = (programs.kerr(ctx, RES, BOUNDS, n0=0.5, a=1.0, ome=0.0).build(order=0, batch=256) for _ in range(2))
    _load(gn, start)
    _load(nr, start)
    ON.gauss_newton_step(ON.XctsProblem(sy), 1e-10)
    conv, err, _, _ = gn.do_newton(1e-10, "gauss-newton")
    first = _field_err(gn, sy.fields)
    print(f"Gauss-Newton first step vs oracle: {first:.2e}", flush=True)
    assert not conv and first <= 1e-3
    for it in range(12):
        conv, err, _, _ = gn.do_newton(1e-10, "gauss-newton")
        print(f"Gauss-Newton it {it + 1}: |F| {err:.3e}", flush=True)
        if conv:
            break
    assert conv
    it_nr, _ = nr.solve(1e-10, "f64")
    worst = 0.0
    for name, f in NAMES.items():
        for d in range(1, len(BOUNDS) + 1):
            a, b = gn.coef(f, d), nr.coef(f, d)
            worst = max(worst, np.abs(a - b).max() / max(1.0, np.abs(b).max()))
    print(f"Gauss-Newton vs Newton solution: {worst:.2e} (Newton: {it_nr} iterations)", flush=True)
    assert worst <= 1e-10
    gn.close()
    nr.close()


def test_reused_f64_factors_reach_the_newton_solution(ctx):
    """Solver 5: the fp64 factors of an earlier Jacobian precondition flexible GMRES on the
    current one (refactored when they stop converging quickly).  From the same guess it
    must reach ||F||_inf < 1e-10 and the same solution as the factor-every-step Newton
    (1e-10), with at least one step that reuses factors (no factorisation)."""
    from oracle import xcts as OX
    sp = OX.Space(RES, BOUNDS)
    sy = OX.XCTSSystem(sp, 1, 3, n0=0.5, a=1.0, ome=0.0)
    sy.set_fields(OX.initial_fields(sy, 1.0, OX.initial_perturbation(sy)))
    a, b = (programs.kerr(ctx, RES, BOUNDS, n0=0.5, a=1.0, ome=0.0).build(order=1, batch=256) for _ in range(2))
    _load(a, sy.fields)
    _load(b, sy.fields)
    reused = 0
    for it in range(15):
        conv, err, _, ir = a.do_newton(1e-10, "f64-reuse")
        print(f"f64-reuse it {it}: |F| {err:.3e} inner {ir[1]} factored {ir[4]}", flush=True)
        if conv:
            break
        reused += ir[4] == 0
    assert conv and reused >= 1
    b.solve(1e-10, "f64")
    worst = 0.0
    for name, f in NAMES.items():
        for d in range(1, len(BOUNDS) + 1):
            x, y = a.coef(f, d), b.coef(f, d)
            worst = max(worst, np.abs(x - y).max() / max(1.0, np.abs(y).max()))
    print(f"f64-reuse vs f64 solution: {worst:.2e}", flush=True)
    assert worst <= 1e-10
    a.close()
    b.close()
