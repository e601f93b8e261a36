"""CPU: pins of the oracle's XCTS restatement that the mathematics fixes.

- Forward-AD Jacobian columns (do_col_J, src/System_of_eqs/solver.cpp:120-214)
  equal central finite differences of the residual F(u) (sec_member,
  solver.cpp:49-85) with h = 1e-6, where u moves along one unknown through
  xx_to_vars + affecte_tau (system_of_eqs.cpp:597-612).  A dropped product-rule
  term, a wrong sign or a transposed seed fails this.
- Newton (do_newton, src_par/do_newton.cpp:65-430) converges quadratically to
  ||F||_inf < 1e-12: a Jacobian inconsistent with F cannot.
- n_unknowns == n_rows (the reference aborts otherwise, do_newton.cpp:113-118).
"""
import numpy as np
import pytest

from oracle import xcts as OX
from oracle.kadath import vadd, vmul_scal

RES, BOUNDS = (5, 5, 4), [1.0, 2.0, 4.0]


def _system():
    sp = OX.Space(RES, BOUNDS)
    sy = OX.XCTSSystem(sp, 1, len(BOUNDS), n0=0.5, a=1.0, ome=0.05)
    X0 = OX.initial_perturbation(sy)
    sy.set_fields(OX.initial_fields(sy, 1.0, X0))
    return sy


def _shifted(sy, base, corr, h):
    return {name: {d: vadd(base[name][d], vmul_scal(corr[name][d], h)) for d in base[name]} for name in base}


def test_oracle_sizes():
    sy = _system()
    lay, n = sy.unknown_layout()
    assert n == sy.residual().size
    assert sum(c for (_, _, _, c) in lay) == n


def test_oracle_jacobian_matches_central_fd():
    sy = _system()
    lay, n = sy.unknown_layout()
    base = {name: dict(sy.fields[name]) for name in sy.fields}
    # first, middle and last unknown of every (field, domain) block
    cols = sorted({off + q for (_, _, off, cnt) in lay for q in (0, cnt // 2, cnt - 1)})
    J = sy.jacobian(cols=cols, chunk=len(cols))
    h = 1e-6
    worst = 0.0
    for j, c in enumerate(cols):
        e = np.zeros(n)
        e[c] = 1.0
        corr, _ = sy.xx_to_fields(e)
        sy.set_fields(_shifted(sy, base, corr, h))
        Fp = sy.residual()
        sy.set_fields(_shifted(sy, base, corr, -h))
        Fm = sy.residual()
        sy.set_fields(base)
        fd = (Fp - Fm) / (2 * h)
        scale = max(1.0, np.abs(J[:, j]).max())
        worst = max(worst, np.abs(fd - J[:, j]).max() / scale)
    assert worst < 1e-7, worst


def test_oracle_newton_is_quadratic():
    sy = _system()
    errs = []
    for _ in range(8):
        err, *_ = OX.newton_step(sy)
        errs.append(err)
        if err < 1e-12:
            break
    assert errs[-1] < 1e-12, errs
    # once in the basin (err < 1e-2): e_{k+1} <= C e_k^2
    tail = [e for e in errs if e < 1e-2]
    for a, b in zip(tail, tail[1:]):
        assert b <= 50.0 * a * a + 1e-13, errs


def test_oracle_observables_closed_forms():
    """Domain_compact::integ surface integrals (reader.cpp:114-125) on fields whose
    ADM mass, Komar mass and angular momentum are known exactly."""
    sy = _system()
    sy.set_fields(OX.analytic_fields(sy, 0.7, 0.0))
    o = sy.observables()
    assert abs(o["M_adm"] - 0.7) < 1e-12, o
    assert abs(o["M_komar"] - 0.7) < 1e-12, o
    assert abs(o["J_adm"]) < 1e-12, o
    sy.set_fields(OX.analytic_fields(sy, 0.0, 0.3))
    o = sy.observables()
    assert abs(o["J_adm"] - 0.3) < 1e-12, o
    assert abs(o["M_adm"]) < 1e-12 and abs(o["M_komar"]) < 1e-12, o
    # the flat guess P = 1 + a/r has M_ADM = 2a, N = 1 and bet = 0 carry nothing
    sy.set_fields(OX.initial_fields(sy, 1.0))
    o = sy.observables()
    assert abs(o["M_adm"] - 2.0) < 1e-12 and abs(o["M_komar"]) < 1e-12 and abs(o["J_adm"]) < 1e-12, o
