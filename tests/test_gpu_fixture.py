"""GPU: the reference's own converged FUKA black hole (tests/golden/
converged_BH_TOTAL_BC.0.5.0.0.09.*, from codes/PythonTools/Example_id/) evaluated by the
device system built from the TOTAL_BC stage's equation strings
(programs.fuka_bh_total_bc; bh_3d_xcts_stages.cpp:36-92, bh_3d_xcts_solver_imp.cpp:118-159).

The reference stopped its Newton iteration there on ||F||_inf < 1e-8 (the .info's
solver_precision; do_newton returns at entry when the error is below it), so the device
residual at the stored fields must be below 1e-8 too; its reader prints the global
quantities (codes/FUKAv2_Solvers/BH/README.md:58-68): Komar mass 4.97314e-01, ADM mass
5.00087e-01, Mirr 0.50000, Chi 0.  These pin the device's equations (operators, tensor
algebra, excision conditions, integrals, the row selection) against the reference's own
discretisation, which no restatement of ours does.

The homothetic radius of domains 1-2 is not an unknown of the device system (it stays at
its stored value), so the system has one more row than unknowns: the residual and the
integrals are evaluated, and do_newton refuses it as the reference does for a non-square
system (do_newton.cpp:113-118).
"""
import os

import numpy as np
import pytest

from kadath_frontera_amd import kadath_io as K, programs
from kadath_frontera_amd._lib import KadError

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
STEM = os.path.join(GOLDEN, "converged_BH_TOTAL_BC.0.5.0.0.09")


def test_converged_black_hole_fixture(ctx):
    sp, fields = K.read_solution(STEM + ".dat")
    info = K.read_info(STEM + ".info")
    s = programs.fuka_bh_total_bc(ctx, sp, info).build(order=0, batch=512)
    programs.load_fuka_bh_fields(s, fields)
    F = s.residual().cpu().numpy()
    errs = s.check_equations()
    print("fixture: n_rows", F.size, "max |F|", np.abs(F).max(), flush=True)
    print("per equation:", " ".join(f"{e:.2e}" for e in errs), flush=True)
    g = s.globals()
    print("globals:", {k: float(v) for k, v in g.items()}, flush=True)
    assert np.abs(F).max() < 1e-8
    assert abs(g["M_komar"] - 4.97314e-01) <= 5e-7
    assert abs(g["M_adm"] - 5.00087e-01) <= 5e-7
    assert abs(np.sqrt(g["Mirr2"]) - 0.5) <= 5e-6
    assert abs(g["S"]) <= 1e-9
    with pytest.raises(KadError):
        s.do_newton(1e-12, "f64")
    s.close()
