"""CPU: the Kadath binary reader/writer (kadath_frontera_amd/kadath_io.py) on the
reference's own converged FUKA black hole (tests/golden/converged_BH_TOTAL_BC.0.5.0.0.09.*,
from codes/PythonTools/Example_id/, solved by the BH solver's TOTAL_BC stage at
resolution 9x9x8; codes/FUKAv2_Solvers/BH/README.md:48-72 prints its diagnostics).

- The whole file is consumed by the save chain (Space_adapted_bh, 4 domains; conf,
  lapse, shift), and writing it back gives the same bytes.
- The domain radii equal the .info file's rin / rmid / rout.
- The stored fields reproduce what the reference's reader prints: on the horizon
  (domain 2, r index 0) the lapse is the .info's fixed_lapse and psi = 1.52539.
- The excised domains 0-1 hold zero fields (the eq_full "N = 0, P = 0, bet^i = 0" rows
  of the reference's system are these coefficients).
- The stored spectral bases are the standard ones this package assigns: theta basis
  COS_EVEN / SIN_ODD by phi coefficient pair for P, N, bet^x, bet^y and COS_ODD /
  SIN_EVEN for bet^z (domain_shell.cpp:301-307 and the anti base).
- The TOTAL_BC system traced on this space has one equation more than unknowns (the
  homothetic radius is not an unknown here).
"""
import os

import numpy as np
import pytest

from kadath_frontera_amd import kadath_io as K
from oracle import spectral as S

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
DAT = os.path.join(GOLDEN, "converged_BH_TOTAL_BC.0.5.0.0.09.dat")
INFO = os.path.join(GOLDEN, "converged_BH_TOTAL_BC.0.5.0.0.09.info")


@pytest.fixture(scope="module")
def solution():
    return K.read_solution(DAT), K.read_info(INFO)


def test_layout_and_round_trip(solution, tmp_path):
    (sp, fields), info = solution
    kinds = [d.kind for d in sp.domains]
    assert kinds == [K.NUCLEUS, K.SHELL_OUTER_ADAPTED, K.SHELL_INNER_ADAPTED, K.COMPACT]
    assert all(d.nbr_points == (9, 9, 8) and d.nbr_coefs == (9, 9, 10) for d in sp.domains)
    out = tmp_path / "rt.dat"
    K.write_solution(str(out), sp, fields)
    assert out.read_bytes() == open(DAT, "rb").read()


def test_radii_and_horizon_values(solution):
    (sp, (conf, lapse, shift)), info = solution
    bh = info["bh"]
    for got, want in zip(sp.radii(), (bh["rin"], bh["rmid"], bh["rout"])):
        assert got == pytest.approx(want, rel=1e-15)
    N = lapse[2].conf_values().reshape(8, 9, 9)[:, :, 0]
    P = conf[2].conf_values().reshape(8, 9, 9)[:, :, 0]
    np.testing.assert_allclose(N, bh["fixed_lapse"], rtol=1e-14)
    np.testing.assert_allclose(P, 1.52539, atol=5e-6)
    for d in (0, 1):
        assert np.abs(conf[d].conf).max() == 0 and np.abs(lapse[d].conf).max() == 0
        assert all(np.abs(c[d].conf).max() == 0 for c in shift["components"])


def test_stored_bases_are_the_standard_ones(solution):
    (sp, (conf, lapse, shift)), info = solution
    npc = sp.domains[2].nbr_coefs[2]
    std = [S.COS_EVEN if (k // 2) % 2 == 0 else S.SIN_ODD for k in range(npc)]
    anti = [S.COS_ODD if (k // 2) % 2 == 0 else S.SIN_EVEN for k in range(npc)]
    for d in range(2, 4):
        for v in (conf[d], lapse[d], shift["components"][0][d], shift["components"][1][d]):
            assert [v.theta_base(k) for k in range(npc)] == std
            assert int(v.bases[2][0][0]) == S.COSSIN
        assert [shift["components"][2][d].theta_base(k) for k in range(npc)] == anti


def test_total_bc_system_traces(solution):
    from kadath_frontera_amd import programs
    (sp, fields), info = solution
    sd = programs.fuka_bh_total_bc(None, sp, info)
    i, rows = sd.host_trace(0)
    assert i["n_rows"] == i["n_unknowns"] + 1
    # two integral rows first, then the domain-2 equations
    assert [r[0] for r in rows[:2]] == [4, 4]
