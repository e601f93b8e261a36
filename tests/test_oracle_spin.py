"""CPU: the oracle's integral equation and double unknown (the spin-fixed hole:
Ome is add_var'd and fixed by integ(A_ij m^i s^j / 8 pi) - chi CM^2 = 0 at the
excision sphere, bh_3d_xcts_stages.cpp:66,90; Eq_int, eq_int.cpp:38-51).

- The Jacobian's Ome column (every equation's tangent in Ome, solver.cpp:179-186) and
  its integral row equal central differences of the residual (h = 1e-6).
- Staged from the fixed-Ome solution with a 5% larger spin target, Newton converges
  quadratically to ||F|| < 1e-13 and moves Ome.
"""
import numpy as np

from oracle import xcts as OX

RES, BOUNDS = (5, 5, 4), [1.0, 2.0, 4.0]


def _perturbed(spin):
    sy = OX.XCTSSystem(OX.Space(RES, BOUNDS), 1, 3, n0=0.5, a=1.0, ome=0.05, spin=spin)
    X0 = OX.initial_perturbation(sy)
    sy.set_fields(OX.initial_fields(sy, 1.0, X0))
    return sy


def test_ome_column_and_integral_row_match_central_differences():
    sy = _perturbed(0.2)
    J = sy.jacobian()
    h = 1e-6
    sy.ome += h
    Fp = sy.residual()
    sy.ome -= 2 * h
    Fm = sy.residual()
    sy.ome += h
    fd = (Fp - Fm) / (2 * h)
    assert np.abs(fd).max() > 1e-3
    assert np.abs(J[:, 0] - fd).max() <= 1e-7 * max(1.0, np.abs(fd).max())
    # integral row against moves of single unknowns of the first domain
    lay, n = sy.unknown_layout()
    base = sy.fields
    for (name, d, off, cnt) in lay:
        if d != 1:
            continue
        for c in (off, off + cnt // 2, off + cnt - 1):
            e = np.zeros(n)
            e[c] = h
            vals = []
            for sgn in (1, -1):
                corr, _ = sy.xx_to_fields(sgn * e)
                sy.set_fields({k: {dd: OX.vadd(base[k][dd], corr[k][dd]) for dd in base[k]} for k in base})
                vals.append(sy.residual()[0])
            sy.set_fields(base)
            fdc = (vals[0] - vals[1]) / (2 * h)
            assert abs(J[0, c] - fdc) <= 1e-7 * max(1.0, abs(fdc)), (name, c)


def test_staged_spin_newton_converges():
    sp = OX.Space(RES, BOUNDS)
    sy = OX.XCTSSystem(sp, 1, 3, n0=0.5, a=1.0, ome=0.05)
    X0 = OX.initial_perturbation(sy)
    sy.set_fields(OX.initial_fields(sy, 1.0, X0))
    for _ in range(12):
        if OX.newton_step(sy)[0] < 1e-11:
            break
    probe = OX.XCTSSystem(sp, 1, 3, n0=0.5, a=1.0, ome=0.05, spin=0.0)
    probe.set_fields(sy.fields)
    jh = probe.residual()[0]
    s3 = OX.XCTSSystem(sp, 1, 3, n0=0.5, a=1.0, ome=0.05, spin=1.05 * jh)
    s3.set_fields(sy.fields)
    errs = []
    for _ in range(8):
        errs.append(OX.newton_step(s3)[0])
        if errs[-1] < 1e-13:
            break
    assert errs[-1] < 1e-13
    assert abs(s3.ome - 0.05) > 1e-3
