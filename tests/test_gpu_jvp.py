"""GPU: the batched JVP Jacobian fill against the oracle's forward-AD columns and central
finite differences, with a horizon rotation (Ome = 0.05) so every term of the Kerr
system is active.

- 9x7x4, 3 domains (the CPU-runnable configuration): F and the whole J of the
  hand-traced system at batch 4096 against the oracle's (rows 1e-10 of their scale).
- 33x13x12, 3 domains: the first (field, domain) block has 33 x 141 = 4653 unknowns,
  so one launch of its tangent program carries a full batch of 4096 unit seeds
  (kad_jvp_fill of columns 0..4095).  Sampled columns of that launch (first, last and
  interior seeds) against the oracle's AD columns (1e-10 of the column scale) and
  against central differences of the device residual, h = 1e-6 (1e-6 of the column
  scale: the O(h^2) truncation of the difference).
"""
import numpy as np
import pytest
import torch

from oracle import xcts as OX
from kadath_frontera_amd import programs

pytestmark = pytest.mark.gpu

FIELD_IDX = {"P": 0, "N": 1, "bx": 2, "by": 3, "bz": 4}


def _oracle(res, bounds, ome):
    sp = OX.Space(res, bounds)
    sy = OX.XCTSSystem(sp, 1, len(bounds), n0=0.5, a=1.0, ome=ome)
    sy.set_fields(OX.initial_fields(sy, 1.0, OX.initial_perturbation(sy)))
    return sy


def _load(s, sy, ndom):
    for name, f in FIELD_IDX.items():
        for d in range(1, ndom + 1):
            v = sy.fields[name][d]
            v.coef()
            s.set_coef(f, d, v.cf)


def test_rotating_hole_F_and_J_match_oracle(ctx):
    from kadath_frontera_amd.xcts import XctsSystem
    res, bounds = (9, 7, 4), [1.0, 2.0, 4.0]
    sy = _oracle(res, bounds, 0.05)
    s = XctsSystem(ctx, res, bounds, n0=0.5, a=1.0, ome=0.05, batch=4096)
    _load(s, sy, 3)
    F_o, J_o = sy.residual(), sy.jacobian()
    F = s.residual().cpu().numpy()
    J = s.jacobian().cpu().numpy()
    scale_F = max(1.0, np.abs(F_o).max())
    row_scale = np.maximum(1.0, np.abs(J_o).max(axis=1))
    print(f"Ome 0.05: |F - F_o| {np.abs(F - F_o).max():.2e}  J rel {(np.abs(J - J_o).T / row_scale).max():.2e}",
          flush=True)
    assert np.abs(F - F_o).max() <= 1e-10 * scale_F
    assert (np.abs(J - J_o).T / row_scale).max() <= 1e-10
    s.close()


def test_full_4096_seed_launch_matches_oracle_and_central_differences(ctx):
    res, bounds = (33, 13, 12), [1.0, 2.0, 4.0]
    sy = _oracle(res, bounds, 0.05)
    lay, n_o = sy.unknown_layout()
    name0, d0, off0, cnt0 = lay[0]
    assert off0 == 0 and cnt0 >= 4096, (name0, cnt0)
    s = programs.kerr(ctx, res, bounds, n0=0.5, a=1.0, ome=0.05).build(order=0, batch=4096)
    assert s.n == n_o
    _load(s, sy, 3)
    print(f"n = {s.n}, first block {name0} in domain {d0}: {cnt0} unknowns", flush=True)
    T = s.jvp_fill(0, 4096)  # one launch: seeds 0..4095 of the first tangent program
    torch.cuda.synchronize()
    cols = [0, 1, 2, 777, 2048, 3001, 4094, 4095]
    Jo = sy.jacobian(cols=cols)
    Jd = T[:, cols].cpu().numpy()
    cscale = np.maximum(1.0, np.abs(Jo).max(axis=0))
    err_ad = (np.abs(Jd - Jo) / cscale).max()
    print(f"4096-seed launch vs oracle AD columns: {err_ad:.2e}", flush=True)
    assert err_ad <= 1e-10
    # central differences of the device residual: fields +- h e_c through kad_system_update
    h = 1e-6
    worst = 0.0
    for c in (0, 2048, 4095):
        e = torch.zeros(s.n, dtype=torch.float64, device="cuda")
        e[c] = -h
        s.update(e)            # u + h e_c (update subtracts)
        Fp = s.residual().clone()
        e[c] = 2 * h
        s.update(e)            # u - h e_c
        Fm = s.residual().clone()
        e[c] = -h
        s.update(e)            # back to u
        fd = ((Fp - Fm) / (2 * h)).cpu().numpy()
        col = T[:, c].cpu().numpy()
        err = np.abs(col - fd).max() / max(1.0, np.abs(col).max())
        worst = max(worst, err)
        print(f"column {c}: |J e_c - central difference| / scale = {err:.2e}", flush=True)
    assert worst <= 1e-6
    s.close()
