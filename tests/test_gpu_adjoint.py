"""GPU: adjoint sensitivities of the converged solution (kad_xcts_sensitivity and the
torch.autograd binding) against central finite differences.

- dQ/dOme for Q = (M_ADM, M_Komar, J_ADM): adjoint solve J^T lambda = dQ/du with the LU
  factors versus Richardson-extrapolated central differences of full Newton solves at
  Ome +- h, Ome +- h/2 (truncation O(h^4)); tolerance 1e-7 relative.
- dQ/du (the observables' tangent programs) versus central differences h = 1e-6 of
  the surface integrals, moving one unknown through xx_to_vars + affecte_tau
  (kad_xcts_update); tolerance 1e-7 relative to the quantity's scale.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

RES, BOUNDS, OME = (9, 7, 4), [1.0, 2.0, 4.0], 0.02
KEYS = ("M_adm", "M_komar", "J_adm")


@pytest.fixture(scope="module")
def sys_dm(ctx):
    from kadath_frontera_amd.xcts import XctsSystem
    s = XctsSystem(ctx, RES, BOUNDS, n0=0.5, a=1.0, ome=OME, batch=512, order=1)
    yield s
    s.close()


def test_adjoint_dQ_dOme_matches_finite_differences(sys_dm):
    from kadath_frontera_amd.adjoint import solved_observables
    X0 = sys_dm.perturbation(20261202)
    ome = torch.tensor(OME, dtype=torch.float64, requires_grad=True)
    Q = solved_observables(sys_dm, ome, guess=X0)
    adj = np.array([torch.autograd.grad(Q[q], ome, retain_graph=True)[0].item() for q in range(3)])

    def q_at(w):
        return solved_observables(sys_dm, torch.tensor(w, dtype=torch.float64), guess=X0).numpy()

    h = 2e-4
    fd_h = (q_at(OME + h) - q_at(OME - h)) / (2 * h)
    fd_h2 = (q_at(OME + h / 2) - q_at(OME - h / 2)) / h
    fd = (4 * fd_h2 - fd_h) / 3
    print("adjoint", adj, "richardson fd", fd, flush=True)
    assert np.all(np.abs(adj - fd) <= 1e-7 * np.maximum(1.0, np.abs(fd))), (adj, fd)


def test_observable_tangents_match_finite_differences(sys_dm):
    X0 = sys_dm.perturbation(20261202)
    sys_dm.set_ome(OME)
    sys_dm.solve(1e-10, "f64", X0)
    _, dqdu = sys_dm.sensitivity(with_dqdu=True)
    # unknowns of the outer domain (the only ones the surface integrals read): first,
    # middle and last of every field block
    cols = sorted({off + q for (f, d, off, cnt) in sys_dm.layout if d == len(BOUNDS)
                   for q in (0, cnt // 2, cnt - 1)})
    h = 1e-6
    q0 = sys_dm.observables()
    scale = {k: max(1.0, abs(q0[k])) for k in KEYS}
    for c in cols:
        e = torch.zeros(sys_dm.n, dtype=torch.float64, device="cuda")
        e[c] = -h                      # update subtracts: fields += h affecte_tau(e_c)
        sys_dm.update(e)
        qp = sys_dm.observables()
        e[c] = 2 * h                   # back to the solution, then h below it
        sys_dm.update(e)
        qm = sys_dm.observables()
        e[c] = -h
        sys_dm.update(e)
        for i, k in enumerate(KEYS):
            fd = (qp[k] - qm[k]) / (2 * h)
            assert abs(fd - dqdu[i, c]) <= 1e-7 * scale[k], (c, k, fd, dqdu[i, c])
    # unknowns of the other domains do not enter the surface integrals
    inner = [off for (f, d, off, cnt) in sys_dm.layout if d < len(BOUNDS)]
    assert np.all(dqdu[:, inner] == 0.0)
