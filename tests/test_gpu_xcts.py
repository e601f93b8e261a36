"""GPU: the device Newton path (residual, batched-JVP Jacobian, LU, update) against
the oracle's restatement of System_of_eqs on the 9x7x4 three-domain space.

Tolerances: F and J rows 1e-10 relative to the row scale (FP64 roundoff of O(10^3)
operations and different summation orders); converged fields 1e-10 relative
(both Newton iterations stop at ||F||_inf < 1e-10, the acceptance bound).
"""
import numpy as np
import pytest
import torch

from oracle import xcts as OX

pytestmark = pytest.mark.gpu

RES, BOUNDS = (9, 7, 4), [1.0, 2.0, 4.0]
FIELD_IDX = {"P": 0, "N": 1, "bx": 2, "by": 3, "bz": 4}


@pytest.fixture(scope="module")
def oracle_sys():
    sp = OX.Space(RES, BOUNDS)
    sy = OX.XCTSSystem(sp, 1, len(BOUNDS), n0=0.5, a=1.0)
    X0 = OX.initial_perturbation(sy)
    sy.set_fields(OX.initial_fields(sy, 1.0, X0))
    return sy, X0


@pytest.fixture(scope="module")
def gpu_sys(ctx):
    from kadath_frontera_amd.xcts import XctsSystem
    s = XctsSystem(ctx, RES, BOUNDS, n0=0.5, a=1.0, batch=128)
    yield s
    s.close()


def test_sizes_and_layout(oracle_sys, gpu_sys):
    sy, _ = oracle_sys
    lay, n = sy.unknown_layout()
    assert gpu_sys.n == n == gpu_sys.m
    assert [(FIELD_IDX[a], b, c, d) for (a, b, c, d) in lay] == gpu_sys.layout


def test_residual_matches_oracle(oracle_sys, gpu_sys):
    sy, X0 = oracle_sys
    gpu_sys.set_guess(X0)
    F = gpu_sys.residual().cpu().numpy()
    Fo = sy.residual()
    scale = max(1.0, np.abs(Fo).max())
    assert np.abs(F - Fo).max() < 1e-11 * scale, np.abs(F - Fo).max()


def test_jacobian_matches_oracle(oracle_sys, gpu_sys):
    sy, X0 = oracle_sys
    gpu_sys.set_guess(X0)
    J = gpu_sys.jacobian().cpu().numpy()
    Jo = sy.jacobian()
    colscale = np.maximum(1.0, np.abs(Jo).max(axis=0))
    err = (np.abs(J - Jo) / colscale).max()
    assert err < 1e-11, err
    # structure: entries the oracle has at exactly zero are zero up to roundoff
    assert (np.abs(J)[Jo == 0] / np.broadcast_to(colscale, J.shape)[Jo == 0]).max() < 1e-12


def test_jacobian_fill_streams_give_the_same_J(oracle_sys, gpu_sys):
    # the tangent programs dealt to 1-4 streams (batch 128: several column tiles per
    # program, run one after another on their stream's arena) fill J bit for bit alike,
    # and the value program split by domain over the same streams gives the same F
    _, X0 = oracle_sys
    gpu_sys.set_guess(X0)
    try:
        gpu_sys.set_option(fill_streams=1)
        F1 = gpu_sys.residual().clone()
        J1 = gpu_sys.jacobian().clone()
        for k in (2, 3, 4):
            gpu_sys.set_option(fill_streams=k)
            assert torch.equal(gpu_sys.residual(), F1), k
            assert torch.equal(gpu_sys.jacobian(), J1), k
        with pytest.raises(RuntimeError):
            gpu_sys.set_option(fill_streams=5)
    finally:
        gpu_sys.set_option(fill_streams=3)


def test_fused_element_wise_instructions_give_the_same_J(oracle_sys, gpu_sys):
    # the element-wise instructions applied in the spectral pass that feeds them, or
    # launched on their own: the same J bit for bit
    _, X0 = oracle_sys
    gpu_sys.set_guess(X0)
    try:
        gpu_sys.set_option(fuse=False)
        J0 = gpu_sys.jacobian().clone()
        gpu_sys.set_option(fuse=True)
        assert torch.equal(gpu_sys.jacobian(), J0)
    finally:
        gpu_sys.set_option(fuse=True)


@pytest.mark.parametrize("solver", ["f64", "f32-gmres"])
def test_newton_converges_to_oracle_solution(oracle_sys, gpu_sys, solver):
    sy, X0 = oracle_sys
    # oracle solution (cached on the fixture)
    if not hasattr(sy, "_solved"):
        sy.set_fields(OX.initial_fields(sy, 1.0, X0))
        for _ in range(12):
            err, *_ = OX.newton_step(sy)
            if err < 1e-10:
                break
        sy._solved = {}
        for k in sy.fields:
            sy._solved[k] = {}
            for d, v in sy.fields[k].items():
                v.coef()
                sy._solved[k][d] = v.cf
        sy._err = sy.check_equations()
    gpu_sys.set_guess(X0)
    conv, it = False, 0
    while not conv and it < 15:
        conv, err, times, ir = gpu_sys.do_newton(1e-10, solver)
        print(f"{solver} it {it} err {err:.3e} times {times} ir {ir}", flush=True)
        it += 1
    assert conv and err < 1e-10
    for name, f in FIELD_IDX.items():
        for d in range(1, len(BOUNDS) + 1):
            ref = sy._solved[name][d]
            got = gpu_sys.coef(f, d)
            assert np.abs(got - ref).max() <= 1e-10 * max(1.0, np.abs(ref).max()), (name, d)


@pytest.fixture(scope="module")
def gpu_sys_dm(ctx):
    from kadath_frontera_amd.xcts import XctsSystem
    s = XctsSystem(ctx, RES, BOUNDS, n0=0.5, a=1.0, batch=128, order=1)
    yield s
    s.close()


def _perms(ref, dm):
    """Row and column permutations taking the reference-order system to the
    domain-major one: dm[i, j] = ref[rows[i], cols[j]]."""
    cols = np.zeros(ref.n, dtype=np.int64)
    for (f, d, off, cnt) in dm.layout:
        (o0,) = [o for (ff, dd, o, c) in ref.layout if (ff, dd) == (f, d)]
        cols[off:off + cnt] = np.arange(o0, o0 + cnt)
    rows = np.zeros(ref.n, dtype=np.int64)
    ref_rows = {}
    for q, (k, d, c, r0, cnt) in enumerate(ref.row_layout):
        ref_rows.setdefault((k, d, c), []).append(r0)
    seen = {}
    for (k, d, c, r0, cnt) in dm.row_layout:
        j = seen.get((k, d, c), 0)
        seen[(k, d, c)] = j + 1
        rows[r0:r0 + cnt] = np.arange(ref_rows[(k, d, c)][j], ref_rows[(k, d, c)][j] + cnt)
    return rows, cols


def test_domain_major_is_a_permutation(oracle_sys, gpu_sys, gpu_sys_dm):
    """The domain-major system is the reference-order one with rows and columns
    permuted: same guess, F and J equal entry for entry (same kernels, same data)."""
    _, X0 = oracle_sys
    rows, cols = _perms(gpu_sys, gpu_sys_dm)
    assert sorted(rows.tolist()) == list(range(gpu_sys.n)) and sorted(cols.tolist()) == list(range(gpu_sys.n))
    gpu_sys.set_guess(X0)
    Xdm = np.empty_like(X0)
    Xdm[:] = X0[cols]
    gpu_sys_dm.set_guess(Xdm)
    F0 = gpu_sys.residual().cpu().numpy()
    F1 = gpu_sys_dm.residual().cpu().numpy()
    assert np.array_equal(F1, F0[rows])
    J0 = gpu_sys.jacobian().cpu().numpy()
    J1 = gpu_sys_dm.jacobian().cpu().numpy()
    assert np.array_equal(J1, J0[np.ix_(rows, cols)])
    # the seeded perturbation is drawn in the reference order whatever the layout
    assert np.array_equal(gpu_sys_dm.perturbation(1234)[:], gpu_sys.perturbation(1234)[cols])


@pytest.mark.parametrize("solver", ["f64", "f32-gmres"])
def test_domain_major_newton_matches_oracle(oracle_sys, gpu_sys_dm, solver):
    sy, X0 = oracle_sys
    if not hasattr(sy, "_solved"):
        pytest.skip("needs the oracle solution of test_newton_converges_to_oracle_solution")
    lay, _ = sy.unknown_layout()
    ref_off = {(FIELD_IDX[a], b): c for (a, b, c, d) in lay}
    Xdm = np.empty_like(X0)
    for (f, d, off, cnt) in gpu_sys_dm.layout:
        Xdm[off:off + cnt] = X0[ref_off[(f, d)]:ref_off[(f, d)] + cnt]
    gpu_sys_dm.set_guess(Xdm)
    conv, it = False, 0
    while not conv and it < 15:
        conv, err, times, ir = gpu_sys_dm.do_newton(1e-10, solver)
        print(f"domain-major {solver} it {it} err {err:.3e} ir {ir}", flush=True)
        it += 1
    assert conv and err < 1e-10
    for name, f in FIELD_IDX.items():
        for d in range(1, len(BOUNDS) + 1):
            ref = sy._solved[name][d]
            got = gpu_sys_dm.coef(f, d)
            assert np.abs(got - ref).max() <= 1e-10 * max(1.0, np.abs(ref).max()), (name, d)


def _load_oracle_fields(sy, s):
    for name, f in FIELD_IDX.items():
        for d in range(1, len(BOUNDS) + 1):
            v = sy.fields[name][d]
            v.coef()
            s.set_coef(f, d, v.cf.ravel())


def test_observables_match_oracle(oracle_sys, gpu_sys):
    """ADM mass, Komar mass and J_ADM (reader.cpp:114-125) of identical fields: the
    device surface integrals against the oracle's, on the oracle's converged solution
    (when the Newton test above produced one) and on closed-form fields."""
    sy, X0 = oracle_sys
    cases = []
    if hasattr(sy, "_solved"):
        cases.append(("solution", None))
    cases += [("mass", (0.7, 0.0)), ("spin", (0.0, 0.3))]
    for label, mj in cases:
        if mj is not None:
            sy.set_fields(OX.analytic_fields(sy, *mj))
        _load_oracle_fields(sy, gpu_sys)
        got = gpu_sys.observables()
        ref = sy.observables()
        print(label, got, ref, flush=True)
        for k in ("M_adm", "M_komar", "J_adm"):
            assert abs(got[k] - ref[k]) <= 1e-9 * max(1.0, abs(ref[k])), (label, k, got, ref)
        if mj is not None:
            m, J = mj
            assert abs(got["M_adm"] - m) < 1e-11 and abs(got["M_komar"] - m) < 1e-11, (label, got)
            assert abs(got["J_adm"] - J) < 1e-11, (label, got)
