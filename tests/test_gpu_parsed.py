"""GPU: systems given as equation strings (kad_sysdef_* / kad_system_*).

- tuto5 (tutorials/tuto5.cpp): one Newton step reaches the reference's known answer
  1 + a/r at every collocation point, and check_equations is then below 1e-10.
- The Kerr program of kerr.cpp:69-99 as strings gives the same F and J as the
  hand-traced XCTS system (itself pinned against the oracle) at the seeded guess, in
  both unknown orders, and Newton from that guess converges to the same fields.
- kad_jvp_fill column tiles (fp64 and fp32) equal the matching columns of the full J.
- kad_vjp(w) = J^T w without forming J, and the torch.autograd.Function's backward is
  that product; dF/dome from the parameter program equals central differences.
"""
import numpy as np
import pytest
import torch

from kadath_frontera_amd import programs
from kadath_frontera_amd.system import Residual

pytestmark = pytest.mark.gpu

RES, BOUNDS = (9, 7, 4), [1.0, 2.0, 4.0]


def _flat_guess(s, sd_coloc, a=1.0):
    for d in range(1, len(BOUNDS) + 1):
        r, t, p = sd_coloc[d]
        P = np.where(np.isfinite(r), 1.0 + a / np.where(np.isfinite(r), r, 1.0), 1.0)
        s.set_field("P", d, P)
        s.set_field("N", d, np.ones_like(P))
        for c in "xyz":
            s.set_field("bet^" + c, d, np.zeros_like(P))


def _kerr(ctx, order, ome=0.05):
    sd = programs.kerr(ctx, RES, BOUNDS, n0=0.5, a=1.0, ome=ome)
    co = {d: sd.coloc(d) for d in range(1, len(BOUNDS) + 1)}
    s = sd.build(order=order, batch=256)
    return s, co


def test_tuto5_known_answer(ctx):
    sd, rad = programs.tuto5(ctx)
    co = {d: sd.coloc(d) for d in range(1, 4)}
    s = sd.build(order=0, batch=256)
    for d in range(1, 4):
        s.set_field("F", d, np.ones_like(co[d][0]))
    print("tuto5: Newton", flush=True)
    conv, err, _, _ = s.do_newton(1e-8)
    assert not conv and err > 1e-3
    conv, err, _, _ = s.do_newton(1e-8)
    assert conv and err < 1e-10, err  # linear problem: one step
    assert s.check_equations().max() < 1e-10
    from oracle.tuto5 import Tuto5
    t = Tuto5()
    t.newton_step()
    for d in range(1, 4):
        r = co[d][0]
        exact = np.where(np.isfinite(r), 1.0 + rad / np.where(np.isfinite(r), r, 1.0), 1.0)
        got = s.get_field("F", d)
        # radial truncation of 13 Chebyshev points: 4.0e-9 in the first shell
        assert np.abs(got - exact).max() < 1e-8, d
        # the same discrete solution as the oracle's restatement
        assert np.abs(got - t.F[d].conf()).max() < 1e-12, d
    s.close()


@pytest.mark.parametrize("order", [0, 1])
def test_kerr_strings_match_hand_traced(ctx, order):
    from kadath_frontera_amd.xcts import XctsSystem
    s, co = _kerr(ctx, order)
    x = XctsSystem(ctx, RES, BOUNDS, n0=0.5, a=1.0, ome=0.05, batch=256, order=order)
    assert (s.n, s.m) == (x.n, x.m)
    assert [L[2:] for L in s.layout] == [L[2:] for L in x.layout]
    X0 = x.perturbation(20261202)
    assert np.array_equal(X0, s.perturbation(20261202))
    x.set_guess(X0)
    _flat_guess(s, co)
    s.add_unknowns(X0)
    print(f"kerr strings order {order}: F, J", flush=True)
    F1, F2 = s.residual().cpu().numpy(), x.residual().cpu().numpy()
    assert np.abs(F1 - F2).max() < 1e-12 * max(1.0, np.abs(F2).max())
    J1, J2 = s.jacobian().cpu().numpy(), x.jacobian().cpu().numpy()
    scale = np.maximum(1.0, np.abs(J2).max(axis=0))
    assert (np.abs(J1 - J2) / scale).max() < 1e-11
    # Newton from this guess diverges at Ome = 0.05 on this space (so does the oracle's
    # restatement: 0.50, 0.048, 0.017, 0.11, 0.053, 0.020, 3.9, 2e3, ...); it converges
    # for the static hole, set through both systems' parameter entry points
    s.set_param("ome", 0.0)
    x.set_ome(0.0)
    for name, sysd in (("strings", s), ("hand-traced", x)):
        for it in range(15):
            conv, err, _, _ = sysd.do_newton(1e-10, "f64")
            print(f"{name} it {it} |F| {err:.3e}", flush=True)
            if conv:
                break
        assert conv, name
    for f in range(5):
        for d in range(1, len(BOUNDS) + 1):
            a, b = s.coef(f, d), x.coef(f, d)
            assert np.abs(a - b).max() <= 1e-10 * max(1.0, np.abs(b).max()), (f, d)
    s.close()
    x.close()


def test_jvp_fill_tiles_and_vjp(ctx):
    s, co = _kerr(ctx, 1)
    _flat_guess(s, co)
    s.add_unknowns(s.perturbation(7))
    J = s.jacobian()
    for (c0, cnt) in [(0, 1), (5, 300), (s.n - 77, 77), (1000, 2048)]:
        T64 = s.jvp_fill(c0, cnt)
        assert torch.equal(T64, J[:, c0:c0 + cnt])
        T32 = s.jvp_fill(c0, cnt, dtype=torch.float32, ld=s.m + 3)
        assert torch.equal(T32, J[:, c0:c0 + cnt].to(torch.float32))
    g = torch.Generator(device="cpu").manual_seed(3)
    w = torch.randn(s.m, generator=g, dtype=torch.float64).cuda()
    v = s.vjp(w)
    ref = J.T @ w
    assert ((v - ref).abs().max() / ref.abs().max()).item() < 1e-13
    # autograd: backward of F(u) is J^T grad_F
    u = torch.zeros(s.n, dtype=torch.float64, device="cuda", requires_grad=True)
    F = Residual.apply(u, s)
    F.backward(w)
    assert ((u.grad - ref).abs().max() / ref.abs().max()).item() < 1e-13
    # at a shifted point the backward linearises there: compare with central differences
    h = 1e-6
    rng = np.random.default_rng(5)
    cols = rng.choice(s.n, 6, replace=False)
    u1 = torch.zeros(s.n, dtype=torch.float64, device="cuda")
    u1[int(cols[0])] = 1e-4
    u1.requires_grad_(True)
    Residual.apply(u1, s).backward(w)
    for c in cols:
        e = torch.zeros(s.n, dtype=torch.float64, device="cuda")
        e[int(c)] = h
        fd = (Residual.apply(u1.detach() + e, s) - Residual.apply(u1.detach() - e, s)) @ w / (2 * h)
        assert abs(u1.grad[int(c)].item() - fd.item()) <= 1e-6 * max(1.0, abs(fd.item())), c
    s.close()


def test_param_tangent_matches_central_difference(ctx):
    s, co = _kerr(ctx, 0)
    _flat_guess(s, co)
    s.add_unknowns(s.perturbation(11))
    dF = s.param_tangent("ome").cpu().numpy()
    h = 1e-6
    s.set_param("ome", 0.05 + h)
    Fp = s.residual().cpu().numpy()
    s.set_param("ome", 0.05 - h)
    Fm = s.residual().cpu().numpy()
    s.set_param("ome", 0.05)
    fd = (Fp - Fm) / (2 * h)
    assert np.abs(dF).max() > 0
    assert np.abs(dF - fd).max() < 1e-7 * max(1.0, np.abs(fd).max())
    s.close()


def test_spin_system_matches_oracle(ctx):
    """Double unknown + integral equation: the spin-fixed hole of bh_3d_xcts_stages.cpp
    (Ome add_var'd, integ(A_ij m^i s^j / 8 pi) - chi CM^2 = 0 at the excision sphere)
    as strings, against the oracle's spin mode: F and J (the Ome column and the integral
    row included) at the fixed-Ome solution, then Newton to the same Ome and fields."""
    from oracle import xcts as OX
    res, bounds = (5, 5, 4), [1.0, 2.0, 4.0]
    sp = OX.Space(res, bounds)
    sy = OX.XCTSSystem(sp, 1, 3, n0=0.5, a=1.0, ome=0.05)
    sy.set_fields(OX.initial_fields(sy, 1.0, OX.initial_perturbation(sy)))
    for _ in range(12):
        if OX.newton_step(sy)[0] < 1e-11:
            break
    probe = OX.XCTSSystem(sp, 1, 3, n0=0.5, a=1.0, ome=0.05, spin=0.0)
    probe.set_fields(sy.fields)
    chi = 1.05 * probe.residual()[0]
    orc = OX.XCTSSystem(sp, 1, 3, n0=0.5, a=1.0, ome=0.05, spin=chi)
    orc.set_fields(sy.fields)
    F_o, J_o = orc.residual(), orc.jacobian()
    sd = programs.kerr(ctx, res, bounds, n0=0.5, a=1.0, ome=0.05, spin=chi, cm=1.0)
    s = sd.build(order=0, batch=256)
    names = {"P": 0, "N": 1, "bx": 2, "by": 3, "bz": 4}
    for name, f in names.items():
        for d in range(1, 4):
            v = sy.fields[name][d]
            v.coef()
            s.set_coef(f, d, v.cf)
    assert s.n == J_o.shape[1] and abs(s.get_double("ome") - 0.05) == 0.0
    print("spin system: F, J", flush=True)
    F = s.residual().cpu().numpy()
    assert np.abs(F - F_o).max() < 1e-12 * max(1.0, np.abs(F_o).max())
    J = s.jacobian().cpu().numpy()
    scale = np.maximum(1.0, np.abs(J_o).max(axis=0))
    assert (np.abs(J - J_o) / scale).max() < 1e-11
    assert np.abs(J[0]).max() > 0 and np.abs(J[:, 0]).max() > 0  # integral row, Ome column
    for it in range(10):
        conv, err, _, _ = s.do_newton(1e-11, "f64")
        print(f"spin device it {it} |F| {err:.3e} ome {s.get_double('ome'):.12f}", flush=True)
        if conv:
            break
    assert conv
    for _ in range(10):
        if OX.newton_step(orc)[0] < 1e-11:
            break
    assert abs(s.get_double("ome") - orc.ome) < 1e-10
    for name, f in names.items():
        for d in range(1, 4):
            v = orc.fields[name][d]
            v.coef()
            assert np.abs(s.coef(f, d) - v.cf).max() <= 1e-10 * max(1.0, np.abs(v.cf).max()), (name, d)
    s.close()


def _staged_spin(ctx, res, bounds, factor=1.05):
    """The spin-fixed hole staged from the fixed-Ome solution (oracle), target spin
    integral `factor` times the fixed-Ome one; returns the device system and the oracle
    fixed-Ome system."""
    from oracle import xcts as OX
    sp = OX.Space(res, bounds)
    sy = OX.XCTSSystem(sp, 1, len(bounds), n0=0.5, a=1.0, ome=0.05)
    sy.set_fields(OX.initial_fields(sy, 1.0, OX.initial_perturbation(sy)))
    for _ in range(12):
        if OX.newton_step(sy)[0] < 1e-11:
            break
    probe = OX.XCTSSystem(sp, 1, len(bounds), n0=0.5, a=1.0, ome=0.05, spin=0.0)
    probe.set_fields(sy.fields)
    chi = factor * probe.residual()[0]
    sd = programs.kerr(ctx, res, bounds, n0=0.5, a=1.0, ome=0.05, spin=chi, cm=1.0)
    s = sd.build(order=0, batch=256)
    for name, f in {"P": 0, "N": 1, "bx": 2, "by": 3, "bz": 4}.items():
        for d in range(1, len(bounds) + 1):
            v = sy.fields[name][d]
            v.coef()
            s.set_coef(f, d, v.cf)
    return s, chi


def test_globals_and_sensitivities_match_hand_traced(ctx):
    """M_ADM, M_Komar, J_ADM given as integral strings equal the hand-traced observables,
    and their adjoint sensitivities to Ome equal kad_xcts_sensitivity's."""
    from kadath_frontera_amd.xcts import XctsSystem
    res, bounds = (5, 5, 4), [1.0, 2.0, 4.0]
    sd = programs.kerr(ctx, res, bounds, n0=0.5, a=1.0, ome=0.05)
    co = {d: sd.coloc(d) for d in range(1, 4)}
    s = sd.build(order=0, batch=256)
    x = XctsSystem(ctx, res, bounds, n0=0.5, a=1.0, ome=0.05, batch=256, order=0)
    X0 = x.perturbation(20261202)
    x.set_guess(X0)
    _flat_guess_b(s, co, bounds)
    s.add_unknowns(X0)
    for sysd in (s, x):
        for _ in range(12):
            if sysd.do_newton(1e-11, "f64")[0]:
                break
    g, o = s.globals(), x.observables()
    print("globals", g, "observables", o, flush=True)
    for k in ("M_adm", "M_komar", "J_adm"):
        assert abs(g[k] - o[k]) < 1e-10, k
    sg, so = s.sensitivity("ome"), x.sensitivity()
    print("d/dOme", sg, so, flush=True)
    for k in ("M_adm", "M_komar", "J_adm"):
        assert abs(sg[k] - so[k]) <= 1e-8 * max(1.0, abs(so[k])), k
    s.close()
    x.close()


def _flat_guess_b(s, co, bounds, a=1.0):
    for d in range(1, len(bounds) + 1):
        r = co[d][0]
        P = np.where(np.isfinite(r), 1.0 + a / np.where(np.isfinite(r), r, 1.0), 1.0)
        s.set_field("P", d, P)
        s.set_field("N", d, np.ones_like(P))
        for c in "xyz":
            s.set_field("bet^" + c, d, np.zeros_like(P))


def test_spin_sensitivity_matches_central_differences(ctx):
    """dQ/dchi of the spin-fixed hole (chi a parameter of the integral equation) by the
    adjoint method against central differences of converged solves at chi +- h."""
    s, chi = _staged_spin(ctx, (5, 5, 4), [1.0, 2.0, 4.0])

    def solve():
        for _ in range(12):
            if s.do_newton(1e-12, "f64")[0]:
                return
        raise AssertionError("no convergence")

    solve()
    adj = s.sensitivity("chi")

    def central(h):
        s.set_param("chi", chi + h)
        solve()
        gp = s.globals()
        s.set_param("chi", chi - h)
        solve()
        gm = s.globals()
        return {k: (gp[k] - gm[k]) / (2 * h) for k in gp}

    # Q is strongly curved in chi at this scale (|chi| ~ 3e-4): the central differences
    # converge as h^2 (errors 1.31 and 0.33 at h = 1e-6, 5e-7 for M_adm ~ 5940), so the
    # check uses their Richardson extrapolation
    c1, c2 = central(1e-6), central(5e-7)
    s.set_param("chi", chi)
    print("dQ/dchi adjoint", adj, flush=True)
    for k in adj:
        rich = (4.0 * c2[k] - c1[k]) / 3.0
        print(k, "richardson", rich, flush=True)
        assert abs(adj[k] - rich) <= 1e-6 * max(1.0, abs(rich)), k
    assert abs(adj["J_adm"]) > 1.0  # the ADM angular momentum follows the horizon spin
    s.close()


def test_nan_residual_never_reads_as_converged(ctx):
    """A NaN in the fields makes ||F||_inf NaN, which must not pass the convergence test
    (do_newton.cpp:71-107 returns true only for error < precision; the device norm
    propagates NaN instead of dropping it as fmax would)."""
    sd, rad = programs.tuto5(ctx)
    s = sd.build(order=0, batch=256)
    cf = s.coef(0, 2)
    cf.flat[3] = np.nan
    s.set_coef(0, 2, cf)
    print("nan residual: tuto5 with one NaN coefficient", flush=True)
    conv, err, _, _ = s.do_newton(1e300, "f64")
    assert not conv and np.isnan(err)
    s.close()
