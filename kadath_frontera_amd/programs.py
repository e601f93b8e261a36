"""The reference's example programs as SystemDef declarations.

tuto5(): tutorials/tuto5.cpp:60-90 -- Laplace equation with a Robin inner boundary
  condition; the exact solution is 1 + a/r (a = inner radius), reached in one Newton
  step (the reference's own known-answer program).
kerr(): codes/Par_version/Kerr/kerr.cpp:69-99 in the Cartesian tensor basis of the
  FUKA BH solver (bh_3d_xcts_solver_imp.cpp:119-159): the XCTS black hole with lapse
  n0 on the excision sphere r = a and horizon rotation ome (a parameter, so dF/dome
  is available), the same system the hand-traced XCTS path builds.
kerr(spin=chi): the horizon rotation becomes a double unknown fixed by the spin
  integral, integ(A_ij m^i s^j / 8 pi) - chi CM^2 = 0 at the excision sphere, as the
  FUKA BH solver fixes it (bh_3d_xcts_stages.cpp:66,90).
"""
import numpy as np

from .system import SystemDef, INNER_BC, OUTER_BC


def tuto5(ctx, res=(13, 5, 4), bounds=(1.0, 2.0, 4.0)):
    sd = SystemDef(ctx, res, bounds)
    ndom = len(bounds) + 1
    sd.add_var("F")
    rad = float(bounds[0])  # radius of the inner boundary of domain 1
    sd.add_cst("a", rad)
    for d in range(1, ndom):
        sd.add_eq_inside(d, "Lap(F)=0")
    sd.add_eq_bc(1, INNER_BC, "dn(F) + 0.5*F/a = 0")
    sd.add_eq_bc(ndom - 1, OUTER_BC, "F=1")
    for d in range(1, ndom - 1):
        sd.add_eq_matching(d, OUTER_BC, "F")
        sd.add_eq_matching(d, OUTER_BC, "dn(F)")
    return sd, rad


def kerr_constants(sd):
    """Radial unit vector s^i, rotation generator m^i = (-y, x, 0) and mh^i = m^i / r
    (bounded at infinity) at the collocation points of the solved domains
    (component-major, then domain)."""
    s_c, m_c, h_c = [[], [], []], [[], [], []], [[], [], []]
    for d in range(sd.dmin, sd.dmax + 1):
        r, t, p = sd.coloc(d)
        st, ct, cp, sp = np.sin(t), np.cos(t), np.cos(p), np.sin(p)
        for i, v in enumerate((st * cp, st * sp, ct)):
            s_c[i].append(v.ravel())
        rr = np.where(np.isfinite(r), r, 0.0)
        for i, v in enumerate((-rr * st * sp, rr * st * cp, np.zeros_like(rr))):
            m_c[i].append(v.ravel())
        for i, v in enumerate((-st * sp, st * cp, np.zeros_like(rr))):
            h_c[i].append(v.ravel())
    cat = lambda c: np.array([np.concatenate(x) for x in c])  # noqa: E731
    return cat(s_c), cat(m_c), cat(h_c)


def kerr(ctx, res, bounds, n0=0.5, a=1.0, ome=0.0, spin=None, cm=1.0):
    sd = SystemDef(ctx, res, bounds)
    sd.add_var("P")
    sd.add_var("N")
    sd.add_var("bet", 1)
    sd.add_cst("a", a)
    s_c, m_c, h_c = kerr_constants(sd)
    sd.add_cst_field("s", s_c, 1)
    sd.add_cst_field("m", m_c, 1)
    sd.add_cst_field("mh", h_c, 1)
    sd.add_cst("n0", n0)
    sd.add_cst("PI", np.pi)
    if spin is None:
        sd.add_param("ome", ome)
    else:
        sd.add_var_double("ome", ome)
        sd.add_param("chi", spin)
        sd.add_cst("CM", cm)
    sd.set_flat_metric("f")
    sd.add_def("A^ij = (D^i bet^j + D^j bet^i - 2. / 3.* D_k bet^k * f^ij) /2. / N")
    if spin is not None:
        sd.add_def("intS = A_ij * m^i * s^j / 8. / PI")
        sd.add_eq_int(1, INNER_BC, "integ(intS) - chi * CM * CM = 0")
    sd.space_add_inner_bc("N=n0")
    sd.space_add_inner_bc("bet^i = n0 / P^2 * s^i + ome * a * m^i")
    sd.space_add_inner_bc("dn(P) + 0.5 * P / a + P^3 * A_ij * s^i * s^j /4.= 0")
    sd.space_add_eq("D_i D^i N + 2 * D_i P * D^i N / P - N * P^4 * A_ij *A^ij= 0", "N", "dn(N)")
    sd.space_add_eq("D_i D^i P + P^5 *A_ij * A^ij / 8= 0", "P", "dn(P)")
    sd.space_add_eq("D_j A^ij + 6 * A^ij * D_j P / P =0", "bet^i", "dn(bet^i)")
    sd.space_add_outer_bc("N=1")
    sd.space_add_outer_bc("P=1")
    sd.space_add_outer_bc("bet^i=0")
    # global quantities at infinity (codes/Par_version/Kerr/reader.cpp:114-125 in the
    # Cartesian basis): ADM mass, Komar mass, ADM angular momentum about z
    last = len(bounds)
    sd.add_global("M_adm", last, OUTER_BC, "-integ(dn(P)) / 2. / PI")
    sd.add_global("M_komar", last, OUTER_BC, "integ(dn(N)) / 4. / PI")
    sd.add_global("J_adm", last, OUTER_BC, "-integ(multr(A_ij * s^i * mh^j)) / 8. / PI")
    return sd
