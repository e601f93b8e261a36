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
fuka_bh_total_bc(): the TOTAL_BC stage of the FUKA isolated black hole
  (bh_3d_xcts_stages.cpp:36-92 with syst_init, bh_3d_xcts_solver_imp.cpp:118-159) on a
  space read from a FUKA solution file (kadath_io.read_solution).
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


def fuka_coord_fields(sd):
    """The coordinate vectors of the FUKA solvers (CoordFields, coord_fields.hpp:260-360)
    about the origin, at the collocation points of the solved domains (component-major,
    then domain): mg = rot_z = (-y, x, 0), sm = einf = e_rad = x^i / r, ex, ey.  In the
    compactified domain the Cartesian coordinates are replaced by the surface ones,
    x^i / r (get_cart_surr), so there mg = (-sin t sin p, sin t cos p, 0)."""
    mg, er, ex, ey = ([[], [], []] for _ in range(4))
    for d in range(sd.dmin, sd.dmax + 1):
        r, t, p = sd.coloc(d)
        st, ct, cp, sp = np.sin(t), np.cos(t), np.cos(p), np.sin(p)
        rr = np.ones_like(r) if d == sd.dmax else r  # the last domain is the compactified one
        for i, v in enumerate((-rr * st * sp, rr * st * cp, np.zeros_like(r))):
            mg[i].append(v.ravel())
        for i, v in enumerate((st * cp, st * sp, ct)):
            er[i].append(v.ravel())
        for i in range(3):
            ex[i].append(np.full(r.size, 1.0 if i == 0 else 0.0))
            ey[i].append(np.full(r.size, 1.0 if i == 1 else 0.0))
    cat = lambda c: np.array([np.concatenate(x) for x in c])  # noqa: E731
    return cat(mg), cat(er), cat(ex), cat(ey)


def fuka_bh_total_bc(ctx, space, info):
    """System of the TOTAL_BC stage of the FUKA black-hole solver on a Space_adapted_bh
    read from file (kadath_io.read_solution), with the .info parameters `info`
    (kadath_io.read_info).  Solved domains: 2 (the shell outside the excision sphere,
    whose homothetic inner radius is held fixed at the file's value) to the compactified
    domain; the excised domains 0-1 (eq_full "N = 0", ...) are not part of it.  The
    homothetic radius is not an unknown here, so the system has one more equation (the
    irreducible-mass integral) than unknowns: it is evaluated (residual, equations,
    global quantities), not solved.
    Global quantities: M_adm and M_komar at infinity (syst_init's intMadm / intMk), Mirr2
    and S on the excision sphere (print_diagnostics, bh_3d_xcts_solver_imp.cpp:161-219)."""
    bh = info["bh"]
    radii = space.radii()
    res = tuple(space.domains[0].nbr_points)
    ndom = len(space.domains)
    sd = SystemDef(ctx, res, radii, dmin=2, dmax=ndom - 1)
    sd.set_adapted(2)  # Domain_shell_inner_homothetic
    sd.set_flat_metric("f")
    sd.add_cst("4piG", bh["qpig"])
    sd.add_cst("PI", np.pi)
    sd.add_cst("M", bh["mirr"])
    sd.add_cst("CM", bh["mch"])
    sd.add_cst("chi", bh["chi"])
    sd.add_var_double("ome", bh["omega"])
    mg, er, ex, ey = fuka_coord_fields(sd)
    sd.add_cst_field("mg", mg, 1)
    sd.add_cst_field("sm", er, 1)
    sd.add_cst_field("ex", ex, 1)
    sd.add_cst_field("ey", ey, 1)
    sd.add_cst_field("einf", er, 1)
    sd.add_var("P")
    sd.add_var("N")
    sd.add_var("bet", 1)
    sd.add_def("NP = P*N")
    sd.add_def("Ntilde = N / P^6")
    sd.add_def("A^ij = (D^i bet^j + D^j bet^i - 2. / 3.* D_k bet^k * f^ij) / 2. / Ntilde")
    sd.add_def("intMsq= P^4 / 16. / PI")
    sd.add_def("intMadm = - einf^i * D_i P / 4piG * 2")
    sd.add_def("intMk = einf^i * D_i N / 4piG")
    sd.add_cst("xboost", bh.get("velx", 0.0))
    sd.add_cst("yboost", bh.get("vely", 0.0))
    sd.add_def("intS  = A_ij * mg^i * sm^j / 8. / PI")
    sd.add_def("eqP = D^i D_i P + A_ij * A^ij / P^7 / 8")
    sd.add_def("eqNP = D^i D_i NP - 7. / 8. * NP / P^8 * A_ij * A^ij")
    sd.add_def("eqbet^i = D_j D^j bet^i + D^i D_j bet^j / 3. - 2. * A^ij * D_j Ntilde")
    sd.space_add_eq("eqNP = 0", "N", "dn(N)")
    sd.space_add_eq("eqP = 0", "P", "dn(P)")
    sd.space_add_eq("eqbet^i = 0", "bet^i", "dn(bet^i)")
    sd.space_add_outer_bc("NP = 1")
    sd.space_add_outer_bc("P = 1")
    sd.space_add_outer_bc("bet^i = xboost * ex^i + yboost * ey^i")
    sd.space_add_inner_bc("dn(NP) = 0")
    sd.space_add_inner_bc("sm^j * D_j P + P / 4 * D^j sm_j + A_ij * sm^i * sm^j / P^3 / 4 = 0")
    sd.space_add_inner_bc("bet^i = N / P^2 * sm^i + ome * mg^i")
    sd.add_eq_int(2, INNER_BC, "integ(intS) - chi * CM * CM = 0 ")
    sd.add_eq_int(2, INNER_BC, "integ(intMsq) - M * M = 0 ")
    sd.add_global("M_adm", ndom - 1, OUTER_BC, "integ(intMadm)")
    sd.add_global("M_komar", ndom - 1, OUTER_BC, "integ(intMk)")
    sd.add_global("Mirr2", 2, INNER_BC, "integ(intMsq)")
    sd.add_global("S", 2, INNER_BC, "integ(intS)")
    return sd


def load_fuka_bh_fields(system, fields, dmin=2):
    """Set P, N and bet^i of a fuka_bh_total_bc system from the fields read with the
    space (conf, lapse, shift: configuration-space values)."""
    conf, lapse, shift = fields
    for d in range(dmin, len(conf)):
        system.set_field("P", d, conf[d].conf_values())
        system.set_field("N", d, lapse[d].conf_values())
        for c, name in enumerate(("bet^x", "bet^y", "bet^z")):
            system.set_field(name, d, shift["components"][c][d].conf_values())
