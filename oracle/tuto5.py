"""ORACLE (test infrastructure only): tutorials/tuto5.cpp restated with the oracle's
domains -- the Laplace equation Lap(F) = 0 in domains 1..ndom-1 of a Space_spheric,
the Robin condition dn(F) + 0.5 F / a = 0 at the inner boundary, F = 1 at infinity,
and the matching of F and dn(F) across every shell boundary (tuto5.cpp:60-90), with
the equation order of those calls.  The exact solution is F = 1 + a / r.

Used only by tests/.  The residual rows follow Equation::export_val
(eq_inside.cpp:34-46, eq_bc.cpp:29-41, eq_matching.cpp:34-64); the problem is linear,
so its Jacobian column c is F(u + e_c) - F(u) for the unit unknown e_c placed by
affecte_tau (System_of_eqs::xx_to_vars, system_of_eqs.cpp:597-612).
"""
import numpy as np
import scipy.linalg

from .kadath import INNER_BC, OUTER_BC, VD, vadd, vsub, vmul_scal, vdiv_scal, vadd_scal
from .xcts import Space


class Tuto5:
    def __init__(self, res=(13, 5, 4), bounds=(1.0, 2.0, 4.0)):
        self.sp = Space(res, bounds)
        self.ndom = len(bounds) + 1
        self.a = float(bounds[0])
        self.F = {d: self.sp.doms[d].conf_field(1.0) for d in range(1, self.ndom)}

    def unknown_layout(self):
        lay, off = [], 0
        for d in range(1, self.ndom):
            dom = self.sp.doms[d]
            n = dom.nbr_unknowns(dom.std_base().bt)
            lay.append((d, off, n))
            off += n
        return lay, off

    def lap(self, dom, v):
        g = dom.der_abs(v)
        terms = [dom.der_abs(g[i])[i] for i in range(3)]
        return vadd(vadd(terms[0], terms[1]), terms[2])

    def residual(self, F=None):
        F = self.F if F is None else F
        doms = self.sp.doms
        rows = []
        for d in range(1, self.ndom):
            rows.append(doms[d].export_tau(self.lap(doms[d], F[d]), 2))
        # inner Robin condition in domain 1
        d1 = doms[1]
        bc = vadd(d1.der_normal(F[1], INNER_BC), vdiv_scal(vmul_scal(F[1], 0.5), self.a))
        rows.append(d1.export_tau_boundary(bc, INNER_BC))
        dl = doms[self.ndom - 1]
        rows.append(dl.export_tau_boundary(vadd_scal(F[self.ndom - 1], -1.0), OUTER_BC))
        for d in range(1, self.ndom - 1):
            a, b = doms[d], doms[d + 1]
            rows.append(a.export_tau_boundary(F[d], OUTER_BC) - b.export_tau_boundary(F[d + 1], INNER_BC))
            rows.append(a.export_tau_boundary(a.der_normal(F[d], OUTER_BC), OUTER_BC)
                        - b.export_tau_boundary(b.der_normal(F[d + 1], INNER_BC), INNER_BC))
        return np.concatenate(rows)

    def shifted(self, X, scale=1.0):
        lay, _ = self.unknown_layout()
        out = {}
        for (d, off, cnt) in lay:
            dom = self.sp.doms[d]
            cf, used = dom.affecte_tau(dom.std_base().bt, X[off:off + cnt])
            assert used == cnt
            out[d] = vadd(self.F[d], VD(dom, cf=scale * cf, base=dom.std_base()))
        return out

    def jacobian(self):
        _, n = self.unknown_layout()
        F0 = self.residual()
        J = np.zeros((F0.size, n))
        for c in range(n):
            e = np.zeros(n)
            e[c] = 1.0
            J[:, c] = self.residual(self.shifted(e)) - F0
        return J

    def newton_step(self):
        F0 = self.residual()
        J = self.jacobian()
        X = scipy.linalg.lu_solve(scipy.linalg.lu_factor(J), F0)
        self.F = self.shifted(X, -1.0)
        return np.abs(F0).max()
