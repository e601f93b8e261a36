"""ORACLE (test infrastructure only): the XCTS black-hole system, residual
F(u), forward-AD Jacobian and Newton step.

Used only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.

System: the Kerr prototype of codes/Par_version/Kerr/kerr.cpp:69-99 (the one the
reference solves by do_newton from a flat guess), written in the Cartesian
tensor basis of the FUKA solvers (bh_3d_xcts_solver_imp.cpp:119-159), on
Space_spheric (src/Space/Spheric/):
  unknowns  P (conformal factor), N (lapse), bet^i (shift)
  A^ij = (D^i bet^j + D^j bet^i - 2/3 D_k bet^k f^ij) / 2 / N
  inner (excision r = a):  N = n0 ; bet^i = n0 / P^2 s^i + Ome m^i ;
                           dn(P) + 0.5 P / a + P^3 A_ij s^i s^j / 4 = 0
  eqN  : D_i D^i N + 2 D_i P D^i N / P - N P^4 A_ij A^ij = 0   (matching N, dn(N))
  eqP  : D_i D^i P + P^5 A_ij A^ij / 8 = 0                    (matching P, dn(P))
  eqbet: D_j A^ij + 6 A^ij D_j P / P = 0                       (matching bet^i, dn(bet^i))
  outer (infinity): N = 1, P = 1, bet^i = 0
with s^i the radial unit vector and m^i = (-y, x, 0) the rotation field (the
spinning variant: kerr.cpp:195, "bet^i = n0 / P^2 * st^i + Ome * m^i * a").

Equation order, unknown order and row export follow System_of_eqs
(src/System_of_eqs/solver.cpp:49-85, 120-214; system_of_eqs.cpp:597-612) and
Space_spheric::add_eq / add_inner_bc / add_outer_bc
(src/Space/Spheric/spheric_add_eq.cpp:26-52).
"""
import numpy as np
import scipy.linalg

from .kadath import (VD, Domain, INNER_BC, OUTER_BC, vadd, vsub, vmul, vdiv, vneg, vmul_scal, vdiv_scal,
                     vadd_scal)


# ---------------------------------------------------------------------------
# Term_eq: (value, tangent) pairs over one domain (src/Term_eq/term_eq_math.cpp)
# ---------------------------------------------------------------------------
class TE:
    __slots__ = ("val", "der")

    def __init__(self, val, der):
        self.val, self.der = val, der


def t_add(a, b):
    return TE(vadd(a.val, b.val), vadd(a.der, b.der))


def t_sub(a, b):
    return TE(vsub(a.val, b.val), vsub(a.der, b.der))


def t_mul(a, b):
    """term_eq_math.cpp:248-342: der = a' b + a b'."""
    return TE(vmul(a.val, b.val), vadd(vmul(a.der, b.val), vmul(a.val, b.der)))


def t_div(a, b):
    """term_eq_math.cpp:344-441: der = (a' b - a b') / (b b)."""
    return TE(vdiv(a.val, b.val), vdiv(vsub(vmul(a.der, b.val), vmul(a.val, b.der)), vmul(b.val, b.val)))


def t_scal(a, x):
    return TE(vmul_scal(a.val, x), vmul_scal(a.der, x))


def t_div_scal(a, x):
    return TE(vdiv_scal(a.val, x), vdiv_scal(a.der, x))


def t_pow(a, n, dom):
    """pow(Term_eq, int) (term_eq_math.cpp:547-567): products for n > 0, 1/a/a/... for n < 0."""
    if n > 0:
        res = a
        for _ in range(1, n):
            res = t_mul(res, a)
        return res
    one = TE(dom.conf_field(1.0), VD(dom, c=np.zeros((dom.np, dom.nt, dom.nr)), base=dom.std_base()))
    res = one
    for _ in range(-n):
        res = t_div(res, a)
    return res


def t_grad(dom, a):
    """Metric_flat::derive_partial_cart -> Domain::partial_cart (metric_flat.cpp:150-195,
    domain.cpp:408-451): Cartesian components of the flat gradient."""
    gv = dom.der_abs(a.val)
    gd = dom.der_abs(a.der)
    return [TE(gv[i], gd[i]) for i in range(3)]


def t_dn(dom, a, bound):
    return TE(dom.der_normal(a.val, bound), dom.der_normal(a.der, bound))


def t_sum(terms):
    res = terms[0]
    for t in terms[1:]:
        res = t_add(res, t)
    return res


# ---------------------------------------------------------------------------
# Space + system
# ---------------------------------------------------------------------------
class Space:
    """Space_spheric(CHEB_TYPE, 0, res, bounds) (src/Space/Spheric/space_spheric.cpp:25-47):
    nucleus [0, b0], shells [b_{i-1}, b_i], compactified [b_last, inf).  The
    nucleus is never solved here (it is excised), so only shells/compact are built."""

    def __init__(self, res, bounds):
        self.bounds = list(bounds)
        self.ndom = len(bounds) + 1
        self.doms = [None]
        for i in range(1, self.ndom - 1):
            self.doms.append(Domain(i, "shell", res, bounds[i - 1], bounds[i]))
        self.doms.append(Domain(self.ndom - 1, "compact", res, bounds[-1]))


def coord_fields(dom):
    """Constant vector fields of the FUKA BH setup, Cartesian components with the
    standard x/y, anti-symmetric z bases (Tensor::std_base, src/Tensor/tensor.cpp:609-640):
    sm = radial unit vector, mg = global rotation (-y, x, 0), einf = radial unit vector."""
    st, ct = np.sin(dom.TH), np.cos(dom.TH)
    cp, sp = np.cos(dom.PH), np.sin(dom.PH)
    sm = [dom.conf_field(st * cp), dom.conf_field(st * sp), dom.conf_field(ct, anti=True)]
    with np.errstate(invalid="ignore"):
        R = dom.R
        mg = [dom.conf_field(-R * st * sp), dom.conf_field(R * st * cp), VD.zeros(dom)]
    return sm, mg


class XCTSSystem:
    FIELDS = ("P", "N", "bx", "by", "bz")

    def __init__(self, space, dmin, dmax, n0=0.5, a=1.0, ome=0.0, spin=None, cm=1.0):
        """n0: lapse on the horizon; a: excision radius; ome: horizon rotation Ome
        (constant), or spin=chi to make ome an unknown fixed by the spin integral
        (add_var("ome") + add_eq_int_bh, bh_3d_xcts_stages.cpp:84-85)."""
        self.sp = space
        self.n0, self.a = float(n0), float(a)
        self.dmin, self.dmax = dmin, dmax
        self.ome = float(ome)
        self.spin = spin
        self.cm = cm
        self.fields = {}
        self.consts = {d: coord_fields(space.doms[d]) for d in range(dmin, dmax + 1)}
        self.ncond = {}

    # ---- fields
    def set_fields(self, fields):
        """fields[name][d] = VD (P, N, bx, by, bz; bz has the anti-symmetric base)."""
        self.fields = fields

    def comp_base_t(self, name, d):
        dom = self.sp.doms[d]
        return (dom.anti_base() if name == "bz" else dom.std_base()).bt

    # ---- unknowns (System_of_eqs::add_var order: var_double, then var x domain x component)
    def unknown_layout(self):
        """Variables in add_var order (P, N, bet); each variable domain by domain,
        a vector's Cartesian components together inside each domain
        (System_of_eqs::xx_to_vars, system_of_eqs.cpp:597-612; Domain_shell::affecte_tau
        on a CARTESIAN vector, domain_shell_affecte_tau.cpp:140-150)."""
        lay = []
        off = 1 if self.spin is not None else 0
        for var in (("P",), ("N",), ("bx", "by", "bz")):
            for d in range(self.dmin, self.dmax + 1):
                for name in var:
                    dom = self.sp.doms[d]
                    n = dom.nbr_unknowns(self.comp_base_t(name, d))
                    lay.append((name, d, off, n))
                    off += n
        return lay, off

    def xx_to_fields(self, X):
        """xx_to_vars (system_of_eqs.cpp:597-612) + affecte_tau: coefficient corrections."""
        lay, n = self.unknown_layout()
        out = {name: {} for name in self.FIELDS}
        for (name, d, off, cnt) in lay:
            dom = self.sp.doms[d]
            cf, used = dom.affecte_tau(self.comp_base_t(name, d), X[off:off + cnt])
            assert used == cnt
            base = dom.anti_base() if name == "bz" else dom.std_base()
            out[name][d] = VD(dom, cf=cf, base=base)
        dome = X[0] if self.spin is not None else 0.0
        return out, dome

    # ---- evaluation of every equation touching domain d
    def _terms(self, d, ders):
        dom = self.sp.doms[d]
        f = {}
        for name in self.FIELDS:
            v = self.fields[name][d]
            f[name] = TE(v, ders.get((name, d), VD.zeros(dom)))
        ome = TE(self._ome_vd(dom, self.ome), ders.get(("ome", d), VD.zeros(dom)))
        return dom, f, ome

    def _ome_vd(self, dom, x):
        return dom.conf_field(x)

    def defs(self, d, ders):
        dom, f, ome = self._terms(d, ders)
        P, N = f["P"], f["N"]
        bet = [f["bx"], f["by"], f["bz"]]
        Dbet = [t_grad(dom, b) for b in bet]          # Dbet[j][i] = d_i bet^j
        divb = t_sum([Dbet[k][k] for k in range(3)])
        A = [[None] * 3 for _ in range(3)]
        for i in range(3):
            for j in range(3):
                s = t_add(Dbet[j][i], Dbet[i][j])
                if i == j:
                    s = t_sub(s, t_scal(divb, 2.0 / 3.0))
                A[i][j] = t_div(t_div_scal(s, 2.0), N)
        AA = t_sum([t_mul(A[i][j], A[i][j]) for i in range(3) for j in range(3)])
        return dict(dom=dom, P=P, N=N, bet=bet, A=A, AA=AA, ome=ome)

    def lap(self, dom, a):
        g = t_grad(dom, a)
        return t_sum([t_grad(dom, g[i])[i] for i in range(3)])

    def eq_inside(self, D, which):
        dom = D["dom"]
        P, N = D["P"], D["N"]
        if which == "N":
            gP, gN = t_grad(dom, P), t_grad(dom, N)
            t2 = t_div(t_scal(t_sum([t_mul(gP[i], gN[i]) for i in range(3)]), 2.0), P)
            t3 = t_mul(t_mul(N, t_pow(P, 4, dom)), D["AA"])
            return [t_sub(t_add(self.lap(dom, N), t2), t3)]
        if which == "P":
            return [t_add(self.lap(dom, P), t_div_scal(t_mul(t_pow(P, 5, dom), D["AA"]), 8.0))]
        gP = t_grad(dom, P)
        gA = [[t_grad(dom, D["A"][i][j]) for j in range(3)] for i in range(3)]
        out = []
        for i in range(3):
            divA = t_sum([gA[i][j][j] for j in range(3)])
            t2 = t_div(t_scal(t_sum([t_mul(D["A"][i][j], gP[j]) for j in range(3)]), 6.0), P)
            out.append(t_add(divA, t2))
        return out

    def match_expr(self, D, which, bound):
        dom = D["dom"]
        if which == "N":
            return [D["N"]]
        if which == "dnN":
            return [t_dn(dom, D["N"], bound)]
        if which == "P":
            return [D["P"]]
        if which == "dnP":
            return [t_dn(dom, D["P"], bound)]
        if which == "bet":
            return D["bet"]
        if which == "dnbet":
            return [t_dn(dom, b, bound) for b in D["bet"]]
        raise ValueError(which)

    def bc_outer(self, D, which):
        if which == "N":
            return [TE(vadd_scal(D["N"].val, -1.0), D["N"].der)]
        if which == "P":
            return [TE(vadd_scal(D["P"].val, -1.0), D["P"].der)]
        return D["bet"]

    def bc_inner(self, D, which, d):
        dom = D["dom"]
        sm, mg = self.consts[d]
        zero = lambda: VD.zeros(dom)  # noqa: E731
        smT = [TE(s, zero()) for s in sm]
        mgT = [TE(m, zero()) for m in mg]
        if which == "N":
            return [TE(vadd_scal(D["N"].val, -self.n0), D["N"].der)]
        if which == "bet":
            nP2 = t_div(TE(dom.conf_field(self.n0), zero()), t_pow(D["P"], 2, dom))
            out = []
            for i in range(3):
                rhs = t_mul(nP2, smT[i])
                rhs = t_add(rhs, t_scal(t_mul(D["ome"], mgT[i]), self.a))
                out.append(t_sub(D["bet"][i], rhs))
            return out
        # dn(P) + 0.5 * P / a + P^3 * A_ij * s^i * s^j / 4 = 0
        t1 = t_dn(dom, D["P"], INNER_BC)
        t2 = t_div_scal(t_scal(D["P"], 0.5), self.a)
        t3 = t_sum([t_mul(t_mul(D["A"][i][j], smT[i]), smT[j]) for i in range(3) for j in range(3)])
        t3 = t_div_scal(t_mul(t_pow(D["P"], 3, dom), t3), 4.0)
        return [t_add(t_add(t1, t2), t3)]

    def eq_list(self):
        """Equations in the order kerr.cpp:84-99 adds them."""
        L = [("inner", self.dmin, "N"), ("inner", self.dmin, "bet"), ("inner", self.dmin, "AH")]
        for which, rac, racd in (("N", "N", "dnN"), ("P", "P", "dnP"), ("bet", "bet", "dnbet")):
            for d in range(self.dmin, self.dmax):
                L.append(("inside", d, which))
                L.append(("match", d, rac))
                L.append(("match", d, racd))
            L.append(("inside", self.dmax, which))
        for which in ("N", "P", "bet"):
            L.append(("outer", self.dmax, which))
        return L

    def _eval(self, eq, cache, ders_for):
        kind, d, which = eq
        if kind == "match":
            Da = cache(d)
            Db = cache(d + 1)
            return (self.match_expr(Da, which, OUTER_BC), self.match_expr(Db, which, INNER_BC))
        D = cache(d)
        if kind == "inside":
            return (self.eq_inside(D, which),)
        if kind == "outer":
            return (self.bc_outer(D, which),)
        return (self.bc_inner(D, which, d),)

    def _export(self, eq, parts, use_der, nrows_hint=None):
        """Equation::export_val / export_der (eq_inside.cpp:34-46, eq_bc.cpp:29-41, eq_matching.cpp:34-64)."""
        kind, d, which = eq
        dom = self.sp.doms[d]
        pick = (lambda t: t.der) if use_der else (lambda t: t.val)
        rows = []
        for ci, t in enumerate(parts[0]):
            key = (eq, ci)
            vd = pick(t)
            if kind == "inside":
                r = dom.export_tau(vd, 2)
            else:
                bound = INNER_BC if kind == "inner" else OUTER_BC
                r = dom.export_tau_boundary(vd, bound)
            if key not in self.ncond:
                assert r is not None
                self.ncond[key] = r.shape[-1]
            nc = self.ncond[key]
            if kind == "match":
                other = self.sp.doms[d + 1]
                r2 = other.export_tau_boundary(pick(parts[1][ci]), INNER_BC)
                if r is None and r2 is None:
                    r = None
                else:
                    base_shape = (r if r is not None else r2).shape
                    r = (r if r is not None else 0.0) - (r2 if r2 is not None else 0.0)
                    r = np.broadcast_to(r, base_shape)
            rows.append((nc, r))
        return rows

    def integral_value(self, D, d, use_der):
        dom = D["dom"]
        sm, mg = self.consts[d]
        zero = lambda: VD.zeros(dom)  # noqa: E731
        smT = [TE(s, zero()) for s in sm]
        mgT = [TE(m, zero()) for m in mg]
        t = t_sum([t_mul(t_mul(D["A"][i][j], mgT[i]), smT[j]) for i in range(3) for j in range(3)])
        t = t_div_scal(t_div_scal(t, 8.0), np.pi)
        vd = t.der if use_der else t.val
        val = dom.integ(vd, INNER_BC)
        if not use_der:
            val = val - self.spin * self.cm * self.cm
        return val

    # ---- global quantities (codes/Par_version/Kerr/reader.cpp:114-125, Cartesian basis)
    def observables(self):
        """M_ADM = -integ(d_r P)/2pi, M_Komar = integ(d_r N)/4pi, J_ADM = -integ(A^ij s_i m_j)/8pi
        at the outer boundary of the compactified domain, integ as Domain_compact::integ
        (domain_compact.cpp:89-121); d_r of the compactified domain is der_normal there
        (domain_compact_ope.cpp:187-189, domain_compact.cpp:68-87)."""
        d = self.dmax
        D = self.defs(d, {})
        dom = D["dom"]
        sm, _ = self.consts[d]
        # m = r mh with mh = sin(theta) (-sin(phi), cos(phi), 0) bounded at infinity; the
        # factor r is applied in coefficient space, as mult_r(a13.mult_sin_theta()) does
        st = np.sin(dom.TH)
        mh = [dom.conf_field(-st * np.sin(dom.PH)), dom.conf_field(st * np.cos(dom.PH))]
        m_adm = -dom.integ(dom.der_normal(D["P"].val, OUTER_BC), OUTER_BC) / (2 * np.pi)
        m_komar = dom.integ(dom.der_normal(D["N"].val, OUTER_BC), OUTER_BC) / (4 * np.pi)
        so = None
        for i in range(3):
            for j in range(2):
                t = vmul(vmul(D["A"][i][j].val, sm[i]), mh[j])
                so = t if so is None else vadd(so, t)
        j_adm = -dom.integ(dom.mult_r(so), OUTER_BC) / (8 * np.pi)
        return dict(M_adm=float(m_adm), M_komar=float(m_komar), J_adm=float(j_adm))

    # ---- residual (System_of_eqs::sec_member, solver.cpp:49-85)
    def residual(self):
        cache_d = {}

        def cache(d):
            if d not in cache_d:
                cache_d[d] = self.defs(d, {})
            return cache_d[d]

        out = []
        if self.spin is not None:
            out.append(np.atleast_1d(self.integral_value(cache(self.dmin), self.dmin, False)))
        for eq in self.eq_list():
            parts = self._eval(eq, cache, None)
            for nc, r in self._export(eq, parts, False):
                out.append(np.zeros(nc) if r is None else r)
        return np.concatenate(out)

    def check_equations(self):
        """Per-equation max |F| (solver.cpp:24-47)."""
        F = self.residual()
        return np.abs(F).max() if F.size else 0.0

    # ---- Jacobian by batched forward AD (do_col_J, solver.cpp:120-214)
    def jacobian(self, cols=None, chunk=512):
        lay, n = self.unknown_layout()
        F0 = self.residual()
        m = F0.size
        J = np.zeros((m, n)) if cols is None else np.zeros((m, len(cols)))
        want = np.arange(n) if cols is None else np.asarray(cols)
        colpos = {int(c): q for q, c in enumerate(want)}
        eqs = self.eq_list()
        # row offsets of each equation
        offs = []
        pos = 1 if self.spin is not None else 0
        for eq in eqs:
            ncs = [self.ncond[(eq, ci)] for ci in range(self._ncomp(eq))]
            offs.append((pos, ncs))
            pos += sum(ncs)
        assert pos == m
        if self.spin is not None and 0 in colpos:
            J[:, colpos[0]] = self._ome_column(eqs, offs, m)
        for (name, d, off, cnt) in lay:
            sel = [u for u in range(cnt) if (off + u) in colpos]
            if not sel:
                continue
            dom = self.sp.doms[d]
            seeds_all = dom.seed_coefs(self.comp_base_t(name, d))
            base = dom.anti_base() if name == "bz" else dom.std_base()
            for s0 in range(0, len(sel), chunk):
                idx = sel[s0:s0 + chunk]
                seeds = VD(dom, cf=seeds_all[idx], base=base)
                block = self._jvp_block(eqs, offs, m, {(name, d): seeds}, d)
                for q, u in enumerate(idx):
                    J[:, colpos[off + u]] = block[:, q]
        return J

    def _ncomp(self, eq):
        return 3 if eq[2] in ("bet", "dnbet") else 1


    def _jvp_block(self, eqs, offs, m, ders, dseed):
        B = next(iter(ders.values())).cf.shape[0]
        out = np.zeros((m, B))
        cache_d = {}

        def cache(d):
            if d not in cache_d:
                cache_d[d] = self.defs(d, ders)
            return cache_d[d]

        if self.spin is not None:
            if dseed == self.dmin:
                out[0] = self.integral_value(cache(self.dmin), self.dmin, True)
        for eq, (pos, ncs) in zip(eqs, offs):
            kind, d, which = eq
            touches = (d == dseed) or (kind == "match" and d + 1 == dseed)
            if not touches:
                continue
            parts = self._eval(eq, cache, ders)
            p = pos
            for nc, r in self._export(eq, parts, True):
                if r is not None:
                    out[p:p + nc] = np.asarray(r).T
                p += nc
        return out

    def _ome_column(self, eqs, offs, m):
        """Column of the var_double ome: every equation is applied (solver.cpp:179-186)."""
        col = np.zeros(m)
        cache_d = {}

        def ders_for(d):
            dom = self.sp.doms[d]
            return {("ome", d): VD(dom, c=np.ones((1, dom.np, dom.nt, dom.nr)), base=dom.std_base())}

        def cache(d):
            if d not in cache_d:
                cache_d[d] = self.defs(d, ders_for(d))
            return cache_d[d]

        col[0] = np.atleast_1d(self.integral_value(cache(self.dmin), self.dmin, True))[0]
        for eq, (pos, ncs) in zip(eqs, offs):
            parts = self._eval(eq, cache, None)
            p = pos
            for nc, r in self._export(eq, parts, True):
                if r is not None:
                    col[p:p + nc] = np.asarray(r).reshape(-1, nc)[0]
                p += nc
        return col


# ---------------------------------------------------------------------------
# Initial data and Newton
# ---------------------------------------------------------------------------
def initial_perturbation(sys, seed=20261202, amp=1e-3):
    """amp * 2^-(i+j+k) U(-1,1) on every admissible coefficient, in unknown order
    (numpy.random.default_rng(seed))."""
    rng = np.random.default_rng(seed)
    lay, n = sys.unknown_layout()
    X = np.zeros(n)
    for (name, d, off, cnt) in lay:
        dom = sys.sp.doms[d]
        u = off
        for (k, j) in dom.unknown_slots(sys.comp_base_t(name, d)):
            for i in range(dom.nr):
                X[u] = amp * 2.0 ** (-(i + j + k)) * rng.uniform(-1, 1)
                u += 1
    return X


def initial_fields(sys, rexc, X0=None):
    """Flat guess P = 1 + rexc/r, N = 1, bet = 0 (codes/Par_version/Kerr/kerr.cpp:36-52),
    plus affecte_tau(X0) (X0 in unknown order, e.g. initial_perturbation)."""
    space, dmin, dmax = sys.sp, sys.dmin, sys.dmax
    fields = {n: {} for n in XCTSSystem.FIELDS}
    for d in range(dmin, dmax + 1):
        dom = space.doms[d]
        with np.errstate(divide="ignore"):
            Pc = 1.0 + rexc / dom.R
        Pc = np.where(np.isfinite(Pc), Pc, 1.0)
        fields["P"][d] = dom.conf_field(Pc)
        fields["N"][d] = dom.conf_field(1.0)
        fields["bx"][d] = dom.conf_field(0.0)
        fields["by"][d] = dom.conf_field(0.0)
        fields["bz"][d] = dom.conf_field(0.0, anti=True)
    if X0 is not None:
        corr, _ = sys.xx_to_fields(X0)
        for name in XCTSSystem.FIELDS:
            for d in range(dmin, dmax + 1):
                fields[name][d] = vadd(fields[name][d], corr[name][d])
    return fields


def analytic_fields(sy, m, J):
    """P = 1 + m/(2r), N = 1 - m/r, bet = (2J/r^3)(-y, x, 0): polynomials in 1/r in the
    compactified domain, so the surface integrals at infinity are exact: M_ADM = m,
    M_Komar = m, and, for m = 0 (A^ij = L bet / 2N with N = 1 stays polynomial),
    J_ADM = J."""
    fields = {n: {} for n in XCTSSystem.FIELDS}
    for d in range(sy.dmin, sy.dmax + 1):
        dom = sy.sp.doms[d]
        with np.errstate(divide="ignore"):
            u = np.where(np.isfinite(dom.R), 1.0 / dom.R, 0.0)
        st = np.sin(dom.TH)
        fields["P"][d] = dom.conf_field(1.0 + 0.5 * m * u)
        fields["N"][d] = dom.conf_field(1.0 - m * u)
        fields["bx"][d] = dom.conf_field(-2.0 * J * u * u * st * np.sin(dom.PH))
        fields["by"][d] = dom.conf_field(2.0 * J * u * u * st * np.cos(dom.PH))
        fields["bz"][d] = dom.conf_field(0.0, anti=True)
    return fields


def newton_step(sys, solver="lapack"):
    """One System_of_eqs::do_newton iteration with the single-process LU of
    do_newton_seq (src_par/do_newton.cpp:519-572; Matrice::set_lu/solve)."""
    F = sys.residual()
    err = np.abs(F).max()
    J = sys.jacobian()
    lu = scipy.linalg.lu_factor(J)
    X = scipy.linalg.lu_solve(lu, F)
    corr, dome = sys.xx_to_fields(X)
    new = {}
    for name in XCTSSystem.FIELDS:
        new[name] = {d: vsub(sys.fields[name][d], corr[name][d]) for d in sys.fields[name]}
    sys.set_fields(new)
    if sys.spin is not None:
        sys.ome -= dome
    return err, F, J, X
