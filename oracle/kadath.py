"""ORACLE (test infrastructure only): numpy restatement of Kadath's spherical
domains, Val_domain algebra, Term_eq forward AD and tau bookkeeping.

Used only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.

Arrays are (..., n_phi, n_theta, n_r) with any leading batch dims (a batch of
tangents for the forward-mode Jacobian); r is the fastest index as in the
reference's Array (include/implementation/index.cpp:84-96).  Bases:
Base.br[k, j] (r lines), Base.bt[k] (theta lines), Base.bp (phi) as
Base_spectral::bases_1d (src/Base_spectral/base_spectral.cpp:92-104).
"""
import numpy as np

from . import spectral as S
from .spectral import CHEB, COSSIN, COS_EVEN, COS_ODD, SIN_EVEN, SIN_ODD

INNER_BC, OUTER_BC = 0, 1


# ---------------------------------------------------------------------------
# 1-D coefficient-space operators, literal restatements of src/Ope_1d/*.cpp.
# Each returns (new_line, output_basis).
# ---------------------------------------------------------------------------
def der_1d(base, tab):
    """src/Ope_1d/der_1d.cpp:31-48, 132-148, 215-279."""
    nr = len(tab)
    res = np.zeros(nr)
    if base == CHEB:
        for i in range(nr):
            for p in range(i + 1, nr, 2):
                res[i] += p * tab[p]
        res *= 2
        res[0] /= 2
        return res, CHEB
    if base == COSSIN:
        for i in range(0, nr - 1, 2):
            res[i + 1] = -(i // 2) * tab[i]
        for i in range(1, nr, 2):
            res[i - 1] = ((i - 1) // 2) * tab[i]
        res[nr - 1] = 0
        return res, COSSIN
    if base == COS_EVEN:
        res = -2 * np.arange(nr) * tab
        res[nr - 1] = 0
        return res, SIN_EVEN
    if base == COS_ODD:
        return -(2 * np.arange(nr) + 1) * tab, SIN_ODD
    if base == SIN_EVEN:
        return 2 * np.arange(nr) * tab, COS_EVEN
    if base == SIN_ODD:
        res = (2 * np.arange(nr) + 1) * tab
        res[nr - 1] = 0
        return res, COS_ODD
    raise NotImplementedError(base)


def mult_sin_1d(base, tab):
    """src/Ope_1d/mult_sin_1d.cpp:29-146."""
    nr = len(tab)
    res = np.zeros(nr)
    if base == COSSIN:
        res[0] = 0.5 * tab[3]
        res[2] = 0.5 * tab[5]
        res[3] = tab[0] - 0.5 * tab[4]
        for k in range(4, nr - 2, 2):
            res[k] = 0.5 * (tab[3 + k] - tab[k - 1])
            res[k + 1] = 0.5 * (tab[k - 2] - tab[k + 2])
        res[nr - 2] = -0.5 * tab[nr - 3]
        return res, COSSIN
    if base == COS_EVEN:
        res[0] = tab[0] - 0.5 * tab[1]
        for k in range(1, nr - 1):
            res[k] = 0.5 * (tab[k] - tab[k + 1])
        return res, SIN_ODD
    if base == COS_ODD:
        for k in range(1, nr):
            res[k] = 0.5 * (tab[k - 1] - tab[k])
        return res, SIN_EVEN
    if base == SIN_EVEN:
        res[0] = 0.5 * tab[1]
        for k in range(1, nr - 1):
            res[k] = 0.5 * (-tab[k] + tab[k + 1])
        return res, COS_ODD
    if base == SIN_ODD:
        res[0] = 0.5 * tab[0]
        for k in range(1, nr):
            res[k] = 0.5 * (-tab[k - 1] + tab[k])
        return res, COS_EVEN
    raise NotImplementedError(base)


def mult_cos_1d(base, tab):
    """src/Ope_1d/mult_cos_1d.cpp:29-142."""
    nr = len(tab)
    res = np.zeros(nr)
    if base == COSSIN:
        res[0] = 0.5 * tab[2]
        res[2] = tab[0] + 0.5 * tab[4]
        res[3] = 0.5 * tab[5]
        for k in range(4, nr - 2):
            res[k] = 0.5 * (tab[2 + k] + tab[k - 2])
        res[nr - 2] = 0.5 * tab[nr - 4]
        return res, COSSIN
    if base == COS_EVEN:
        res[0] = tab[0] + 0.5 * tab[1]
        for k in range(1, nr - 1):
            res[k] = 0.5 * (tab[k] + tab[k + 1])
        return res, COS_ODD
    if base == COS_ODD:
        res[0] = 0.5 * tab[0]
        for k in range(1, nr):
            res[k] = 0.5 * (tab[k - 1] + tab[k])
        return res, COS_EVEN
    if base == SIN_EVEN:
        res[0] = 0.5 * tab[1]
        for k in range(1, nr - 1):
            res[k] = 0.5 * (tab[k] + tab[k + 1])
        return res, SIN_ODD
    if base == SIN_ODD:
        for k in range(1, nr):
            res[k] = 0.5 * (tab[k - 1] + tab[k])
        return res, SIN_EVEN
    raise NotImplementedError(base)


def div_sin_1d(base, tab):
    """src/Ope_1d/div_sin_1d.cpp:75-134."""
    nr = len(tab)
    res = np.zeros(nr)
    somme = 0.0
    if base == COS_EVEN:
        for k in range(nr - 2, -1, -1):
            somme -= 2 * tab[k + 1]
            res[k] = somme
        return res, SIN_ODD
    if base == COS_ODD:
        for k in range(nr - 2, -1, -1):
            somme -= 2 * tab[k]
            res[k] = somme
        return res, SIN_EVEN
    if base == SIN_EVEN:
        for k in range(nr - 2, -1, -1):
            somme += 2 * tab[k + 1]
            res[k] = somme
        return res, COS_ODD
    if base == SIN_ODD:
        for k in range(nr - 2, -1, -1):
            somme += 2 * tab[k]
            res[k] = somme
        res[0] *= 0.5
        return res, COS_EVEN
    raise NotImplementedError(base)


def mult_x_1d(base, so):
    """src/Ope_1d/mult_x_1d.cpp:29-41 (CHEB)."""
    assert base == CHEB
    nr = len(so)
    res = np.zeros(nr)
    res[0] = 0.5 * so[1]
    res[1] = so[0] + 0.5 * so[2]
    for i in range(2, nr - 1):
        res[i] = 0.5 * (so[i + 1] + so[i - 1])
    res[nr - 1] = 0.5 * so[nr - 2]
    return res, CHEB


def mult_xm1_1d(base, so):
    """src/Ope_1d/mult_xm1_1d.cpp:29-41 (CHEB)."""
    assert base == CHEB
    nr = len(so)
    res = np.zeros(nr)
    res[0] = -so[0] + 0.5 * so[1]
    res[1] = so[0] - so[1] + 0.5 * so[2]
    for i in range(2, nr - 1):
        res[i] = 0.5 * so[i - 1] - so[i] + 0.5 * so[i + 1]
    res[nr - 1] = 0.5 * so[nr - 2] - so[nr - 1]
    return res, CHEB


def div_xm1_1d(base, so):
    """src/Ope_1d/div_xm1_1d.cpp:29-40 (CHEB)."""
    assert base == CHEB
    nr = len(so)
    res = np.zeros(nr)
    for i in range(nr - 1):
        for j in range(i + 1, nr):
            res[i] += 2 * (j - i) * so[j]
    res[0] *= 0.5
    return res, CHEB


_cache = {}


def op_mat(fn, base, n):
    """Matrix of a linear 1-D routine (applied to unit vectors) and its output basis."""
    key = (fn.__name__, base, n)
    if key not in _cache:
        M = np.zeros((n, n))
        ob = None
        for c in range(n):
            e = np.zeros(n)
            e[c] = 1.0
            out, ob = fn(base, e)
            M[:, c] = out
        _cache[key] = (M, ob)
    return _cache[key]


def tr_mat(kind, base, nconf):
    """Matrix of the oracle's coef_1d (kind 0) / coef_i_1d (kind 1) restatement."""
    key = ("tr", kind, base, nconf)
    if key not in _cache:
        nc = nconf + 2 if base == COSSIN else nconf
        nin = nconf if kind == 0 else nc
        cols = []
        for c in range(nin):
            e = np.zeros(nin)
            e[c] = 1.0
            cols.append(S.coef_1d(base, e) if kind == 0 else S.coef_i_1d(base, e))
        _cache[key] = np.stack(cols, axis=1)
    return _cache[key]


# ---------------------------------------------------------------------------
# Bases and line application
# ---------------------------------------------------------------------------
class Base:
    __slots__ = ("br", "bt", "bp")

    def __init__(self, br, bt, bp=COSSIN):
        self.br = np.array(br, dtype=int)
        self.bt = np.array(bt, dtype=int)
        self.bp = int(bp)

    def copy(self):
        return Base(self.br.copy(), self.bt.copy(), self.bp)

    def __eq__(self, o):
        return (isinstance(o, Base) and self.bp == o.bp and np.array_equal(self.bt, o.bt)
                and np.array_equal(self.br, o.br))

    def __repr__(self):
        return f"Base(bt={self.bt.tolist()})"


def apply_phi(x, M):
    return np.einsum("oq,...qjr->...ojr", M, x)


def apply_theta(x, mats):
    out = np.empty(x.shape[:-2] + (mats[0].shape[0], x.shape[-1]))
    for k, M in enumerate(mats):
        out[..., k, :, :] = np.einsum("oq,...qr->...or", M, x[..., k, :, :])
    return out


def apply_r(x, br, matfn):
    out = np.empty(x.shape)
    for b in np.unique(br):
        mask = br == b
        out[..., mask, :] = x[..., mask, :] @ matfn(b).T
    return out


# theta-basis product rule, Domain_shell::mult (src/Domain/Spheric/domain_shell.cpp:722-798):
# value = (basis if k%4 < 2, basis otherwise)
_MULT_T = {
    (COS_EVEN, COS_EVEN): (COS_EVEN, SIN_ODD), (COS_EVEN, COS_ODD): (COS_ODD, SIN_EVEN),
    (COS_EVEN, SIN_EVEN): (SIN_EVEN, COS_ODD), (COS_EVEN, SIN_ODD): (SIN_ODD, COS_EVEN),
    (COS_ODD, COS_EVEN): (COS_ODD, SIN_EVEN), (COS_ODD, COS_ODD): (COS_EVEN, SIN_ODD),
    (COS_ODD, SIN_EVEN): (SIN_ODD, COS_EVEN), (COS_ODD, SIN_ODD): (SIN_EVEN, COS_ODD),
    (SIN_EVEN, COS_EVEN): (SIN_EVEN, COS_ODD), (SIN_EVEN, COS_ODD): (SIN_ODD, COS_EVEN),
    (SIN_EVEN, SIN_EVEN): (COS_EVEN, SIN_ODD), (SIN_EVEN, SIN_ODD): (COS_ODD, SIN_EVEN),
    (SIN_ODD, COS_EVEN): (SIN_ODD, COS_EVEN), (SIN_ODD, COS_ODD): (SIN_EVEN, COS_ODD),
    (SIN_ODD, SIN_EVEN): (COS_ODD, SIN_EVEN), (SIN_ODD, SIN_ODD): (COS_EVEN, SIN_ODD),
}
# theta-basis inversion of mult_cos_phi / mult_sin_phi (domain_shell_ope.cpp:36-55)
_FLIP_T = {COS_EVEN: SIN_ODD, COS_ODD: SIN_EVEN, SIN_EVEN: COS_ODD, SIN_ODD: COS_EVEN}


class Domain:
    """Domain_shell (src/Domain/Spheric/domain_shell.cpp:28-35) or Domain_compact
    (src/Domain/Spheric/domain_compact.cpp:28-33), CHEB_TYPE."""

    def __init__(self, num, kind, nbr, rint=None, rext=None):
        self.num = num
        self.kind = kind
        self.nr, self.nt, self.np = nbr
        self.npc = self.np + 2
        if kind == "shell":
            self.alpha = (rext - rint) / 2.0
            self.beta = (rext + rint) / 2.0
        elif kind == "compact":
            self.alpha = -0.5 / rint
            self.beta = None
        else:
            raise ValueError(kind)
        self.xr = -np.cos(np.pi * np.arange(self.nr) / (self.nr - 1))
        self.th = np.pi / 2 * np.arange(self.nt) / (self.nt - 1)
        self.ph = 2 * np.pi * np.arange(self.np) / self.np
        PH, TH, XR = np.meshgrid(self.ph, self.th, self.xr, indexing="ij")
        with np.errstate(divide="ignore"):
            if kind == "shell":
                R = self.alpha * XR + self.beta
            else:
                R = 1.0 / (self.alpha * (XR - 1.0))
        self.R, self.TH, self.PH = R, TH, PH

    # ---- bases
    def std_base(self):
        bt = [COS_EVEN if (k // 2) % 2 == 0 else SIN_ODD for k in range(self.npc)]
        return Base(np.full((self.npc, self.nt), CHEB), bt, COSSIN)

    def anti_base(self):
        bt = [COS_ODD if (k // 2) % 2 == 0 else SIN_EVEN for k in range(self.npc)]
        return Base(np.full((self.npc, self.nt), CHEB), bt, COSSIN)

    def mult_base(self, a, b):
        if a is None or b is None:
            return None
        if a.bp != COSSIN or b.bp != COSSIN:
            return None
        bt = []
        for k in range(self.npc):
            pair = _MULT_T.get((int(a.bt[k]), int(b.bt[k])))
            if pair is None:
                return None
            bt.append(pair[0] if k % 4 < 2 else pair[1])
        if not (np.all(a.br == CHEB) and np.all(b.br == CHEB)):
            return None
        return Base(np.full((self.npc, self.nt), CHEB), bt, COSSIN)

    def radius_vd(self):
        return VD(self, c=self.R.copy(), base=None)

    def conf_field(self, values, anti=False):
        vals = np.broadcast_to(values, (self.np, self.nt, self.nr)).astype(float).copy()
        return VD(self, c=vals, base=self.anti_base() if anti else self.std_base())

    # ---- transforms
    def coef_arr(self, c, base):
        x = apply_phi(c, tr_mat(0, base.bp, self.np))
        x = apply_theta(x, [tr_mat(0, int(b), self.nt) for b in base.bt])
        return apply_r(x, base.br, lambda b: tr_mat(0, int(b), self.nr))

    def coef_i_arr(self, cf, base):
        x = apply_r(cf, base.br, lambda b: tr_mat(1, int(b), self.nr))
        x = apply_theta(x, [tr_mat(1, int(b), self.nt) for b in base.bt])
        return apply_phi(x, tr_mat(1, base.bp, self.np))

    # ---- coefficient-space 1-D operators (Base_spectral::ope_1d, ope_1d.cpp:25-82)
    def ope_theta(self, vd, fn):
        vd.coef()
        mats, nb = [], []
        for k in range(self.npc):
            M, ob = op_mat(fn, int(vd.base.bt[k]), self.nt)
            mats.append(M)
            nb.append(ob)
        base = vd.base.copy()
        base.bt = np.array(nb)
        return VD(self, cf=apply_theta(vd.cf, mats), base=base)

    def ope_r(self, vd, fn, scale=1.0):
        vd.coef()
        ob = op_mat(fn, CHEB, self.nr)[1]
        assert np.all(vd.base.br == CHEB)
        M = op_mat(fn, CHEB, self.nr)[0]
        base = vd.base.copy()
        base.br = np.full_like(base.br, ob)
        return VD(self, cf=(vd.cf @ M.T) * scale, base=base)

    def ope_phi_flip(self, vd, fn):
        """mult_cos_phi / mult_sin_phi (domain_shell_ope.cpp:31-107)."""
        vd.coef()
        M, _ = op_mat(fn, COSSIN, self.npc)
        base = vd.base.copy()
        base.bt = np.array([_FLIP_T[int(b)] for b in vd.base.bt])
        return VD(self, cf=apply_phi(vd.cf, M), base=base)

    def der_var(self, vd, var):
        """Val_domain::compute_der_var (val_domain.cpp:771-782), var = 1, 2, 3."""
        if vd.zero:
            return vd
        vd.coef()
        if var == 1:
            return self.ope_r(vd, der_1d)
        if var == 2:
            return self.ope_theta(vd, der_1d)
        M, ob = op_mat(der_1d, COSSIN, self.npc)
        base = vd.base.copy()
        return VD(self, cf=apply_phi(vd.cf, M), base=base)

    def mult_cos_phi(self, vd):
        return vd if vd.zero else self.ope_phi_flip(vd, mult_cos_1d)

    def mult_sin_phi(self, vd):
        return vd if vd.zero else self.ope_phi_flip(vd, mult_sin_1d)

    def mult_cos_theta(self, vd):
        return vd if vd.zero else self.ope_theta(vd, mult_cos_1d)

    def mult_sin_theta(self, vd):
        return vd if vd.zero else self.ope_theta(vd, mult_sin_1d)

    def div_sin_theta(self, vd):
        return vd if vd.zero else self.ope_theta(vd, div_sin_1d)

    def mult_xm1(self, vd):
        return vd if vd.zero else self.ope_r(vd, mult_xm1_1d)

    def div_xm1(self, vd):
        return vd if vd.zero else self.ope_r(vd, div_xm1_1d)

    def mult_r(self, vd):
        if vd.zero:
            return vd
        if self.kind == "shell":  # domain_shell_ope.cpp:155-164
            vd.coef()
            M, _ = op_mat(mult_x_1d, CHEB, self.nr)
            return VD(self, cf=(vd.cf @ M.T) * self.alpha + vd.cf * self.beta, base=vd.base.copy())
        r = self.div_xm1(vd)  # domain_compact_ope.cpp:173-178
        return vdiv_scal(r, self.alpha)

    def div_r(self, vd):
        if vd.zero:
            return vd
        if self.kind == "shell":  # domain_shell_ope.cpp:215-219
            res = vdiv(vd, self.radius_vd())
            res.base = vd.base.copy()
            return res
        return vmul_scal(self.mult_xm1(vd), self.alpha)  # domain_compact_ope.cpp:180-185

    def der_abs(self, vd):
        """do_der_abs_from_der_var: shell domain_shell.cpp:665-683, compact domain_compact.cpp:674-693."""
        if vd.zero:
            return [vd, vd, vd]
        d0, d1, d2 = (self.der_var(vd, v) for v in (1, 2, 3))
        if self.kind == "shell":
            sintdr = vdiv_scal(self.mult_sin_theta(d0), self.alpha)
            rad = self.radius_vd()
            dtsr = vdiv(d1, rad)
            dtsr.base = d1.base.copy() if d1.base is not None else None
            dpsr = vdiv(d2, rad)
            dpsr.base = d2.base.copy() if d2.base is not None else None
            costdtsr = self.mult_cos_theta(dtsr)
            dpsrssint = self.div_sin_theta(dpsr)
            s = vadd(sintdr, costdtsr)
            dx = vsub(self.mult_cos_phi(s), self.mult_sin_phi(dpsrssint))
            dy = vadd(self.mult_sin_phi(s), self.mult_cos_phi(dpsrssint))
            dz = vsub(vdiv_scal(self.mult_cos_theta(d0), self.alpha), self.mult_sin_theta(dtsr))
            return [dx, dy, dz]
        dr = vneg(self.mult_xm1(d0))
        sintdr = self.mult_sin_theta(dr)
        costdt = self.mult_cos_theta(d1)
        dpssint = self.div_sin_theta(d2)
        s = vadd(sintdr, costdt)
        ax = vsub(self.mult_cos_phi(s), self.mult_sin_phi(dpssint))
        dx = vmul_scal(self.mult_xm1(ax), self.alpha)
        ay = vadd(self.mult_sin_phi(s), self.mult_cos_phi(dpssint))
        dy = vmul_scal(self.mult_xm1(ay), self.alpha)
        az = vsub(self.mult_cos_theta(dr), self.mult_sin_theta(d1))
        dz = vmul_scal(self.mult_xm1(az), self.alpha)
        return [dx, dy, dz]

    def der_normal(self, vd, bound):
        """shell domain_shell.cpp:71-87 (der_var(1)/alpha); compact domain_compact.cpp:68-87."""
        if vd.zero:
            return vd
        d = self.der_var(vd, 1)
        if self.kind == "shell":
            return vdiv_scal(d, self.alpha)
        d = self.mult_xm1(self.mult_xm1(d))
        return vmul_scal(d, -self.alpha)

    def val_boundary_all(self, vd, bound):
        """val_boundary for every (k, j) at once (domain_shell_systems.cpp:43-74)."""
        vd.coef()
        sign = (-1.0) ** np.arange(self.nr) if bound == INNER_BC else np.ones(self.nr)
        return vd.cf @ sign

    def integ(self, vd, bound):
        """Domain_shell::integ (domain_shell.cpp:862-897); Domain_compact::integ (domain_compact.cpp:89-121)."""
        if vd.zero:
            return 0.0
        rr = self.mult_r(self.mult_r(self.mult_sin_theta(vd)))
        vb = self.val_boundary_all(rr, bound)  # (..., npc, nt)
        bt0 = int(rr.base.bt[0])
        if bt0 == COS_EVEN:
            res = np.pi * vb[..., 0, 0]
        elif bt0 == SIN_ODD:
            w = 2.0 / (2 * np.arange(self.nt) + 1.0)
            res = vb[..., 0, :] @ w
        else:
            res = 0.0 * vb[..., 0, 0]
        return res * 2 * np.pi

    # ---- tau bookkeeping (domain_shell_nbr_unknowns.cpp:26-66,
    #      domain_shell_affecte_tau_one_coef.cpp:26-105, domain_shell_affecte_tau.cpp:26-130,
    #      domain_shell_export_tau.cpp:27-120, domain_shell_export_tau_boundary.cpp:27-120;
    #      Domain_compact identical for these valences)
    def _true_tet(self, baset, j, k, kmin, quirk=False):
        nt = self.nt
        if baset == COS_EVEN:
            return not (j == 0 and k >= kmin)
        if baset == COS_ODD:
            return not (j == nt - 1 or (j == 0 and k >= kmin))
        if baset == SIN_EVEN:
            return not ((j == 1 and k >= kmin + 2) or j == 0 or j == nt - 1)
        if baset == SIN_ODD:
            # affecte_tau uses k > kmin+2 (domain_shell_affecte_tau.cpp:58), the others k >= kmin+2
            cond = (k > kmin + 2) if quirk else (k >= kmin + 2)
            return not ((j == 0 and cond) or j == nt - 1)
        raise ValueError(baset)

    def unknown_slots(self, bt, mlim=0):
        """(k, j) lines carrying unknowns, in the reference's order (k outer, j inner)."""
        kmin = 2 * mlim + 2
        out = []
        for k in range(self.npc - 1):
            if k == 1:
                continue
            for j in range(self.nt):
                if self._true_tet(int(bt[k]), j, k, kmin):
                    out.append((k, j))
        return out

    def nbr_unknowns(self, bt, mlim=0):
        return len(self.unknown_slots(bt, mlim)) * self.nr

    def seed_coefs(self, bt, mlim=0):
        """affecte_tau_one_coef for every unknown of one component: array (n_unknowns, npc, nt, nr)."""
        kmin = 2 * mlim + 2
        slots = self.unknown_slots(bt, mlim)
        n = len(slots) * self.nr
        out = np.zeros((n, self.npc, self.nt, self.nr))
        u = 0
        for (k, j) in slots:
            baset = int(bt[k])
            for i in range(self.nr):
                out[u, k, j, i] = 1.0
                if baset in (COS_EVEN, COS_ODD) and k >= kmin:
                    out[u, k, 0, i] = -1.0
                if baset == SIN_EVEN and k >= kmin + 2:
                    out[u, k, 1, i] = -j
                if baset == SIN_ODD and k >= kmin + 2:
                    out[u, k, 0, i] = -(2 * j + 1)
                u += 1
        return out

    def affecte_tau(self, bt, values, mlim=0):
        """affecte_tau_val_domain: consumes nbr of values, returns the coefficient array."""
        kmin = 2 * mlim + 2
        cf = np.zeros((self.npc, self.nt, self.nr))
        pos = 0
        for k in range(self.npc - 1):
            if k == 1:
                continue
            baset = int(bt[k])
            for j in range(self.nt):
                if self._true_tet(baset, j, k, kmin, quirk=True):
                    cf[k, j, :] += values[pos:pos + self.nr]
                    pos += self.nr
        for k in range(self.npc - 1):
            baset = int(bt[k])
            if baset in (COS_EVEN, COS_ODD) and k >= kmin:
                cf[k, 0, :] = -cf[k, 1:, :].sum(axis=0)
            elif baset == SIN_EVEN and k >= kmin + 2:
                cf[k, 1, :] = -(np.arange(2, self.nt)[:, None] * cf[k, 2:, :]).sum(axis=0)
            elif baset == SIN_ODD and k >= kmin + 2:
                cf[k, 0, :] = -((2 * np.arange(1, self.nt) + 1)[:, None] * cf[k, 1:, :]).sum(axis=0)
        return cf, pos

    def _tau_rows(self, bt, order, mlim=0):
        """Row recipe of export_tau_val_domain: list of (k, j, gal_j, gal_coeff) lines;
        each line exports i in [0, nr - order) (order = -1 marks the boundary version)."""
        kmin = 2 * mlim + 2
        rows = []
        for k in range(self.npc - 1):
            if k == 1:
                continue
            baset = int(bt[k])
            for j in range(self.nt):
                if baset == COS_EVEN:
                    if k < kmin:
                        rows.append((k, j, None, 0.0))
                    elif j != 0:
                        rows.append((k, j, 0, 2.0))
                elif baset == COS_ODD:
                    if j != self.nt - 1:
                        if k < kmin:
                            rows.append((k, j, None, 0.0))
                        elif j != 0:
                            rows.append((k, j, 0, 1.0))
                elif baset == SIN_EVEN:
                    if j != 0 and j != self.nt - 1:
                        if k < kmin + 2:
                            rows.append((k, j, None, 0.0))
                        elif j != 1:
                            rows.append((k, j, 1, float(j)))
                elif baset == SIN_ODD:
                    if j != self.nt - 1:
                        if k < kmin + 2:
                            rows.append((k, j, None, 0.0))
                        elif j != 0:
                            rows.append((k, j, 0, float(2 * j + 1)))
                else:
                    raise ValueError(baset)
        return rows

    def export_tau(self, vd, order, ncond=None):
        """export_tau_val_domain (domain_shell_export_tau.cpp:27-120): rows (k, j, i < nr-order)."""
        if vd.zero:
            return None
        vd.coef()
        rows = self._tau_rows(vd.base.bt, order)
        nri = self.nr - order
        out = []
        for (k, j, gj, g) in rows:
            v = vd.cf[..., k, j, :nri]
            if gj is not None:
                v = v - g * vd.cf[..., k, gj, :nri]
            out.append(v)
        return np.concatenate(out, axis=-1)

    def export_tau_boundary(self, vd, bound):
        """export_tau_val_domain_boundary (domain_shell_export_tau_boundary.cpp:27-120)."""
        if vd.zero:
            return None
        vb = self.val_boundary_all(vd, bound)
        rows = self._tau_rows(vd.base.bt, -1)
        out = []
        for (k, j, gj, g) in rows:
            v = vb[..., k, j]
            if gj is not None:
                v = v - g * vb[..., k, gj]
            out.append(v[..., None])
        return np.concatenate(out, axis=-1)

    def ncond(self, bt, order):
        rows = self._tau_rows(bt, order)
        return len(rows) * (self.nr - order) if order >= 0 else len(rows)


# ---------------------------------------------------------------------------
# Val_domain: lazy conf / coef representation (src/Val_domain/val_domain.cpp:617-653)
# ---------------------------------------------------------------------------
class VD:
    __slots__ = ("dom", "c", "cf", "base", "zero")

    def __init__(self, dom, c=None, cf=None, base=None, zero=False):
        self.dom, self.c, self.cf, self.base, self.zero = dom, c, cf, base, zero

    @staticmethod
    def zeros(dom):
        return VD(dom, zero=True)

    def coef(self):
        if self.zero or self.cf is not None:
            return
        if self.base is None:
            raise RuntimeError("Base not defined in Val_domain::coef")
        self.cf = self.dom.coef_arr(self.c, self.base)

    def coef_i(self):
        if self.zero or self.c is not None:
            return
        if self.base is None:
            raise RuntimeError("Base not defined in Val_domain::coef_i")
        self.c = self.dom.coef_i_arr(self.cf, self.base)

    def conf(self):
        if self.zero:
            return np.zeros((self.dom.np, self.dom.nt, self.dom.nr))
        self.coef_i()
        return self.c


def _chk_base(a, b):
    if a.base is not None and b.base is not None:
        if not (a.base == b.base):
            raise AssertionError(f"base mismatch in Val_domain +/-: {a.base} vs {b.base}")
        return a.base.copy()
    return None


def vadd(a, b):
    """operator+ (val_domain_math.cpp:150-195)."""
    if a.zero:
        return b
    if b.zero:
        return a
    base = _chk_base(a, b)
    if a.c is not None:
        b.coef_i()
        return VD(a.dom, c=a.c + b.c, base=base)
    if b.c is not None:
        a.coef_i()
        return VD(a.dom, c=a.c + b.c, base=base)
    return VD(a.dom, cf=a.cf + b.cf, base=base)


def vneg(a):
    """operator- (unary), val_domain_math.cpp:140-147."""
    if a.zero:
        return a
    return VD(a.dom, c=None if a.c is None else -a.c, cf=None if a.cf is None else -a.cf,
              base=None if a.base is None else a.base.copy())


def vsub(a, b):
    """operator- (val_domain_math.cpp:220-262)."""
    if a.zero:
        return vneg(b)
    if b.zero:
        return a
    base = _chk_base(a, b)
    if a.c is not None:
        b.coef_i()
        return VD(a.dom, c=a.c - b.c, base=base)
    if b.c is not None:
        a.coef_i()
        return VD(a.dom, c=a.c - b.c, base=base)
    return VD(a.dom, cf=a.cf - b.cf, base=base)


def vmul(a, b):
    """operator* (val_domain_math.cpp:302-318): conf product, base by Domain::mult."""
    if a.zero:
        return a
    if b.zero:
        return b
    a.coef_i()
    b.coef_i()
    return VD(a.dom, c=a.c * b.c, base=a.dom.mult_base(a.base, b.base))


def vdiv(a, b):
    """operator/ (val_domain_math.cpp:382-399)."""
    if a.zero:
        return a
    if b.zero:
        raise ZeroDivisionError("Division by zero")
    a.coef_i()
    b.coef_i()
    return VD(a.dom, c=a.c / b.c, base=a.dom.mult_base(a.base, b.base))


def vmul_scal(a, x):
    """operator*(Val_domain, double) (val_domain_math.cpp:320-331)."""
    if a.zero:
        return a
    return VD(a.dom, c=None if a.c is None else a.c * x, cf=None if a.cf is None else a.cf * x,
              base=None if a.base is None else a.base.copy())


def vdiv_scal(a, x):
    """operator/(Val_domain, double) (val_domain_math.cpp:401-412)."""
    if a.zero:
        return a
    return VD(a.dom, c=None if a.c is None else a.c / x, cf=None if a.cf is None else a.cf / x,
              base=None if a.base is None else a.base.copy())


def vadd_scal(a, x):
    """operator+(Val_domain, double) (val_domain_math.cpp:197-214)."""
    if a.zero:
        return VD(a.dom, c=np.full((a.dom.np, a.dom.nt, a.dom.nr), float(x)), base=a.dom.std_base())
    a.coef_i()
    return VD(a.dom, c=a.c + x, base=None if a.base is None else a.base.copy())


def vsub_scal(a, x):
    return vadd_scal(a, -x)
