"""Python handle on systems given as Kadath equation strings (C-ABI kad_sysdef_* /
kad_system_*), the counterpart of the reference's System_of_eqs front end
(src/System_of_eqs/system_of_eqs.cpp, add_eq.cpp, give_ope.cpp).

    sd = SystemDef(ctx, (13, 5, 4), [1.0, 2.0, 4.0])      # Space_spheric + System_of_eqs(space, 1, ndom-1)
    sd.add_var("F"); sd.add_cst("a", 1.0)
    for d in range(1, 4): sd.add_eq_inside(d, "Lap(F)=0")
    sd.add_eq_bc(1, INNER_BC, "dn(F) + 0.5*F/a = 0"); sd.add_eq_bc(3, OUTER_BC, "F=1")
    for d in range(1, 3): sd.add_eq_matching(d, OUTER_BC, "F"); sd.add_eq_matching(d, OUTER_BC, "dn(F)")
    s = sd.build(); s.set_field("F", ...); s.do_newton(1e-8)

Residual(s) is a torch.autograd.Function u -> F(u) whose backward is the VJP kernel
(kad_vjp: J^T w with the tangent programs' row exports transposed).
"""
import ctypes

import numpy as np
import torch

from ._lib import lib, check, c_vp
from .xcts import SOLVERS

INNER_BC, OUTER_BC = 0, 1


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _b(s):
    return s.encode()


class SystemDef:
    """Declaration of a system (host only): unknowns, constants, definitions, equations."""

    def __init__(self, ctx, res, bounds, dmin=1, dmax=None):
        self.ctx = ctx
        self.res = tuple(res)
        self.bounds = np.asarray(bounds, dtype=np.float64)
        self.dmin = dmin
        self.dmax = len(self.bounds) if dmax is None else dmax
        h = c_vp()
        check(lib.kad_sysdef_create(res[0], res[1], res[2], _dp(self.bounds), len(self.bounds), self.dmin, self.dmax,
                                    ctypes.byref(h)), "kad_sysdef_create")
        self.h = h
        self.fields = []   # scalar field components in declaration order
        self.params = {}
        self.doubles = {}
        self.globals = []
        self._npar = 0

    def close(self):
        if self.h:
            lib.kad_sysdef_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_var(self, name, valence=0):
        check(lib.kad_sysdef_add_var(self.h, _b(name), valence), f"add_var({name})")
        self.fields += [name] if valence == 0 else [f"{name}^{c}" for c in "xyz"]

    def add_cst(self, name, value):
        check(lib.kad_sysdef_add_cst(self.h, _b(name), float(value)), f"add_cst({name})")

    def add_param(self, name, value):
        idx = check(lib.kad_sysdef_add_param(self.h, _b(name), float(value)), f"add_param({name})")
        self.params[name] = idx
        self._npar += 1
        return idx

    def add_var_double(self, name, value):
        """A global (double) unknown; returns its index among the unknowns."""
        idx = check(lib.kad_sysdef_add_var_double(self.h, _b(name), float(value)), f"add_var({name})")
        self.doubles[name] = idx
        self.params[name] = self._npar  # it is also a parameter (set_param, param_tangent)
        self._npar += 1
        return idx

    def add_global(self, name, dom, bound, expr):
        """A global quantity (integral expression), reported by DeviceSystem.globals()."""
        check(lib.kad_sysdef_add_global(self.h, _b(name), dom, bound, _b(expr)), f"add_global({name})")
        self.globals.append(name)

    def add_eq_int(self, dom, bound, eq):
        check(lib.kad_sysdef_add_eq_int(self.h, dom, bound, _b(eq)), f"add_eq_int({eq})")

    def add_cst_field(self, name, conf, valence=0):
        """conf: array (components, domains dmin..dmax, nr*nt*np) of collocation values."""
        c = np.ascontiguousarray(conf, dtype=np.float64).ravel()
        check(lib.kad_sysdef_add_cst_field(self.h, _b(name), valence, _dp(c), c.size), f"add_cst({name})")

    def set_flat_metric(self, name="f"):
        check(lib.kad_sysdef_set_flat_metric(self.h, _b(name)), "set_flat_metric")

    def add_def(self, text):
        check(lib.kad_sysdef_add_def(self.h, _b(text)), f"add_def({text})")

    def add_eq_inside(self, dom, eq, order=2):
        check(lib.kad_sysdef_add_eq_inside(self.h, dom, _b(eq), order), f"add_eq_inside({eq})")

    def add_eq_bc(self, dom, bound, eq):
        check(lib.kad_sysdef_add_eq_bc(self.h, dom, bound, _b(eq)), f"add_eq_bc({eq})")

    def add_eq_matching(self, dom, bound, expr):
        check(lib.kad_sysdef_add_eq_matching(self.h, dom, bound, _b(expr)), f"add_eq_matching({expr})")

    def space_add_eq(self, eq, m1, m2):
        check(lib.kad_sysdef_space_add_eq(self.h, _b(eq), _b(m1), _b(m2)), f"add_eq({eq})")

    def space_add_inner_bc(self, eq):
        check(lib.kad_sysdef_space_add_bc(self.h, INNER_BC, _b(eq)), f"add_inner_bc({eq})")

    def space_add_outer_bc(self, eq):
        check(lib.kad_sysdef_space_add_bc(self.h, OUTER_BC, _b(eq)), f"add_outer_bc({eq})")

    def set_adapted(self, d):
        """Domain d is a round shell of an adapted space (Space_adapted_bh's homothetic
        shells): its surface integrals weight in configuration space."""
        check(lib.kad_sysdef_set_adapted(self.h, int(d)), f"set_adapted({d})")

    def coloc(self, d):
        """(r, theta, phi) at the collocation points of domain d, each (np, nt, nr)."""
        n = self.res[0] * self.res[1] * self.res[2]
        r, t, p = np.zeros(n), np.zeros(n), np.zeros(n)
        check(lib.kad_sysdef_coloc(self.h, d, _dp(r), _dp(t), _dp(p)), "kad_sysdef_coloc")
        shp = (self.res[2], self.res[1], self.res[0])
        return r.reshape(shp), t.reshape(shp), p.reshape(shp)

    def host_trace(self, order=1):
        """Host-only trace (no device): sizes and row layout; the declaration is consumed."""
        info = (ctypes.c_int * 6)()
        cap = 5 * 4096
        rl = (ctypes.c_int * cap)()
        d, self.h = self.h, None
        q = lib.kad_sysdef_host_trace(d, order, info, rl, cap)
        lib.kad_sysdef_destroy(d)
        check(q, "kad_sysdef_host_trace")
        keys = ("n_unknowns", "n_rows", "value_instr", "tangent_instr", "value_slots", "tangent_programs")
        return dict(zip(keys, list(info))), [tuple(rl[5 * i:5 * i + 5]) for i in range(q)]

    def build(self, order=1, batch=4096):
        """Trace and upload; the declaration is consumed."""
        h = c_vp()
        d, self.h = self.h, None
        check(lib.kad_system_build(self.ctx.h, d, order, batch, ctypes.byref(h)), "kad_system_build")
        return DeviceSystem(self.ctx, h, self.res, self.dmin, self.dmax, list(self.fields), dict(self.params),
                            dict(self.doubles), list(self.globals))


class DeviceSystem:
    def __init__(self, ctx, h, res, dmin, dmax, fields, params, doubles=None, globals_=None):
        self.ctx, self.h, self.res, self.dmin, self.dmax = ctx, h, res, dmin, dmax
        self.fields, self.params, self.doubles = fields, params, doubles or {}
        self.global_names = globals_ or []
        ctx.adopt(self)
        n, m, nf = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(lib.kad_system_sizes(h, ctypes.byref(n), ctypes.byref(m), ctypes.byref(nf)), "kad_system_sizes")
        self.n, self.m = n.value, m.value
        cap = 4 * 4096
        lay = (ctypes.c_int * cap)()
        nl = check(lib.kad_system_layout(h, lay, cap), "kad_system_layout")
        self.layout = [tuple(lay[4 * q:4 * q + 4]) for q in range(nl)]
        rl = (ctypes.c_int * (5 * 4096))()
        nr = check(lib.kad_system_row_layout(h, rl, 5 * 4096), "kad_system_row_layout")
        self.row_layout = [tuple(rl[5 * q:5 * q + 5]) for q in range(nr)]

    def close(self):
        if self.h:
            lib.kad_system_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _f(self, f):
        return self.fields.index(f) if isinstance(f, str) else f

    def set_field(self, f, d, conf):
        c = np.ascontiguousarray(conf, dtype=np.float64).ravel()
        check(lib.kad_system_set_field(self.h, self._f(f), d, _dp(c)), "kad_system_set_field")

    def get_field(self, f, d):
        out = np.zeros(self.res[0] * self.res[1] * self.res[2])
        check(lib.kad_system_get_field(self.h, self._f(f), d, _dp(out)), "kad_system_get_field")
        return out.reshape(self.res[2], self.res[1], self.res[0])

    def coef(self, f, d):
        nr, nt, npt = self.res
        out = np.zeros((npt + 2) * nt * nr)
        check(lib.kad_system_get_coef(self.h, self._f(f), d, _dp(out)), "kad_system_get_coef")
        return out.reshape(npt + 2, nt, nr)

    def set_coef(self, f, d, cf):
        cf = np.ascontiguousarray(cf, dtype=np.float64).ravel()
        check(lib.kad_system_set_coef(self.h, self._f(f), d, _dp(cf)), "kad_system_set_coef")

    def add_unknowns(self, X0):
        X0 = np.ascontiguousarray(X0, dtype=np.float64)
        check(lib.kad_system_add_unknowns(self.h, _dp(X0)), "kad_system_add_unknowns")

    def update(self, X):
        self.ctx._enter()
        check(lib.kad_system_update(self.h, ctypes.c_void_p(X.data_ptr())), "kad_system_update")

    def perturbation(self, seed=20261202, amp=1e-3):
        out = np.zeros(self.n)
        check(lib.kad_system_perturbation(self.h, seed, amp, _dp(out)), "kad_system_perturbation")
        return out

    def globals(self):
        out = np.zeros(max(1, len(self.global_names)))
        check(lib.kad_system_globals(self.h, _dp(out)), "kad_system_globals")
        return dict(zip(self.global_names, out))

    def sensitivity(self, param, with_dqdu=False):
        """Adjoint dQ/dparam of the global quantities at the current (converged) fields."""
        nq = len(self.global_names)
        out = np.zeros(max(1, nq))
        dq = np.zeros((nq, self.n)) if with_dqdu else None
        self.ctx._enter()
        check(lib.kad_system_sensitivity(self.h, self.params[param], _dp(out), _dp(dq) if dq is not None else None),
              "kad_system_sensitivity")
        res = dict(zip(self.global_names, out))
        return (res, dq) if with_dqdu else res

    def get_double(self, name):
        v = ctypes.c_double()
        check(lib.kad_system_get_double(self.h, self.doubles[name], ctypes.byref(v)), "kad_system_get_double")
        return v.value

    def set_param(self, name, value):
        check(lib.kad_system_set_param(self.h, self.params[name], float(value)), "kad_system_set_param")

    def residual(self):
        F = torch.empty(self.m, dtype=torch.float64, device="cuda")
        self.ctx._enter()
        check(lib.kad_residual(self.h, ctypes.c_void_p(F.data_ptr())), "kad_residual")
        return F

    def check_equations(self):
        e = np.zeros(4096)
        q = check(lib.kad_check_equations(self.h, _dp(e), e.size), "kad_check_equations")
        return e[:q]

    def jvp_fill(self, col_begin, col_count, dtype=torch.float64, ld=None):
        """Columns [col_begin, col_begin + col_count) of J as a (m, count) tensor."""
        ld = self.m if ld is None else ld
        T = torch.empty((col_count, ld), dtype=dtype, device="cuda")  # column-major tile = T.T
        self.ctx._enter()
        check(lib.kad_jvp_fill(self.h, col_begin, col_count, ctypes.c_void_p(T.data_ptr()), ld,
                               0 if dtype == torch.float64 else 1), "kad_jvp_fill")
        return T.T[:self.m]

    def jacobian(self):
        return self.jvp_fill(0, self.n)

    def vjp(self, w):
        """J^T w, matrix-free."""
        w = w.to(device="cuda", dtype=torch.float64).contiguous()
        out = torch.empty(self.n, dtype=torch.float64, device="cuda")
        self.ctx._enter()
        check(lib.kad_vjp(self.h, ctypes.c_void_p(w.data_ptr()), ctypes.c_void_p(out.data_ptr())), "kad_vjp")
        return out

    def param_tangent(self, name):
        out = torch.empty(self.m, dtype=torch.float64, device="cuda")
        self.ctx._enter()
        check(lib.kad_param_tangent(self.h, self.params[name], ctypes.c_void_p(out.data_ptr())), "kad_param_tangent")
        return out

    def set_option(self, forcing=None, reuse_inner=None, drop_factors=False, fill_streams=None,
                   static_pivoting=None, fuse=None):
        """Iterative solver options: forcing factor of the refinement target, the
        inner-step threshold for reusing earlier factors, and dropping the factors kept
        for reuse (the next step factors its own Jacobian, with partial pivoting); the
        number of streams (1-4) of the Jacobian fill; static pivoting in the
        domain-decomposed solvers (True, the default, or False); element-wise
        instructions applied in the spectral pass that feeds them (True, the default)."""
        if fill_streams is not None:
            check(lib.kad_system_set_option(self.h, 3, float(fill_streams)), "kad_system_set_option")
        if static_pivoting is not None:
            check(lib.kad_system_set_option(self.h, 4, 1.0 if static_pivoting else 0.0), "kad_system_set_option")
        if fuse is not None:
            check(lib.kad_system_set_option(self.h, 5, 1.0 if fuse else 0.0), "kad_system_set_option")
        if drop_factors:
            check(lib.kad_system_set_option(self.h, 2, 0.0), "kad_system_set_option")
        if forcing is not None:
            check(lib.kad_system_set_option(self.h, 0, float(forcing)), "kad_system_set_option")
        if reuse_inner is not None:
            check(lib.kad_system_set_option(self.h, 1, float(reuse_inner)), "kad_system_set_option")

    def dd_blocks(self):
        """The (rows, unknowns) of each domain block of the domain-decomposed solvers
        (kad_system_dd_blocks)."""
        sizes = (ctypes.c_int * 64)()
        rows = np.empty(self.n, dtype=np.int32)
        cols = np.empty(self.n, dtype=np.int32)
        nb = check(lib.kad_system_dd_blocks(self.h, sizes, 64, rows.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                            cols.ctypes.data_as(ctypes.POINTER(ctypes.c_int))),
                   "kad_system_dd_blocks")
        out, o = [], 0
        for b in range(nb):
            out.append((rows[o:o + sizes[b]].copy(), cols[o:o + sizes[b]].copy()))
            o += sizes[b]
        return out

    def dd_info(self):
        """Sizes and set-up times (s) of the last domain-decomposed factorisation."""
        sz = (ctypes.c_int * 4)()
        tm = (ctypes.c_double * 5)()
        check(lib.kad_system_dd_info(self.h, sz, tm), "kad_system_dd_info")
        st = ctypes.c_int()
        g = ctypes.c_double()
        check(lib.kad_system_dd_pivoting(self.h, ctypes.byref(st), ctypes.byref(g)), "kad_system_dd_pivoting")
        return dict(nd=sz[0], coupling_rows=sz[1], kmax=sz[2], blocks=sz[3],
                    times=dict(gather=tm[0], block_lu=tm[1], spikes=tm[2], capacitance=tm[3], capacitance_lu=tm[4]),
                    static_pivoting=bool(st.value), growth=g.value)

    def do_newton(self, prec, solver="f64"):
        err = ctypes.c_double()
        times = np.zeros(5)
        self.ctx._enter()
        rc = lib.kad_system_do_newton(self.h, prec, ctypes.byref(err), SOLVERS[solver], _dp(times))
        if rc == -5:
            raise RuntimeError("Jacobian is singular! Stopping.")
        check(rc, "kad_system_do_newton")
        it = (ctypes.c_int * 4)()
        res = ctypes.c_double()
        lib.kad_system_last_solve(self.h, it, ctypes.byref(res))
        return rc == 1, err.value, times, (it[0], it[1], res.value, it[2], it[3])

    def do_newton_with_linesearch(self, prec, ntrymax=10, stepmax=1.0, solver="f64"):
        """System_of_eqs::do_newton_with_linesearch (src_par/do_newton.cpp:454-517):
        returns (converged at entry, ||F||_inf at entry, lambda applied, residual
        evaluations of the search)."""
        err = ctypes.c_double()
        lam = ctypes.c_double()
        ntry = ctypes.c_int()
        self.ctx._enter()
        rc = lib.kad_system_do_newton_linesearch(self.h, prec, ctypes.byref(err), int(ntrymax), float(stepmax),
                                                 SOLVERS[solver], ctypes.byref(lam), ctypes.byref(ntry))
        if rc == -5:
            raise RuntimeError("Jacobian is singular! Stopping.")
        if rc == -4:
            raise RuntimeError("roundoff error in do_newton_with_linesearch")
        check(rc, "kad_system_do_newton_linesearch")
        return rc == 1, err.value, lam.value, ntry.value

    def solve(self, precision=1e-10, solver="f64", max_iter=20):
        for it in range(max_iter + 1):
            conv, err, _, _ = self.do_newton(precision, solver)
            if conv:
                return it, err
        raise RuntimeError(f"Newton did not reach {precision} in {max_iter} iterations (|F| = {err:.3e})")


class Residual(torch.autograd.Function):
    """u -> F(u) for a DeviceSystem: u is the unknown vector added to the fields held by
    the system (u = 0: the current fields).  backward: grad_u = J^T grad_F by the VJP
    kernels at the linearisation point, with no J stored."""

    @staticmethod
    def _at(s, u, fn):
        """fn() evaluated with the fields moved by u, then restored bit for bit."""
        if not bool(torch.any(u != 0)):
            return fn()
        saved = [(f, d, s.coef(f, d)) for f in range(len(s.fields)) for d in range(s.dmin, s.dmax + 1)]
        s.update(-u.to(device="cuda", dtype=torch.float64).contiguous())  # fields += affecte_tau(u)
        try:
            return fn()
        finally:
            for f, d, cf in saved:
                s.set_coef(f, d, cf)

    @staticmethod
    def forward(ctx_, u, system):
        ctx_.system = system
        ctx_.save_for_backward(u)
        return Residual._at(system, u, system.residual)

    @staticmethod
    def backward(ctx_, gF):
        (u,) = ctx_.saved_tensors
        s = ctx_.system
        return Residual._at(s, u, lambda: s.vjp(gF)), None
