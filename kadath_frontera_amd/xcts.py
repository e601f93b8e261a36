"""Python handle on the device XCTS Newton system (C-ABI kad_xcts_*).

Mirrors the driver loop of the reference (codes/Par_version/Kerr/kerr.cpp:102-112):
    while not converged: converged = system.do_newton(prec)
"""
import ctypes

import numpy as np
import torch

from ._lib import lib, check, c_vp

FIELDS = ("P", "N", "bx", "by", "bz")
SOLVERS = {"f64": 0, "f32-ir": 1, "f32-gmres": 2, "bf16-gmres": 3, "gauss-newton": 4, "f64-reuse": 5,
           "f64-dd": 6, "f64-dd-reuse": 7}


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def host_info(res, bounds, n0=0.5, a=1.0, ome=0.0, order=0):
    """Host-only trace of the system: sizes and (field, domain, offset, count) layout."""
    b = np.asarray(bounds, dtype=np.float64)
    info = (ctypes.c_int * 6)()
    cap = 4 * 5 * 64
    lay = (ctypes.c_int * cap)()
    nl = check(lib.kad_xcts_host_info(res[0], res[1], res[2], _dp(b), len(b), 1, len(b), n0, a, ome, order, info, lay,
                                      cap),
               "kad_xcts_host_info")
    layout = [tuple(lay[4 * q:4 * q + 4]) for q in range(nl)]
    keys = ("n_unknowns", "n_rows", "value_instr", "tangent_instr", "value_slots", "tangent_programs")
    return dict(zip(keys, list(info))), layout


def host_dd_blocks(res, bounds, order=0):
    """Host-only trace of the system: the (rows, unknowns) of each domain block of the
    domain-decomposed solvers (kad_xcts_host_dd_blocks)."""
    b = np.asarray(bounds, dtype=np.float64)
    info, _ = host_info(res, bounds, order=order)
    n = info["n_unknowns"]
    sizes = (ctypes.c_int * 64)()
    rows = np.empty(n, dtype=np.int32)
    cols = np.empty(n, dtype=np.int32)
    ip = ctypes.POINTER(ctypes.c_int)
    nb = check(lib.kad_xcts_host_dd_blocks(res[0], res[1], res[2], _dp(b), len(b), 1, len(b), order, sizes, 64,
                                           rows.ctypes.data_as(ip), cols.ctypes.data_as(ip)),
               "kad_xcts_host_dd_blocks")
    out, o = [], 0
    for q in range(nb):
        out.append((rows[o:o + sizes[q]].copy(), cols[o:o + sizes[q]].copy()))
        o += sizes[q]
    return out


class XctsSystem:
    def __init__(self, ctx, res, bounds, n0=0.5, a=1.0, ome=0.0, batch=256, order=0):
        """order 0: the reference's unknown/row order; 1: domain-major (block-tridiagonal J)."""
        self.ctx = ctx
        self.res = tuple(res)
        self.bounds = np.asarray(bounds, dtype=np.float64)
        h = c_vp()
        check(lib.kad_xcts_create(ctx.h, res[0], res[1], res[2], _dp(self.bounds), len(self.bounds), 1,
                                  len(self.bounds), n0, a, ome, order, batch, ctypes.byref(h)), "kad_xcts_create")
        self.h = h
        ctx.adopt(self)
        self.order = order
        n, m = ctypes.c_int(), ctypes.c_int()
        st = (ctypes.c_int * 4)()
        check(lib.kad_xcts_sizes(h, ctypes.byref(n), ctypes.byref(m), st), "kad_xcts_sizes")
        self.n, self.m = n.value, m.value
        self.stats = dict(value_instr=st[0], tangent_instr=st[1], value_slots=st[2], tangent_slots=st[3])
        cap = 4 * 5 * 64
        lay = (ctypes.c_int * cap)()
        nl = check(lib.kad_xcts_layout(h, lay, cap), "kad_xcts_layout")
        self.layout = [tuple(lay[4 * q:4 * q + 4]) for q in range(nl)]
        rcap = 5 * 1024
        rl = (ctypes.c_int * rcap)()
        nr = check(lib.kad_xcts_row_layout(h, rl, rcap), "kad_xcts_row_layout")
        # (kind, domain, component, first row, row count) per equation component
        self.row_layout = [tuple(rl[5 * q:5 * q + 5]) for q in range(nr)]

    def close(self):
        if self.h:
            lib.kad_xcts_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_guess(self, X0=None):
        if X0 is None:
            check(lib.kad_xcts_set_guess(self.h, None), "kad_xcts_set_guess")
        else:
            X0 = np.ascontiguousarray(X0, dtype=np.float64)
            check(lib.kad_xcts_set_guess(self.h, _dp(X0)), "kad_xcts_set_guess")

    def perturbation(self, seed=20261202, amp=1e-3):
        out = np.zeros(self.n)
        check(lib.kad_xcts_perturbation(self.h, seed, amp, _dp(out)), "kad_xcts_perturbation")
        return out

    def residual(self):
        F = torch.empty(self.m, dtype=torch.float64, device="cuda")
        self.ctx._enter()
        check(lib.kad_xcts_residual(self.h, ctypes.c_void_p(F.data_ptr())), "kad_xcts_residual")
        return F

    def jacobian(self):
        J = torch.empty((self.n, self.m), dtype=torch.float64, device="cuda")  # column-major view J.T
        self.ctx._enter()
        check(lib.kad_xcts_jacobian(self.h, ctypes.c_void_p(J.data_ptr()), self.m), "kad_xcts_jacobian")
        return J.T

    def update(self, X):
        self.ctx._enter()
        check(lib.kad_xcts_update(self.h, ctypes.c_void_p(X.data_ptr())), "kad_xcts_update")

    def set_coef(self, f, d, cf):
        cf = np.ascontiguousarray(cf, dtype=np.float64)
        check(lib.kad_xcts_set_coef(self.h, f, d, _dp(cf)), "kad_xcts_set_coef")

    def observables(self):
        """(M_ADM, M_Komar, J_ADM) of the current fields (surface integrals at infinity)."""
        out = np.zeros(3)
        check(lib.kad_xcts_observables(self.h, _dp(out)), "kad_xcts_observables")
        return dict(M_adm=out[0], M_komar=out[1], J_adm=out[2])

    def set_ome(self, value):
        """Horizon rotation Ome of the inner shift condition."""
        check(lib.kad_xcts_set_param(self.h, 0, float(value)), "kad_xcts_set_param")

    def sensitivity(self, with_dqdu=False):
        """Adjoint dQ/dOme of Q = (M_ADM, M_Komar, J_ADM) at the current (converged) fields:
        J^T lambda = dQ/du with the LU factors of J, dQ/dOme = -lambda . dF/dOme."""
        out = np.zeros(3)
        dq = np.zeros((3, self.n)) if with_dqdu else None
        check(lib.kad_xcts_sensitivity(self.h, _dp(out), _dp(dq) if dq is not None else None),
              "kad_xcts_sensitivity")
        res = dict(M_adm=out[0], M_komar=out[1], J_adm=out[2])
        return (res, dq) if with_dqdu else res

    def solve(self, precision=1e-10, solver="f64", guess=None, max_iter=20):
        """Newton iterations (kerr.cpp:102-112) until ||F||_inf < precision; returns the
        number of iterations and the final error."""
        self.set_guess(guess)
        for it in range(max_iter + 1):
            conv, err, _, _ = self.do_newton(precision, solver)
            if conv:
                return it, err
        raise RuntimeError(f"Newton did not reach {precision} in {max_iter} iterations (|F| = {err:.3e})")

    def coef(self, f, d):
        nr, nt, npt = self.res
        out = np.zeros((npt + 2) * nt * nr)
        check(lib.kad_xcts_get_coef(self.h, f, d, _dp(out)), "kad_xcts_get_coef")
        return out.reshape(npt + 2, nt, nr)

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
        rc = lib.kad_xcts_do_newton(self.h, prec, ctypes.byref(err), SOLVERS[solver], _dp(times))
        if rc == -5:
            raise RuntimeError("Jacobian is singular! Stopping.")
        check(rc, "kad_xcts_do_newton")
        it = (ctypes.c_int * 4)()
        res = ctypes.c_double()
        lib.kad_xcts_last_solve(self.h, it, ctypes.byref(res))
        # (outer refinement steps, inner GMRES steps, relative residual, fp64 fallback, factored)
        return rc == 1, err.value, times, (it[0], it[1], res.value, it[2], it[3])
