"""Python handle on a device context of libkadath_amd (torch tensors in, HIP kernels out).

torch supplies device memory; all work runs on the context's own HIP stream,
exposed to torch as an ExternalStream so callers can time it with torch.cuda.Event
or order torch work after it.
"""
import ctypes
import weakref

import numpy as np
import torch

from ._lib import lib, check, c_vp

PREC = {"f64": 0, "f32": 1, "bf16": 2}


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _iarr(a):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.int32))
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


class Context:
    def __init__(self, device=0):
        if not torch.cuda.is_available():
            raise RuntimeError("kadath_amd.Context needs a GPU (torch.cuda.is_available() is False)")
        self.device = device
        h = c_vp()
        check(lib.kad_ctx_create(ctypes.byref(h), device), "kad_ctx_create")
        self.h = h
        self.stream = torch.cuda.ExternalStream(lib.kad_ctx_stream(h), device=torch.device("cuda", device))
        # systems built on this context: a system's device state belongs to its context,
        # so they are destroyed before it (kad_system_destroy uses the context's stream)
        self._systems = weakref.WeakSet()

    def adopt(self, system):
        self._systems.add(system)

    def close(self):
        if self.h:
            for sysd in list(self._systems):
                sysd.close()
            lib.kad_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ordering helpers: torch work on the current stream must finish before ours
    def _enter(self):
        torch.cuda.current_stream(self.device).synchronize()

    def sync(self):
        check(lib.kad_sync(self.h), "kad_sync")

    # ---- dense --------------------------------------------------------------
    def gemm(self, prec, op, A, B, C):
        """C -= A@B (op 0) or C = A@B (op 1); column-major views: pass tensors whose
        transposes are contiguous, i.e. A.T.is_contiguous()."""
        m, k = A.shape
        k2, n = B.shape
        assert k == k2 and C.shape == (m, n)
        self._enter()
        check(lib.kad_gemm(self.h, PREC[prec], op, m, n, k, _ptr(A), A.stride(1), _ptr(B), B.stride(1), _ptr(C),
                           C.stride(1)), "kad_gemm")

    def lu_factor(self, prec, A, nb=0):
        """In-place LU of the column-major n x n A (A.T contiguous, dtype f64 or f32)."""
        n = A.shape[0]
        ipiv = torch.empty(n, dtype=torch.int32, device=A.device)
        perm = torch.empty(n, dtype=torch.int32, device=A.device)
        self._enter()
        info = check(lib.kad_lu_factor(self.h, PREC[prec], n, _ptr(A), A.stride(1), _ptr(ipiv), _ptr(perm), nb),
                     "kad_lu_factor")
        return ipiv, perm, info

    def lu_solve(self, prec, LU, perm, b, trans=0):
        n = LU.shape[0]
        self._enter()
        check(lib.kad_lu_solve(self.h, PREC[prec], n, _ptr(LU), LU.stride(1), _ptr(perm), _ptr(b), trans),
              "kad_lu_solve")
        self.sync()
        return b

    def lu_factor_batched(self, prec, A):
        """In-place LU of the column-major matrices A[z] (A of shape (nbat, n, n) with
        A[z].T contiguous), factored together."""
        nbat, n = A.shape[0], A.shape[1]
        if A.stride(1) != 1:
            raise ValueError("lu_factor_batched: A[z] must be column-major")
        ipiv = torch.empty(nbat, n, dtype=torch.int32, device=A.device)
        perm = torch.empty(nbat, n, dtype=torch.int32, device=A.device)
        info = (ctypes.c_int * nbat)()
        self._enter()
        check(lib.kad_lu_factor_batched(self.h, PREC[prec], n, nbat, _ptr(A), A.stride(2), A.stride(0), _ptr(ipiv),
                                        _ptr(perm), info), "kad_lu_factor_batched")
        return ipiv, perm, list(info)

    def lu_factor_batched_static(self, prec, A):
        """In-place LU of the column-major matrices A[z] without row interchanges (their rows
        already in an earlier factorisation's pivot order); returns (info, growth) per matrix."""
        nbat, n = A.shape[0], A.shape[1]
        if A.stride(1) != 1:
            raise ValueError("lu_factor_batched_static: A[z] must be column-major")
        ipiv = torch.empty(nbat, n, dtype=torch.int32, device=A.device)
        info = (ctypes.c_int * nbat)()
        growth = (ctypes.c_double * nbat)()
        self._enter()
        check(lib.kad_lu_factor_batched_static(self.h, PREC[prec], n, nbat, _ptr(A), A.stride(2), A.stride(0),
                                               _ptr(ipiv), info, growth), "kad_lu_factor_batched_static")
        return list(info), list(growth)

    def lu_solve_batched(self, LU, perm, B):
        """B[z] <- A[z]^{-1} B[z] with fp64 batched factors (B: (nbat, n, nrhs), B[z].T contiguous)."""
        nbat, n, nrhs = B.shape
        if B.stride(1) != 1 or LU.stride(1) != 1:
            raise ValueError("lu_solve_batched: B[z] and LU[z] must be column-major")
        ldb = B.stride(2) if nrhs > 1 else n
        self._enter()
        check(lib.kad_lu_solve_batched(self.h, n, nbat, _ptr(LU), LU.stride(2), LU.stride(0), _ptr(perm), _ptr(B),
                                       ldb, B.stride(0), nrhs), "kad_lu_solve_batched")
        return B

    def gemv64(self, A, x, b=None, trans=0):
        n = A.shape[0]
        y = torch.empty(n, dtype=torch.float64, device=A.device)
        self._enter()
        check(lib.kad_gemv64(self.h, n, _ptr(A), A.stride(1), _ptr(x), _ptr(b), _ptr(y), trans), "kad_gemv64")
        self.sync()
        return y

    def cast_f32(self, A, out):
        n = A.shape[0]
        self._enter()
        check(lib.kad_cast_f64_to_f32(self.h, n, _ptr(A), A.stride(1), _ptr(out), out.stride(1)), "kad_cast")

    def ir_solve(self, A64, prec_lo, LUlo, perm, b, tol=1e-13, max_outer=20, mode=0, trans=0):
        n = A64.shape[0]
        x = torch.empty(n, dtype=torch.float64, device=A64.device)
        iters = (ctypes.c_int * 2)()
        resid = ctypes.c_double()
        self._enter()
        rc = lib.kad_ir_solve(self.h, n, _ptr(A64), A64.stride(1), PREC[prec_lo], _ptr(LUlo), LUlo.stride(1),
                              _ptr(perm), _ptr(b), _ptr(x), tol, max_outer, mode, trans, iters, ctypes.byref(resid))
        if rc < 0 and rc != -4:
            check(rc, "kad_ir_solve")
        return x, (iters[0], iters[1]), resid.value, rc == 0

    # ---- spectral transforms ------------------------------------------------
    def coef_3d(self, conf, base_r, base_t, base_p):
        """conf: (nbatch, np, nt, nr) contiguous (r fastest) -> coef (nbatch, np+2, nt, nr)."""
        nb, np_, nt, nr = conf.shape
        out = torch.empty((nb, np_ + 2, nt, nr), dtype=torch.float64, device=conf.device)
        br, pbr = _iarr(base_r)
        bt, pbt = _iarr(base_t)
        self._enter()
        check(lib.kad_coef_3d(self.h, nr, nt, np_, pbr, pbt, int(base_p), nb, _ptr(conf), _ptr(out)), "kad_coef_3d")
        self.sync()
        return out

    def coef_i_3d(self, coef, base_r, base_t, base_p):
        nb, npc, nt, nr = coef.shape
        out = torch.empty((nb, npc - 2, nt, nr), dtype=torch.float64, device=coef.device)
        br, pbr = _iarr(base_r)
        bt, pbt = _iarr(base_t)
        self._enter()
        check(lib.kad_coef_i_3d(self.h, nr, nt, npc - 2, pbr, pbt, int(base_p), nb, _ptr(coef), _ptr(out)),
              "kad_coef_i_3d")
        self.sync()
        return out

    def coef_1d(self, base, conf):
        L, n = conf.shape
        nout = n + 2 if base == 4 else n
        out = torch.empty((L, nout), dtype=torch.float64, device=conf.device)
        self._enter()
        check(lib.kad_coef_1d(self.h, base, n, L, _ptr(conf), _ptr(out)), "kad_coef_1d")
        self.sync()
        return out

    def coef_i_1d(self, base, coef, nconf):
        L, _ = coef.shape
        out = torch.empty((L, nconf), dtype=torch.float64, device=coef.device)
        self._enter()
        check(lib.kad_coef_i_1d(self.h, base, nconf, L, _ptr(coef), _ptr(out)), "kad_coef_i_1d")
        self.sync()
        return out


def host_transform_matrix(kind, base, nconf):
    """Analysis (kind 0) or synthesis (kind 1) matrix of a 1-D basis, from the C++ host tables."""
    nc = nconf + 2 if base == 4 else nconf
    buf = (ctypes.c_double * (nc * nconf))()
    n = check(lib.kad_host_transform_matrix(kind, base, nconf, buf, nc * nconf), "kad_host_transform_matrix")
    a = np.frombuffer(buf, dtype=np.float64, count=n).copy()
    return a.reshape((nc, nconf) if kind == 0 else (nconf, nc))


def host_op1d_matrix(op, base, n):
    buf = (ctypes.c_double * (n * n))()
    ob = ctypes.c_int()
    check(lib.kad_host_op1d_matrix(op, base, n, buf, n * n, ctypes.byref(ob)), "kad_host_op1d_matrix")
    return np.frombuffer(buf, dtype=np.float64).copy().reshape(n, n), ob.value


class DomainSolve:
    """Domain-decomposed solve of J x = b (kad_dd_*): blocks = [(rows, cols), ...], the
    rows and unknowns of J in each (square) block; J column-major fp64 on the device."""

    def __init__(self, ctx, n, blocks):
        self.ctx = ctx
        sizes = _iarr([len(r) for r, _ in blocks])
        rows = _iarr(np.concatenate([np.asarray(r) for r, _ in blocks]))
        cols = _iarr(np.concatenate([np.asarray(c) for _, c in blocks]))
        h = ctypes.c_void_p()
        ctx._enter()
        check(lib.kad_dd_create(ctx.h, n, len(blocks), sizes[1], rows[1], cols[1], ctypes.byref(h)), "kad_dd_create")
        self.h = h
        self.n = n

    def factor(self, J):
        self.ctx._enter()
        rc = lib.kad_dd_factor(self.h, _ptr(J), J.stride(1))
        check(rc, "kad_dd_factor")
        return rc

    def factor_static(self, J, max_growth=8.0):
        """kad_dd_factor_static: 0 factored with the pivot orders of the last factor(),
        > 0 when they do not apply (then call factor)."""
        self.ctx._enter()
        rc = lib.kad_dd_factor_static(self.h, _ptr(J), J.stride(1), float(max_growth))
        check(rc, "kad_dd_factor_static")
        return rc

    def factor_keep(self, J, max_growth=1e3):
        """kad_dd_factor_keep: partial pivoting in the blocks, the capacitance matrix in
        the last factorisation's order while its multipliers stay within max_growth."""
        self.ctx._enter()
        rc = lib.kad_dd_factor_keep(self.h, _ptr(J), J.stride(1), float(max_growth))
        check(rc, "kad_dd_factor_keep")
        return rc

    def pivoting(self):
        """(the last factorisation took the kept orders, its largest |multiplier|)"""
        st = ctypes.c_int()
        g = ctypes.c_double()
        check(lib.kad_dd_pivoting(self.h, ctypes.byref(st), ctypes.byref(g)), "kad_dd_pivoting")
        return bool(st.value), g.value

    def solve(self, J, b, tol=1e-13, max_outer=8):
        x = torch.empty_like(b)
        it = (ctypes.c_int * 2)()
        res = ctypes.c_double()
        self.ctx._enter()
        rc = lib.kad_dd_solve(self.h, _ptr(J), J.stride(1), _ptr(b), _ptr(x), tol, max_outer, it, ctypes.byref(res))
        if rc != -4:
            check(rc, "kad_dd_solve")
        return x, (it[0], it[1]), res.value

    def apply(self, v):
        self.ctx._enter()
        check(lib.kad_dd_apply(self.h, _ptr(v)), "kad_dd_apply")
        return v

    def info(self):
        sz = (ctypes.c_int * 4)()
        tm = (ctypes.c_double * 5)()
        check(lib.kad_dd_info(self.h, sz, tm), "kad_dd_info")
        return dict(nd=sz[0], coupling_rows=sz[1], kmax=sz[2], blocks=sz[3], times=list(tm))

    def close(self):
        if self.h:
            lib.kad_dd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
