"""GPU: kad_lu_solve_noalloc (triangular solves from the library's static per-GPU
scratch, no allocation) gives the same result as kad_lu_solve, bit for bit, on two
calls in a row -- the second with a smaller system -- for fp64 and fp32 factors and
both transposes; sizes beyond its maximum (n = 8192) are rejected before any device
work."""
import ctypes

import numpy as np
import pytest
import torch

from kadath_frontera_amd._lib import lib

pytestmark = pytest.mark.gpu


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("trans", [0, 1])
def test_noalloc_solve_matches_lu_solve(ctx, prec, trans):
    dt = torch.float64 if prec == "f64" else torch.float32
    code = 0 if prec == "f64" else 1
    rng = np.random.default_rng(17 + trans)
    for n in (3000, 1500):  # second call smaller: the static scratch keeps the first call's contents
        print(f"noalloc {prec} trans {trans} n {n}", flush=True)
        A = rng.standard_normal((n, n)) + n ** 0.5 * np.eye(n)
        dA = torch.tensor(np.ascontiguousarray(A.T), dtype=dt, device="cuda").T  # column-major
        ipiv, perm, info = ctx.lu_factor(prec, dA)
        assert info == 0
        b = torch.tensor(rng.standard_normal(n), dtype=torch.float64, device="cuda")
        x_ref = ctx.lu_solve(prec, dA, perm, b.clone(), trans)
        x = b.clone()
        ctx._enter()
        rc = lib.kad_lu_solve_noalloc(ctx.h, code, n, _ptr(dA), dA.stride(1), _ptr(perm), _ptr(x), trans)
        assert rc == 0
        ctx.sync()
        assert torch.equal(x, x_ref)
        # and it solves the system (fp64: to rounding; fp32 factors: to their accuracy)
        At = torch.tensor(A.T if trans else A, device="cuda")
        r = (At @ x - b).abs().max().item()
        assert r < (1e-9 if prec == "f64" else 1e-3) * b.abs().max().item() * n ** 0.5


def test_noalloc_rejects_oversize(ctx):
    dummy = torch.zeros(4, dtype=torch.float64, device="cuda")
    iperm = torch.zeros(4, dtype=torch.int32, device="cuda")
    rc = lib.kad_lu_solve_noalloc(ctx.h, 0, 8193, _ptr(dummy), 8193, _ptr(iperm), _ptr(dummy), 0)
    assert rc == -1
