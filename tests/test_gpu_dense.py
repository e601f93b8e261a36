"""GPU: MFMA GEMM, LU with partial pivoting, triangular solves and mixed-precision
refinement through the C-ABI, against LAPACK (scipy) and float64 references.

Tolerances: f64 paths 1e-12 relative (backward-stable LU: ||PA - LU|| / ||A|| ~ n eps);
f32 GEMM 1e-5 relative; bf16 GEMM 1e-2 relative (8-bit mantissa operands, fp32
accumulation); refined solves reach ||b - Ax||_inf <= 1e-13 ||b||_inf.
"""
import numpy as np
import pytest
import scipy.linalg
import torch

pytestmark = pytest.mark.gpu


def colmajor(a_np, dtype, dev="cuda"):
    """Device tensor holding a_np in column-major order; returns the (m, n) view."""
    t = torch.tensor(np.ascontiguousarray(a_np.T), dtype=dtype, device=dev)
    return t.T


@pytest.mark.parametrize("prec,tol", [("f64", 1e-13), ("f32", 1e-5), ("bf16", 2e-2)])
@pytest.mark.parametrize("m,n,k", [(128, 128, 64), (300, 257, 129), (1, 5, 3), (517, 200, 512)])
def test_gemm(ctx, prec, tol, m, n, k):
    print(f"gemm {prec} {m}x{n}x{k}", flush=True)
    rng = np.random.default_rng(m * 7 + n + k)
    A = rng.standard_normal((m, k))
    B = rng.standard_normal((k, n))
    C = rng.standard_normal((m, n))
    dt = torch.float64 if prec == "f64" else torch.float32
    if prec == "bf16":
        # reference with operands rounded to bf16, fp64 accumulation
        Ar = torch.tensor(A).to(torch.bfloat16).double().numpy()
        Br = torch.tensor(B).to(torch.bfloat16).double().numpy()
    elif prec == "f32":
        Ar, Br = A.astype(np.float32).astype(np.float64), B.astype(np.float32).astype(np.float64)
    else:
        Ar, Br = A, B
    for op in (0, 1):
        dA, dB, dC = colmajor(A, dt), colmajor(B, dt), colmajor(C, dt)
        ctx.gemm(prec, op, dA, dB, dC)
        ctx.sync()
        got = dC.double().cpu().numpy()
        ref = (C - Ar @ Br) if op == 0 else Ar @ Br
        scale = np.abs(Ar) @ np.abs(Br) + (np.abs(C) if op == 0 else 0)
        err = np.max(np.abs(got - ref) / scale)
        assert err < tol, (prec, op, err)


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("m", [64, 600, 4096])
def test_gemm_update_rounds_once(ctx, prec, m):
    # C -= A B with A's first column all ones and its other columns zero: every
    # product is exact, so each entry must be C - B[0, j] rounded once -- the same bits
    # as numpy's subtraction, denormal operands and results included (the update is a
    # no-return atomic add in L2; 600 and 4096 give the 64- and 128-row tiles, edges too)
    print(f"gemm update rounding {prec} m={m}", flush=True)
    npt = np.float64 if prec == "f64" else np.float32
    dt = torch.float64 if prec == "f64" else torch.float32
    tiny = np.finfo(npt).tiny
    rng = np.random.default_rng(m)
    k = 4
    A = np.zeros((m, k), npt)
    A[:, 0] = 1.0
    B = np.zeros((k, m), npt)
    B[0] = rng.standard_normal(m).astype(npt)
    B[0, ::7] = (tiny * rng.uniform(0.5, 3.0, B[0, ::7].size)).astype(npt)
    C = rng.standard_normal((m, m)).astype(npt)
    C[::5] = (tiny * rng.uniform(0.25, 4.0, C[::5].shape)).astype(npt)  # denormal and just-normal rows
    C[1::11, ::7] = 0.0
    dA, dB, dC = colmajor(A, dt), colmajor(B, dt), colmajor(C, dt)
    ctx.gemm(prec, 0, dA, dB, dC)
    ctx.sync()
    got = dC.cpu().numpy()
    want = C - B[0][None, :]
    assert np.array_equal(got.view(np.uint64 if prec == "f64" else np.uint32),
                          want.view(np.uint64 if prec == "f64" else np.uint32))
    assert (np.abs(want[want != 0]) < tiny).any()  # the case includes denormal results


def _spd_like(n, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((n, n)) + 0.5 * np.sqrt(n) * np.eye(n)


def _block_tridiagonal(nblk, m, seed, coupling_rows=24):
    """Block-tridiagonal matrix shaped like a domain-major Kadath Jacobian: dense
    diagonal blocks, off-diagonal blocks nonzero only in a few (matching) rows."""
    rng = np.random.default_rng(seed)
    n = nblk * m
    A = np.zeros((n, n))
    for b in range(nblk):
        s = slice(b * m, (b + 1) * m)
        A[s, s] = rng.standard_normal((m, m)) + 0.5 * np.sqrt(m) * np.eye(m)
        for nb_ in (b - 1, b + 1):
            if 0 <= nb_ < nblk:
                rows = rng.choice(m, coupling_rows, replace=False) + b * m
                A[rows, nb_ * m:(nb_ + 1) * m] = rng.standard_normal((coupling_rows, m))
    return A


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_lu_block_sparse_matches_lapack(ctx, prec):
    """Tile skipping (structural-zero masks) must not change the factorisation:
    same pivots as LAPACK and the same factors."""
    print(f"lu block-sparse {prec}", flush=True)
    A = _block_tridiagonal(5, 700, 7)
    n = A.shape[0]
    dt = torch.float64 if prec == "f64" else torch.float32
    dA = colmajor(A, dt)
    ipiv, perm, info = ctx.lu_factor(prec, dA, nb=256)
    assert info == 0
    if prec == "f64":
        lu_ref, piv_ref = scipy.linalg.lu_factor(A)
        assert np.array_equal(ipiv.cpu().numpy(), piv_ref)
        np.testing.assert_allclose(dA.cpu().numpy(), lu_ref, atol=1e-10 * np.abs(lu_ref).max(), rtol=1e-9)
    else:
        # backward error within a small factor of LAPACK sgetrf's on the same matrix
        def bwd(lu, p):
            L = np.tril(lu, -1) + np.eye(n)
            return np.abs(A[p] - L @ np.triu(lu)).max() / np.abs(A).max()
        got = dA.cpu().numpy().astype(np.float64)
        lu32, piv32 = scipy.linalg.lu_factor(A.astype(np.float32))
        p32 = np.arange(n)
        for i, q in enumerate(piv32):
            p32[i], p32[q] = p32[q], p32[i]
        ref = bwd(lu32.astype(np.float64), p32)
        assert bwd(got, perm.cpu().numpy()) < 4 * ref + 1e-7, (bwd(got, perm.cpu().numpy()), ref)


@pytest.mark.parametrize("n,nb", [(64, 32), (257, 64), (1000, 256), (2100, 256)])
def test_lu_f64_matches_lapack(ctx, n, nb):
    print(f"lu f64 n={n} nb={nb}", flush=True)
    A = np.random.default_rng(n).standard_normal((n, n))
    dA = colmajor(A, torch.float64)
    ipiv, perm, info = ctx.lu_factor("f64", dA, nb=nb)
    assert info == 0
    lu_ref, piv_ref = scipy.linalg.lu_factor(A)
    got = dA.cpu().numpy()
    # same pivot sequence as LAPACK dgetrf (partial pivoting, first max)
    assert np.array_equal(ipiv.cpu().numpy(), piv_ref)
    np.testing.assert_allclose(got, lu_ref, atol=1e-10 * np.abs(lu_ref).max(), rtol=1e-9)
    # backward error ||P A - L U|| / ||A||
    L = np.tril(got, -1) + np.eye(n)
    U = np.triu(got)
    p = perm.cpu().numpy()
    be = np.abs(A[p] - L @ U).max() / np.abs(A).max()
    assert be < 1e-13 * n


def _per_column_factor(monkeypatch, prec, A):
    """Factor A with the per-column base-panel kernels (a context created with
    KAD_LU_PANEL=percolumn) instead of the persistent one."""
    from kadath_frontera_amd.device import Context
    monkeypatch.setenv("KAD_LU_PANEL", "percolumn")
    c = Context(0)
    try:
        ipiv, perm, info = c.lu_factor(prec, A)
        c.sync()
    finally:
        c.close()
        monkeypatch.delenv("KAD_LU_PANEL")
    return ipiv, info


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_persistent_panel_ties_match_per_column(ctx, prec, monkeypatch):
    """The persistent base panel (one launch per 32 columns, one candidate exchange
    per column, rows held in registers by 12 workgroups here) performs the per-column
    kernels' operations in the same order: pivots and factors agree bit for bit.  Integer
    entries make exact ties in |a|, where the choice must be the first index (the
    first column's pivot is checked against argmax directly)."""
    print(f"persistent panel ties {prec}", flush=True)
    n = 3000
    A = np.random.default_rng(11).integers(-2, 3, (n, n)).astype(np.float64)
    dt = torch.float64 if prec == "f64" else torch.float32
    d1, d2 = colmajor(A, dt), colmajor(A, dt)
    ipiv1, _, info1 = ctx.lu_factor(prec, d1)
    ctx.sync()
    ipiv2, info2 = _per_column_factor(monkeypatch, prec, d2)
    assert info1 == info2 == 0
    assert int(ipiv1[0]) == int(np.argmax(np.abs(A[:, 0])))
    assert torch.equal(ipiv1, ipiv2)
    assert torch.equal(d1, d2)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_persistent_panel_large_grid_matches_per_column(ctx, prec, monkeypatch):
    """Dense n = 40000: the first panels are taller than the persistent panel takes
    (64 workgroups of 256 rows) and run on the per-column kernels; from 16384 rows down
    the persistent panel runs with up to 64 workgroups exchanging candidates.  Bitwise
    agreement with the per-column kernels throughout."""
    print(f"persistent panel n=40000 {prec}", flush=True)
    n = 40000
    dt = torch.float64 if prec == "f64" else torch.float32
    g = torch.Generator(device="cuda").manual_seed(n)
    d1 = torch.randn(n, n, generator=g, dtype=dt, device="cuda").T  # column-major view
    d2 = torch.empty_like(d1)
    d2.copy_(d1)
    assert d2.stride() == d1.stride() == (1, n)
    ipiv1, _, info1 = ctx.lu_factor(prec, d1)
    ctx.sync()
    ipiv2, info2 = _per_column_factor(monkeypatch, prec, d2)
    assert info1 == info2 == 0
    assert torch.equal(ipiv1, ipiv2)
    assert torch.equal(d1, d2)
    del d1, d2
    torch.cuda.empty_cache()


def test_persistent_panel_abandoned_barrier_reports_error(ctx, monkeypatch):
    """The persistent panel's waits are bounded: with a spin limit of 0 (set from the
    host, KAD_LU_SPIN) its waits for the other workgroups' candidates are abandoned at
    once, every workgroup exits, and kad_lu_factor reports an error instead of waiting.
    A context with the default limit then factors the same matrix with LAPACK's
    pivots."""
    from kadath_frontera_amd.device import Context
    from kadath_frontera_amd._lib import KadError
    print("persistent panel abandoned barrier", flush=True)
    n = 3000
    A = np.random.default_rng(12).standard_normal((n, n))
    monkeypatch.setenv("KAD_LU_SPIN", "0")
    c = Context(0)
    try:
        with pytest.raises(KadError):
            c.lu_factor("f64", colmajor(A, torch.float64))
        c.sync()
    finally:
        c.close()
        monkeypatch.delenv("KAD_LU_SPIN")
    dA = colmajor(A, torch.float64)
    ipiv, perm, info = ctx.lu_factor("f64", dA)
    assert info == 0
    _, piv_ref = scipy.linalg.lu_factor(A)
    assert np.array_equal(ipiv.cpu().numpy(), piv_ref)


def test_persistent_panel_back_to_back_factorisations(ctx):
    """The persistent panel's exchange slots are reused by every factorisation, each
    word flagged with a launch sequence that must never repeat while a stale word
    with that flag is still there.  Different matrices factored back to back on one
    context (and the first again) each give LAPACK's pivots and a backward-stable
    factorisation."""
    print("persistent panel back-to-back factorisations", flush=True)
    n = 2600  # 11 workgroups for the first panels, a partial last panel
    mats = [np.random.default_rng(40 + q).standard_normal((n, n)) for q in range(3)]
    for q in (0, 1, 2, 0):
        A = mats[q]
        dA = colmajor(A, torch.float64)
        ipiv, perm, info = ctx.lu_factor("f64", dA)
        assert info == 0
        lu_ref, piv_ref = scipy.linalg.lu_factor(A)
        assert np.array_equal(ipiv.cpu().numpy(), piv_ref), q
        np.testing.assert_allclose(dA.cpu().numpy(), lu_ref, rtol=0, atol=1e-9 * np.abs(lu_ref).max())


def test_persistent_panel_interleaved_precisions(ctx):
    """FP64 and FP32 factorisations share one context's exchange slots and launch
    sequence.  Interleaved, the FP64 ones keep LAPACK's pivots and the two FP32
    factorisations of the same matrix agree bit for bit."""
    print("persistent panel interleaved precisions", flush=True)
    n = 2100
    A64 = np.random.default_rng(90).standard_normal((n, n))
    A32 = np.random.default_rng(91).standard_normal((n, n))
    _, piv_ref = scipy.linalg.lu_factor(A64)
    out32 = []
    for q in range(2):
        d64 = colmajor(A64, torch.float64)
        ipiv, _, info = ctx.lu_factor("f64", d64)
        assert info == 0 and np.array_equal(ipiv.cpu().numpy(), piv_ref), q
        d32 = colmajor(A32, torch.float32)
        ipiv32, _, info32 = ctx.lu_factor("f32", d32)
        assert info32 == 0
        out32.append((ipiv32.cpu().clone(), d32.cpu().clone()))
    assert torch.equal(out32[0][0], out32[1][0])
    assert torch.equal(out32[0][1], out32[1][1])


def test_lu_block_size_override(ctx, monkeypatch):
    """KAD_LU_NB (read when the block size is left at 0) changes the blocking, not the
    pivoting: LAPACK's pivots at NB = 128 and 512."""
    print("lu block size override", flush=True)
    n = 1500
    A = np.random.default_rng(77).standard_normal((n, n))
    _, piv_ref = scipy.linalg.lu_factor(A)
    for nb in ("128", "512"):
        monkeypatch.setenv("KAD_LU_NB", nb)
        dA = colmajor(A, torch.float64)
        ipiv, _, info = ctx.lu_factor("f64", dA)
        assert info == 0
        assert np.array_equal(ipiv.cpu().numpy(), piv_ref), nb
    monkeypatch.delenv("KAD_LU_NB")


@pytest.mark.parametrize("trans", [0, 1])
def test_lu_solve_f64(ctx, trans):
    n = 777
    A = _spd_like(n, 3)
    b = np.random.default_rng(4).standard_normal(n)
    dA = colmajor(A, torch.float64)
    ipiv, perm, info = ctx.lu_factor("f64", dA)
    x = ctx.lu_solve("f64", dA, perm, torch.tensor(b, device="cuda"), trans=trans).cpu().numpy()
    ref = np.linalg.solve(A.T if trans else A, b)
    np.testing.assert_allclose(x, ref, rtol=1e-10, atol=1e-12)


def test_singular_reports_info(ctx):
    n = 96
    A = np.random.default_rng(5).standard_normal((n, n))
    A[:, 40] = 0.0  # exactly singular: zero column
    dA = colmajor(A, torch.float64)
    ipiv, perm, info = ctx.lu_factor("f64", dA, nb=32)
    assert info == 41


def test_gemv64(ctx):
    n = 1500
    rng = np.random.default_rng(6)
    A = rng.standard_normal((n, n))
    x = rng.standard_normal(n)
    b = rng.standard_normal(n)
    dA = colmajor(A, torch.float64)
    y = ctx.gemv64(dA, torch.tensor(x, device="cuda"), torch.tensor(b, device="cuda")).cpu().numpy()
    np.testing.assert_allclose(y, b - A @ x, rtol=1e-12, atol=1e-11)
    yt = ctx.gemv64(dA, torch.tensor(x, device="cuda"), None, trans=1).cpu().numpy()
    np.testing.assert_allclose(yt, A.T @ x, rtol=1e-12, atol=1e-11)


@pytest.mark.parametrize("prec,mode", [("f32", 0), ("f32", 1), ("bf16", 1)])
@pytest.mark.parametrize("trans", [0, 1])
def test_mixed_precision_refinement(ctx, prec, mode, trans):
    n = 1200
    print(f"ir {prec} mode={mode} trans={trans}", flush=True)
    A = _spd_like(n, 8)
    b = np.random.default_rng(9).standard_normal(n)
    dA64 = colmajor(A, torch.float64)
    dlo = colmajor(A, torch.float32)
    ipiv, perm, info = ctx.lu_factor(prec, dlo)
    assert info == 0
    x, iters, resid, ok = ctx.ir_solve(dA64, prec, dlo, perm, torch.tensor(b, device="cuda"), tol=1e-13,
                                       max_outer=30, mode=mode, trans=trans)
    assert ok, (iters, resid)
    ref = np.linalg.solve(A.T if trans else A, b)
    np.testing.assert_allclose(x.cpu().numpy(), ref, rtol=1e-9, atol=1e-11)
