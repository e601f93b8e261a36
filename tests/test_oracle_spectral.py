"""CPU: pin the oracle's 1-D/3-D transforms and the product's host tables.

Pins (what a plausible mistake would break):
  * coef_i of a unit coefficient vector == the closed-form basis function sampled
    on the reference's grid (wrong sign convention, grid or index fails here);
  * coef of a sampled basis function == the unit vector (catches a dropped factor 2);
  * CHEB / COS_EVEN coefficients == scipy's DCT-I (independent textbook routine);
  * the product's host matrices (C++ restatement) == the oracle (numpy FFT restatement).
"""
import numpy as np
import pytest
import scipy.fft

from oracle import spectral as S
from kadath_frontera_amd.device import host_transform_matrix, host_op1d_matrix

BASES_THETA = [S.COS_EVEN, S.COS_ODD, S.SIN_EVEN, S.SIN_ODD]


def closed_form(base, n, j):
    """Basis function j of `base` on the reference grid (domain_shell.cpp:252-267)."""
    if base == S.CHEB:
        x = -np.cos(np.pi * np.arange(n) / (n - 1))
        return np.cos(j * np.arccos(np.clip(x, -1, 1)))
    if base == S.COSSIN:
        ph = 2 * np.pi * np.arange(n) / n
        m = j // 2
        return np.sin(m * ph) if j % 2 else np.cos(m * ph)
    th = np.pi / 2 * np.arange(n) / (n - 1)
    return {S.COS_EVEN: np.cos(2 * j * th), S.COS_ODD: np.cos((2 * j + 1) * th),
            S.SIN_EVEN: np.sin(2 * j * th), S.SIN_ODD: np.sin((2 * j + 1) * th)}[base]


def valid_slots(base, n):
    nc = n + 2 if base == S.COSSIN else n
    if base == S.COSSIN:
        return [j for j in range(nc) if j not in (1, nc - 1)]
    if base in (S.COS_ODD, S.SIN_ODD):
        return list(range(nc - 1))
    if base == S.SIN_EVEN:
        return list(range(1, nc - 1))
    return list(range(nc))


CASES = [(S.CHEB, n) for n in (5, 9, 13, 17, 25, 33)] + \
        [(b, n) for b in BASES_THETA for n in (5, 7, 9, 13, 17)] + \
        [(S.COSSIN, n) for n in (4, 6, 8, 12, 16, 32)]


@pytest.mark.parametrize("base,n", CASES)
def test_oracle_coef_i_unit_vectors_are_basis_functions(base, n):
    nc = n + 2 if base == S.COSSIN else n
    for j in valid_slots(base, n):
        e = np.zeros(nc)
        e[j] = 1.0
        np.testing.assert_allclose(S.coef_i_1d(base, e), closed_form(base, n, j), atol=2e-14)


@pytest.mark.parametrize("base,n", CASES)
def test_oracle_coef_of_basis_function_is_unit(base, n):
    nc = n + 2 if base == S.COSSIN else n
    for j in valid_slots(base, n):
        e = np.zeros(nc)
        e[j] = 1.0
        np.testing.assert_allclose(S.coef_1d(base, closed_form(base, n, j)), e, atol=5e-14)


@pytest.mark.parametrize("n", [5, 9, 17, 33])
def test_cheb_matches_scipy_dct1(n):
    rng = np.random.default_rng(n)
    f = rng.uniform(-1, 1, n)
    # f_i = sum_k c_k T_k(x_i), x_i = -cos(pi i/(n-1)): T_k(x_i) = (-1)^k cos(pi k i/(n-1))
    d = scipy.fft.dct(f, type=1) / (n - 1)
    d[0] /= 2
    d[-1] /= 2
    c_ref = d * (-1.0) ** np.arange(n)
    np.testing.assert_allclose(S.coef_1d(S.CHEB, f), c_ref, atol=1e-14)
    # COS_EVEN is the same DCT-I in psi = 2 theta, without the sign flip
    np.testing.assert_allclose(S.coef_1d(S.COS_EVEN, f), d, atol=1e-14)


@pytest.mark.parametrize("base,n", CASES)
def test_host_matrices_match_oracle(base, n):
    E = host_transform_matrix(1, base, n)
    C = host_transform_matrix(0, base, n)
    nc = E.shape[1]
    rng = np.random.default_rng(7)
    for _ in range(3):
        c = np.zeros(nc)
        vs = valid_slots(base, n)
        c[vs] = rng.uniform(-1, 1, len(vs))
        np.testing.assert_allclose(E @ c, S.coef_i_1d(base, c), atol=1e-13)
        v = S.coef_i_1d(base, c)
        np.testing.assert_allclose(C @ v, S.coef_1d(base, v), atol=1e-13)


def test_3d_roundtrip_oracle():
    rng = np.random.default_rng(1)
    nr, nt, npt = 9, 7, 4
    for anti in (False, True):
        br, bt, bp = S.std_base(nt, npt, anti)
        # random valid coefficient array: project through coef o coef_i
        conf = rng.uniform(-1, 1, (npt, nt, nr))
        cf = S.coef_3d(conf, br, bt, bp)
        v = S.coef_i_3d(cf, br, bt, bp)
        cf2 = S.coef_3d(v, br, bt, bp)
        np.testing.assert_allclose(cf2, cf, atol=1e-13)


def test_der_cheb_closed_form():
    # d/dx T_k sampled == coef_i(der_1d(e_k)) sampled
    n = 13
    M, ob = host_op1d_matrix(0, S.CHEB, n)
    assert ob == S.CHEB
    x = -np.cos(np.pi * np.arange(n) / (n - 1))
    E = host_transform_matrix(1, S.CHEB, n)
    for k in range(n):
        e = np.zeros(n)
        e[k] = 1
        dk = np.polynomial.chebyshev.chebder(e)
        ref = np.polynomial.chebyshev.chebval(x, dk) if len(dk) else np.zeros(n)
        np.testing.assert_allclose(E @ (M @ e), ref, atol=1e-10 * max(1, k * k))


def test_theta_ops_closed_form():
    # mult_sin / mult_cos / div_sin / der on COS_EVEN lines against sampled functions.
    n = 9
    th = np.pi / 2 * np.arange(n) / (n - 1)
    for op, f in [(1, np.sin), (2, np.cos)]:
        M, ob = host_op1d_matrix(op, S.COS_EVEN, n)
        E_in = host_transform_matrix(1, S.COS_EVEN, n)
        E_out = host_transform_matrix(1, ob, n)
        for j in range(n - 1):  # top mode is truncated by the reference
            e = np.zeros(n)
            e[j] = 1
            np.testing.assert_allclose(E_out @ (M @ e), f(th) * (E_in @ e), atol=1e-13)
    # der COS_EVEN -> SIN_EVEN: d/dth cos(2 j th) = -2 j sin(2 j th)
    M, ob = host_op1d_matrix(0, S.COS_EVEN, n)
    assert ob == S.SIN_EVEN
    E_out = host_transform_matrix(1, ob, n)
    for j in range(n - 1):
        e = np.zeros(n)
        e[j] = 1
        np.testing.assert_allclose(E_out @ (M @ e), -2 * j * np.sin(2 * j * th), atol=1e-12)
    # div_sin on SIN_ODD -> COS_EVEN: sin((2j+1)th)/sin(th)
    M, ob = host_op1d_matrix(3, S.SIN_ODD, n)
    assert ob == S.COS_EVEN
    E_out = host_transform_matrix(1, ob, n)
    for j in range(n - 1):
        e = np.zeros(n)
        e[j] = 1
        ref = np.where(th > 0, np.sin((2 * j + 1) * th) / np.where(th > 0, np.sin(th), 1), 2 * j + 1)
        np.testing.assert_allclose(E_out @ (M @ e), ref, atol=1e-12)
