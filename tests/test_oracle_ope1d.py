"""CPU: every 1-D coefficient-space operator of src/Ope_1d, on every basis it accepts,
against the closed form it must equal -- for the oracle's restatement
(oracle/kadath.py) and for the library's operator matrices (kad_host_op1d_matrix),
which also agree with each other.

Basis functions: CHEB T_j(x); COS_EVEN cos(2j th), COS_ODD cos((2j+1) th), SIN_EVEN
sin(2j th), SIN_ODD sin((2j+1) th) on th in [0, pi/2]; COSSIN [1, 0, cos phi,
sin phi, cos 2phi, ...] (coef_1d.cpp:254-292).  Closed forms:
  der_1d      d/dx f                       (der_1d.cpp)
  mult_sin_1d sin(.) f,  mult_cos_1d cos(.) f
  div_sin_1d  f / sin th (SIN_*), (f - f(0)) / sin th (COS_EVEN),
              (f - f(0) cos th) / sin th (COS_ODD)     (div_sin_1d.cpp:75-134)
  mult_x_1d   x f,  mult_xm1_1d (x - 1) f,  div_xm1_1d (f - f(1)) / (x - 1)
The top mode of a line is outside the closed form where the reference truncates the
result to the line's length (a product raising the degree past the last
coefficient): each case below names the modes it excludes.  A wrong recurrence index,
a sign or a dropped factor 1/2 on the first coefficient fails these.
"""
import numpy as np
import pytest

from oracle import kadath as K
from oracle import spectral as S
from kadath_frontera_amd.device import host_op1d_matrix

CHEB, COSSIN = S.CHEB, S.COSSIN
CE, CO, SE, SO = S.COS_EVEN, S.COS_ODD, S.SIN_EVEN, S.SIN_ODD
OPS = {"der": (K.der_1d, 0), "mult_sin": (K.mult_sin_1d, 1), "mult_cos": (K.mult_cos_1d, 2),
       "div_sin": (K.div_sin_1d, 3), "mult_x": (K.mult_x_1d, 4), "mult_xm1": (K.mult_xm1_1d, 5),
       "div_xm1": (K.div_xm1_1d, 6)}
# (operator, basis) -> output basis, per the reference's switch statements
CASES = {
    ("der", CHEB): CHEB, ("der", CE): SE, ("der", CO): SO, ("der", SE): CE, ("der", SO): CO, ("der", COSSIN): COSSIN,
    ("mult_sin", CE): SO, ("mult_sin", CO): SE, ("mult_sin", SE): CO, ("mult_sin", SO): CE,
    ("mult_sin", COSSIN): COSSIN,
    ("mult_cos", CE): CO, ("mult_cos", CO): CE, ("mult_cos", SE): SO, ("mult_cos", SO): SE,
    ("mult_cos", COSSIN): COSSIN,
    ("div_sin", CE): SO, ("div_sin", CO): SE, ("div_sin", SE): CO, ("div_sin", SO): CE,
    ("mult_x", CHEB): CHEB, ("mult_xm1", CHEB): CHEB, ("div_xm1", CHEB): CHEB,
}


def basis(b, j, t):
    if b == CHEB:
        return np.polynomial.chebyshev.chebval(t, np.eye(j + 1)[j])
    if b == CE:
        return np.cos(2 * j * t)
    if b == CO:
        return np.cos((2 * j + 1) * t)
    if b == SE:
        return np.sin(2 * j * t)
    if b == SO:
        return np.sin((2 * j + 1) * t)
    if j == 0:
        return np.ones_like(t)
    if j == 1:
        return np.zeros_like(t)
    return np.cos((j // 2) * t) if j % 2 == 0 else np.sin((j // 2) * t)


def dbasis(b, j, t):
    if b == CHEB:
        return np.polynomial.chebyshev.chebval(t, np.polynomial.chebyshev.chebder(np.eye(j + 1)[j])) if j else 0 * t
    if b == CE:
        return -2 * j * np.sin(2 * j * t)
    if b == CO:
        return -(2 * j + 1) * np.sin((2 * j + 1) * t)
    if b == SE:
        return 2 * j * np.cos(2 * j * t)
    if b == SO:
        return (2 * j + 1) * np.cos((2 * j + 1) * t)
    if j < 2:
        return 0 * t
    k = j // 2
    return -k * np.sin(k * t) if j % 2 == 0 else k * np.cos(k * t)


def series(b, c, t):
    return sum(c[j] * basis(b, j, t) for j in range(len(c)))


def closed_form(op, b, j, t):
    f = basis(b, j, t)
    if op == "der":
        return dbasis(b, j, t)
    if op == "mult_sin":
        return np.sin(t) * f
    if op == "mult_cos":
        return np.cos(t) * f
    if op == "div_sin":
        if b == CE:
            return (f - 1.0) / np.sin(t)
        if b == CO:
            return (f - np.cos(t)) / np.sin(t)
        return f / np.sin(t)
    if op == "mult_x":
        return t * f
    if op == "mult_xm1":
        return (t - 1) * f
    return (f - 1.0) / (t - 1)  # div_xm1: T_j(1) = 1


def excluded(op, b, n):
    """Top modes whose exact result needs a coefficient past the end of the line."""
    if b == COSSIN:
        # the Nyquist pair cos(n/2 phi) (index n - 2) and sin (n - 1, zero on the grid);
        # products with sin / cos phi also lift the mode below it
        if op == "der":
            return {n - 2, n - 1}
        return {n - 4, n - 3, n - 2, n - 1} if op in ("mult_sin", "mult_cos") else set()
    if op == "div_sin" and b in (SE, SO):
        return {n - 1} if b == SO else set()
    if op in ("der", "mult_sin", "mult_cos", "mult_x", "mult_xm1", "div_sin"):
        return {n - 1}
    return set()


def points(b, rng):
    if b == CHEB:
        return rng.uniform(-1, 0.99, 64)
    if b == COSSIN:
        return rng.uniform(0, 2 * np.pi, 64)
    return rng.uniform(0.02, np.pi / 2, 64)


@pytest.mark.parametrize("case", sorted(CASES, key=str))
@pytest.mark.parametrize("impl", ["oracle", "library"])
def test_ope1d_closed_form(case, impl):
    op, b = case
    n = 14 if b == COSSIN else 9
    fn, code = OPS[op]
    rng = np.random.default_rng(hash(case) % 2**32)
    t = points(b, rng)
    if impl == "library":
        M, ob = host_op1d_matrix(code, b, n)
    else:
        M, ob = K.op_mat(fn, b, n)
    assert ob == CASES[case]
    skip = excluded(op, b, n)
    checked = 0
    for j in range(n):
        if j in skip or (b == COSSIN and j == 1):
            continue
        got = series(ob, M[:, j], t)
        ref = closed_form(op, b, j, t)
        tol = 1e-12 * max(1.0, float(np.abs(ref).max()))
        assert np.abs(got - ref).max() <= tol, (op, b, j, np.abs(got - ref).max())
        checked += 1
    assert checked >= n - 5


@pytest.mark.parametrize("case", sorted(CASES, key=str))
def test_library_matrices_equal_oracle(case):
    op, b = case
    n = 14 if b == COSSIN else 9
    fn, code = OPS[op]
    Ml, obl = host_op1d_matrix(code, b, n)
    Mo, obo = K.op_mat(fn, b, n)
    assert obl == obo
    np.testing.assert_allclose(Ml, Mo, rtol=0, atol=1e-15)
