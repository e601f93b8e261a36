"""CPU: the oracle's Cartesian derivatives and Laplacian (Domain::partial_cart via
do_der_abs_from_der_var, domain_shell.cpp:665-683, domain_compact.cpp:674-693) on
fields whose derivatives are known in closed form, in shells and in the compactified
domain; and tutorials/tuto5.cpp's known answer through the oracle's tau machinery.

Exactly representable fields (x, y, z, r^2 in a shell; 1/r and x/r^3 in the
compactified domain, polynomials of the radial variable there) must come out to
rounding; 1/r in a shell to spectral accuracy.  A swapped sin/cos factor, a missing
1/r or 1/alpha, or the wrong sign of the compactified map fails these.
"""
import numpy as np
import pytest

from oracle.kadath import Domain, VD, vadd, vsub
from oracle.tuto5 import Tuto5


def _shell(res=(9, 7, 4), rint=1.0, rext=2.0):
    return Domain(1, "shell", res, rint, rext)


def _compact(res=(9, 7, 4), rint=4.0):
    return Domain(3, "compact", res, rint)


def _xyz(dom):
    st, ct, cp, sp = np.sin(dom.TH), np.cos(dom.TH), np.cos(dom.PH), np.sin(dom.PH)
    return dom.R * st * cp, dom.R * st * sp, dom.R * ct


def _conf(v):
    return v.conf() if not v.zero else 0.0


def test_gradient_of_coordinates_in_a_shell():
    dom = _shell()
    x, y, z = _xyz(dom)
    for comp, (vals, anti) in enumerate([(x, False), (y, False), (z, True)]):
        g = dom.der_abs(dom.conf_field(vals, anti=anti))
        for i in range(3):
            np.testing.assert_allclose(_conf(g[i]), 1.0 if i == comp else 0.0, atol=1e-12)


def test_gradient_and_laplacian_of_r2_in_a_shell():
    dom = _shell()
    x, y, z = _xyz(dom)
    f = dom.conf_field(dom.R ** 2)
    g = dom.der_abs(f)
    for i, c in enumerate((x, y, z)):
        np.testing.assert_allclose(_conf(g[i]), 2 * c, atol=1e-11)
    lap = vadd(vadd(dom.der_abs(g[0])[0], dom.der_abs(g[1])[1]), dom.der_abs(g[2])[2])
    np.testing.assert_allclose(_conf(lap), 6.0, atol=1e-10)


@pytest.mark.parametrize("kind", ["shell", "compact"])
def test_gradient_and_laplacian_of_inverse_r(kind):
    dom = _shell((25, 7, 4)) if kind == "shell" else _compact()
    with np.errstate(divide="ignore", invalid="ignore"):
        u = np.where(np.isfinite(dom.R), 1.0 / dom.R, 0.0)
        x, y, z = _xyz(dom)
        # -x_i / r^3 = -(unit vector)_i u^2, finite at infinity
        st, ct, cp, sp = np.sin(dom.TH), np.cos(dom.TH), np.cos(dom.PH), np.sin(dom.PH)
        ref = [-st * cp * u * u, -st * sp * u * u, -ct * u * u]
    f = dom.conf_field(u)
    g = dom.der_abs(f)
    tol = 1e-12 if kind == "compact" else 1e-10
    for i in range(3):
        np.testing.assert_allclose(_conf(g[i]), ref[i], atol=tol)
    lap = vadd(vadd(dom.der_abs(g[0])[0], dom.der_abs(g[1])[1]), dom.der_abs(g[2])[2])
    np.testing.assert_allclose(_conf(lap), 0.0, atol=100 * tol)


def test_normal_derivative_of_inverse_r():
    # dn = d/dr: -1/r^2 in a shell and in the compactified domain
    for dom in (_shell((25, 5, 4)), _compact((9, 5, 4))):
        with np.errstate(divide="ignore"):
            u = np.where(np.isfinite(dom.R), 1.0 / dom.R, 0.0)
        dn = dom.der_normal(dom.conf_field(u), 1)
        np.testing.assert_allclose(_conf(dn), -u * u, atol=1e-10)


def test_tuto5_known_answer():
    t = Tuto5()
    lay, n = t.unknown_layout()
    assert n == t.residual().size
    err0 = t.newton_step()
    assert err0 > 1e-3
    assert np.abs(t.residual()).max() < 1e-11
    for d in range(1, t.ndom):
        dom = t.sp.doms[d]
        with np.errstate(divide="ignore"):
            exact = np.where(np.isfinite(dom.R), 1.0 + t.a / dom.R, 1.0)
        # the discrete solution differs from 1 + a/r by the radial truncation error of
        # 13 Chebyshev points (4.0e-9 in the first shell)
        assert np.abs(t.F[d].conf() - exact).max() < 1e-8, d


def test_tuto5_planted_error_fails():
    # a wrong Robin coefficient moves the solution away from 1 + a/r
    t = Tuto5()
    t.a *= 1.5
    t.newton_step()
    dom = t.sp.doms[1]
    exact = 1.0 + 1.0 / dom.R
    assert np.abs(t.F[1].conf() - exact).max() > 1e-3
