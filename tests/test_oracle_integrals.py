"""Oracle quadrature rules and closed-form potential integrals."""

import numpy as np
import pytest

from oracle.integrals_ref import TRI3_BARY, TRI3_W, TRI7_BARY, TRI7_W, points, wilton


def test_potential_integrals_vs_brute_force():
    """Mirrors Engine/tests/test_mom25d.py:36-50."""
    V = np.array([[[0, 0], [1.0, 0.2], [0.3, 0.9]]])
    r = np.array([[0.4, 0.3], [2.0, -0.5], [0.5, 0.1]])
    I0, I1 = wilton(V, r)
    n = 1200
    u, v = np.meshgrid((np.arange(n) + 0.5) / n, (np.arange(n) + 0.5) / n)
    m = u + v < 1
    P = V[0, 0] + u[m, None] * (V[0, 1] - V[0, 0]) + v[m, None] * (V[0, 2] - V[0, 0])
    e1, e2 = V[0, 1] - V[0, 0], V[0, 2] - V[0, 0]
    w = 0.5 * abs(e1[0] * e2[1] - e1[1] * e2[0]) / m.sum()
    for k in range(len(r)):
        d = P - r[k]
        Rk = np.hypot(d[:, 0], d[:, 1])
        assert I0[0, k] == pytest.approx((w / Rk).sum(), rel=2e-3)
        np.testing.assert_allclose(I1[0, k], (w * d / Rk[:, None]).sum(0), atol=1e-3)


@pytest.mark.parametrize("d", [0.05, 0.3])
def test_potential_integrals_off_plane(d):
    """Observer above the plane: the integrand is smooth, so a collapsed
    (Duffy) tensor Gauss-Legendre rule converges to many digits."""
    V = np.array([[[0, 0], [1.0, 0.2], [0.3, 0.9]]])
    r = np.array([[0.4, 0.3], [1.5, -0.2], [0.0, 0.0]])
    I0, I1 = wilton(V, r, d)
    x, w = np.polynomial.legendre.leggauss(400)
    x, w = 0.5 * (x + 1), 0.5 * w
    U, Vv = np.meshgrid(x, x, indexing="ij")
    Wt = np.outer(w, w) * U
    a, b, c = V[0]
    P = a + U[..., None] * (b - a) + (U * Vv)[..., None] * (c - b)
    jac = abs((b - a)[0] * (c - b)[1] - (b - a)[1] * (c - b)[0])
    for k in range(len(r)):
        dd = P - r[k]
        Rk = np.sqrt(dd[..., 0] ** 2 + dd[..., 1] ** 2 + d * d)
        assert I0[0, k] == pytest.approx(jac * np.sum(Wt / Rk), rel=1e-9)
        np.testing.assert_allclose(I1[0, k], jac * np.einsum("ij,ijd->d", Wt / Rk, dd), rtol=1e-8, atol=1e-12)


def _monomial_integral(p, q):
    """int over the unit right triangle of x^p y^q = p! q! / (p+q+2)!"""
    from math import factorial

    return factorial(p) * factorial(q) / factorial(p + q + 2)


@pytest.mark.parametrize("bary,w,deg", [(TRI7_BARY, TRI7_W, 5), (TRI3_BARY, TRI3_W, 2)])
def test_rules_exact_on_polynomials(bary, w, deg):
    V = np.array([[[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]]])
    pts = points(V, bary)[0]
    for p in range(deg + 1):
        for q in range(deg + 1 - p):
            approx = 0.5 * np.sum(w * pts[:, 0] ** p * pts[:, 1] ** q)
            assert approx == pytest.approx(_monomial_integral(p, q), rel=1e-13, abs=1e-15)
    # degree deg + 1 is not integrated exactly by at least one monomial
    errs = [abs(0.5 * np.sum(w * pts[:, 0] ** p * pts[:, 1] ** (deg + 1 - p)) - _monomial_integral(p, deg + 1 - p))
            for p in range(deg + 2)]
    assert max(errs) > 1e-8
