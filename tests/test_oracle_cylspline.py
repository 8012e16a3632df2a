"""CylSpline oracle pins (CPU): Legendre Q against mpmath, the Green's-function
expansion it implements, the not-a-knot spline and the separable interpolant
on functions they must reproduce exactly, and the whole builder against the
closed-form potential of an off-centre Plummer sphere."""
import math

import numpy as np
import pytest

from oracle import cylspline as C
from oracle.constants import G

PL_M, PL_A, PL_OFF = 1e10, 1.5, (0.6, 0.4)


def plummer_rho(x, y, z, params=None):
    r2 = (x - PL_OFF[0]) ** 2 + (y - PL_OFF[1]) ** 2 + z ** 2
    return 3 * PL_M / (4 * np.pi * PL_A ** 3) * (1 + r2 / PL_A ** 2) ** -2.5


def plummer_phi(x, y, z):
    return -G * PL_M / np.sqrt((x - PL_OFF[0]) ** 2 + (y - PL_OFF[1]) ** 2 + z ** 2 + PL_A ** 2)


def plummer_acc(x, y, z):
    d = np.array([x - PL_OFF[0], y - PL_OFF[1], z])
    return -G * PL_M * d / (np.sum(d * d, axis=0) + PL_A ** 2) ** 1.5


def test_legendre_q_against_mpmath():
    mp = pytest.importorskip("mpmath")
    mp.mp.dps = 30
    chis = np.concatenate([1 + np.logspace(-11, 0, 23), np.linspace(1.02, 1.6, 12), [3.0, 10.0, 100.0]])
    for m in range(13):
        for chi in chis:
            ref = float(mp.legenq(m - 0.5, 0, mp.mpf(chi), type=3).real)
            tol = max(1e-13, 2e-17 / (chi - 1.0))   # 1 - 1/chi^2 carries 1e-16/(chi - 1)
            assert abs(C.legendre_q_half(m, chi) - ref) <= tol * abs(ref), (m, chi)
    # the floor: at chi = 1 (y = 1 - 1/chi^2 = 0) only the n = 0 term of the
    # logarithmic series survives, with ln y replaced by ln 1e-12:
    # Q = P_m Gamma(m+1)/(Gamma(a) Gamma(b)) [2 psi(1) - psi(a) - psi(b) - ln 1e-12]
    for m in (0, 5, 12):
        a, b = mp.mpf(2 * m + 1) / 4, mp.mpf(2 * m + 3) / 4
        P = mp.sqrt(mp.pi) * mp.gamma(m + mp.mpf(1) / 2) / (mp.gamma(m + 1) * mp.mpf(2) ** (m + mp.mpf(1) / 2))
        ref = float(P * mp.gamma(m + 1) / (mp.gamma(a) * mp.gamma(b))
                    * (2 * mp.digamma(1) - mp.digamma(a) - mp.digamma(b) - mp.log(mp.mpf('1e-12'))))
        assert abs(C.legendre_q_half(m, 1.0) - ref) <= 1e-13 * abs(ref), m


def test_green_function_expansion():
    """sum_m eps_m Xi_m cos(m dphi) = 1/|x - x'| (eps_0 = 1, eps_m = 2): the
    mode kernel's normalisation and the factor 2 of the m > 0 modes."""
    R, z, Rp, zp, dphi = 3.0, 0.5, 1.2, -0.7, 0.9
    s = sum((1.0 if m == 0 else 2.0) * C.kernel_xi(m, R, z, Rp, zp) * math.cos(m * dphi)
            for m in range(60))
    d2 = R * R + Rp * Rp - 2 * R * Rp * math.cos(dphi) + (z - zp) ** 2
    assert abs(s - 1.0 / math.sqrt(d2)) <= 1e-12 / math.sqrt(d2)


def test_not_a_knot_spline_exact_on_cubics():
    """A cubic is its own not-a-knot spline: M = p'' at the knots; and for any
    data the third derivative is continuous at the second and penultimate
    knots."""
    rng = np.random.default_rng(0)
    x = np.cumsum(rng.uniform(0.2, 1.0, 9))
    c = [0.3, -1.1, 0.7, 0.25]
    y = c[0] + c[1] * x + c[2] * x ** 2 + c[3] * x ** 3
    assert np.allclose(C.second_derivatives(x, y), 2 * c[2] + 6 * c[3] * x, rtol=0, atol=1e-11)
    y = rng.normal(size=9)
    M = C.second_derivatives(x, y)
    h = np.diff(x)
    third = np.diff(M) / h                      # piecewise-constant third derivative
    assert abs(third[0] - third[1]) < 1e-10 * np.abs(third).max()
    assert abs(third[-1] - third[-2]) < 1e-10 * np.abs(third).max()
    # interior rows: continuity of the first derivative
    for i in range(1, 8):
        left = (y[i] - y[i - 1]) / h[i - 1] + h[i - 1] * (2 * M[i] + M[i - 1]) / 6
        right = (y[i + 1] - y[i]) / h[i] - h[i] * (2 * M[i] + M[i + 1]) / 6
        assert abs(left - right) < 1e-10


def test_separable_interpolant_exact_cases():
    """The separable scheme reproduces p(u) (a + b v) and (a + b u) q(v) for
    cubic p, q exactly (M_y is interpolated linearly in u), with exact
    derivatives."""
    rng = np.random.default_rng(1)
    gu = np.cumsum(rng.uniform(0.1, 0.5, 8))
    gv = np.cumsum(rng.uniform(0.1, 0.5, 7))
    U, V = np.meshgrid(gu, gv, indexing='ij')
    p = lambda t: 0.2 - t + 0.5 * t ** 2 - 0.3 * t ** 3  # noqa: E731
    dp = lambda t: -1 + t - 0.9 * t ** 2  # noqa: E731
    pts = np.column_stack([rng.uniform(gu[0], gu[-1], 50), rng.uniform(gv[0], gv[-1], 50)])
    for f, fu, fv in ((lambda u, v: p(u) * (1.5 - 0.4 * v), lambda u, v: dp(u) * (1.5 - 0.4 * v),
                       lambda u, v: -0.4 * p(u)),
                      (lambda u, v: (0.7 + 2 * u) * p(v), lambda u, v: 2 * p(v),
                       lambda u, v: (0.7 + 2 * u) * dp(v))):
        vals = f(U, V)
        mx, my = C.precompute_splines((gu, gv), vals)
        val, du, dv = C.spline_eval_with_derivs(pts, (gu, gv), vals, mx, my)
        assert np.allclose(val, f(pts[:, 0], pts[:, 1]), rtol=0, atol=1e-11)
        assert np.allclose(du, fu(pts[:, 0], pts[:, 1]), rtol=0, atol=1e-10)
        assert np.allclose(dv, fv(pts[:, 0], pts[:, 1]), rtol=0, atol=1e-10)
    out = C.spline_eval_with_derivs(np.array([[gu[0] - 0.1, gv[1]]]), (gu, gv), vals, mx, my, 0.0)
    assert out[0][0] == 0.0 and out[1][0] == 0.0 and out[2][0] == 0.0


@pytest.fixture(scope="module")
def plummer_tables():
    return C.get_phi_m(plummer_rho, None, 24, 16, 0.05, 20.0, 0.05, 15.0, 6, 24, 1600)


def test_builder_against_closed_form(plummer_tables):
    """The whole builder (density modes, splines, Simpson nodes, Legendre
    kernel, mode sum) on an off-centre Plummer sphere, which has modes of every
    m with real and imaginary parts: Phi to 5e-3, -grad Phi to 0.1 of its
    largest value (the reference's quadrature accuracy at this resolution)."""
    d = plummer_tables
    rng = np.random.default_rng(0)
    R, ph, z = rng.uniform(0.3, 6, 300), rng.uniform(0, 2 * np.pi, 300), rng.uniform(-4, 4, 300)
    x, y = R * np.cos(ph), R * np.sin(ph)
    pe = plummer_phi(x, y, z)
    assert np.max(np.abs(C.evaluate_phi(x, y, z, d) - pe) / np.abs(pe)) < 5e-3
    ae = plummer_acc(x, y, z)
    assert np.max(np.abs(C.get_acc(x, y, z, d) - ae)) < 0.1 * np.abs(ae).max()
    # the m = 0 part alone misses the off-centre structure
    assert np.max(np.abs(C.evaluate_phi(x, y, z, d, modes=1) - pe) / np.abs(pe)) > 2e-2
