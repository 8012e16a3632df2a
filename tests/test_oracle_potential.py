"""Pins of the oracle's potential / density / gradient (CPU).

Each check is chosen so that a plausible slip in the restatement (a dropped
term, a wrong power, a swapped a/b) fails it: Poisson's equation ties every
potential to its density, a brute-force needle convolution ties the A_n / C_n
recurrences to the integral they stand for, finite differences tie the dual
numbers to the potential, and closed forms fix the axisymmetric limits.
"""
import numpy as np
import pytest
from scipy import integrate

from oracle import constants as C
from oracle import potential as P


def test_constants_closed_form():
    # astropy: G*M_sun = IAU nominal 1.3271244e20 m^3/s^2; Julian Gyr; IAU pc
    assert C.G == 4.498502151469554e-06
    assert C.KPCGYR_TO_KMS == 0.9777922216807892
    assert abs(C.KMS_TO_KPCGYR * C.KPCGYR_TO_KMS - 1.0) < 1e-15


def _laplacian(f, x, y, z, h=2e-3):
    """4th-order central-difference Laplacian."""
    def d2(fp2, fp1, f0, fm1, fm2):
        return (-fp2 + 16 * fp1 - 30 * f0 + 16 * fm1 - fm2) / (12 * h * h)
    f0 = f(x, y, z)
    lx = d2(f(x + 2 * h, y, z), f(x + h, y, z), f0, f(x - h, y, z), f(x - 2 * h, y, z))
    ly = d2(f(x, y + 2 * h, z), f(x, y + h, z), f0, f(x, y - h, z), f(x, y - 2 * h, z))
    lz = d2(f(x, y, z + 2 * h), f(x, y, z + h), f0, f(x, y, z - h), f(x, y, z - 2 * h))
    return lx + ly + lz


PTS = np.array([[1.0, 0.5, 0.2], [3.0, -1.0, -0.4], [-2.5, 2.0, 0.7], [5.0, 0.3, 0.1],
                [0.7, -2.2, -0.3], [6.5, 4.0, 1.1]])


@pytest.mark.parametrize("which", ["mn", "t3", "v4"])
def test_poisson_pairs_potential_and_density(which, canon_params):
    """nabla^2 Phi = 4 pi G rho for each component (pins both formulas)."""
    _, ph, pb = canon_params
    x, y, z = PTS.T
    Mb = 10.0 ** pb['logM_bar']
    if which == "mn":
        mp = P._mn_params(pb)
        phi = lambda a, b, c: P.mn_potential(a, b, c, mp)  # noqa: E731
        rho = P.mn_density(x, y, z, mp)
    elif which == "t3":
        args = (Mb, pb['a_bar'], pb['b_bar'], pb['L_bar'], P.GAMMA_BAR)
        phi = lambda a, b, c: P.t3_potential(a, b, c, *args)  # noqa: E731
        rho = P.t3_density(x, y, z, *args)
    else:
        # the V4 bulge is compact (a = b = 0.5 kpc): test where its density is
        # not swamped by finite-difference roundoff of Phi
        x, y, z = np.array([[0.3, 0.2, 0.1], [0.5, -0.4, -0.2], [-0.8, 0.6, 0.3],
                            [1.2, 0.1, 0.05], [0.05, 0.9, -0.4]]).T
        args = (Mb, P.V4_A, P.V4_B, P.V4_L, P.V4_GAMMA)
        phi = lambda a, b, c: P.v4_potential(a, b, c, *args)  # noqa: E731
        rho = P.v4_density(x, y, z, *args)
    lap = _laplacian(phi, x, y, z)
    np.testing.assert_allclose(lap, 4 * np.pi * C.G * rho, rtol=2e-5)


@pytest.mark.parametrize("n", [1, 3, 5, 7, 9])
@pytest.mark.parametrize("gamma", [0.0, 1.0, 0.4])
def test_needle_convolution_brute_force(n, gamma):
    """convolve_val == integral of the needle kernel times M / r^n."""
    M, L = 2.5, 1.7
    for (x, y, z) in [(0.3, 0.8, 1.1), (-2.0, 0.2, 0.9), (4.0, -1.0, 1.6)]:
        def integrand(xp):
            k = (1 - gamma) / (2 * L) + gamma * (L - abs(xp)) / (L * L)
            return k * M / ((x - xp) ** 2 + y * y + z * z) ** (n / 2)
        ref, _ = integrate.quad(integrand, -L, L, epsabs=0, epsrel=1e-13, limit=200, points=[0.0])
        got = P.convolve_val(M, x, L, gamma, y, z, n)
        assert abs(got - ref) <= 1e-9 * abs(ref), (x, y, z, got, ref)


def test_axisymmetric_T1_is_miyamoto_nagai():
    """T1 with L = 0 is the Miyamoto-Nagai potential; small L converges to it."""
    M, a, b = 3e10, 2.0, 0.4
    x, y, z = PTS.T
    mn = -C.G * M / np.sqrt(x * x + y * y + (a + np.sqrt(z * z + b * b)) ** 2)
    np.testing.assert_allclose(P.t1_potential(x, y, z, M, a, b, 0.0, 0.0), mn, rtol=1e-14)
    np.testing.assert_allclose(P.t1_potential(x, y, z, M, a, b, 1e-3, 0.0), mn, rtol=1e-6)


def test_far_field_point_mass(canon_params):
    _, ph, pb = canon_params
    Mb = 10.0 ** pb['logM_bar']
    r = 2000.0
    d = np.array([[1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0], [0.6, 0.0, 0.8]]) * r
    for fn, args in [(P.t3_potential, (Mb, pb['a_bar'], pb['b_bar'], pb['L_bar'], 1.0)),
                     (P.v4_potential, (Mb, 0.5, 0.5, 0.1, 0.0))]:
        v = fn(d[:, 0], d[:, 1], d[:, 2], *args)
        np.testing.assert_allclose(v, -C.G * Mb / r, rtol=5e-3)


def test_nfw_closed_form(canon_params):
    _, ph, _ = canon_params
    x, y, z = PTS.T
    r = np.sqrt(x * x + y * y + z * z)
    M = 10.0 ** ph['logM']
    ref = -C.G * M * np.log(1 + r / ph['Rs']) / r
    np.testing.assert_allclose(P.nfw_potential(x, y, z, ph), ref, rtol=1e-11)


def test_gradient_matches_finite_differences(canon_params):
    from oracle.model import baryon_params
    _, ph, pb = canon_params
    bp = baryon_params(pb)
    x, y, z = PTS.T
    ax, ay, az, _ = P.accel(x, y, z, bp, ph)
    h = 1e-5
    f = lambda a, b, c: P.potential_func(a, b, c, bp, ph)  # noqa: E731
    fx = -(f(x + h, y, z) - f(x - h, y, z)) / (2 * h)
    fy = -(f(x, y + h, z) - f(x, y - h, z)) / (2 * h)
    fz = -(f(x, y, z + h) - f(x, y, z - h)) / (2 * h)
    scale = np.sqrt(ax ** 2 + ay ** 2 + az ** 2)
    for g, fd in ((ax, fx), (ay, fy), (az, fz)):
        assert np.max(np.abs(g - fd) / scale) < 1e-8


def test_rotation_identity_for_default_direction():
    R = P.get_mat(0.0, 0.0, 1.0)
    assert np.array_equal(R, np.eye(3))
    R2 = P.get_mat(0.3, -0.2, 0.9)
    np.testing.assert_allclose(R2 @ R2.T, np.eye(3), atol=1e-12)
    v = np.array([0.3, -0.2, 0.9]) / np.linalg.norm([0.3, -0.2, 0.9])
    np.testing.assert_allclose(R2 @ np.array([0, 0, 1.0]), v, atol=1e-12)
