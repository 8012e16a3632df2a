"""Pins of the oracle's BS3 integrator and grid / Gauss-Hermite rules (CPU)."""
import numpy as np
import pytest

from oracle import integrate as I


def _run(y0, acc_fn, Omega, N, dt0, atol=1e-8, rtol=1e-6, dt_min=1e-5, dt_max=0.1):
    y = np.asarray(y0, dtype=np.float64).reshape(6, -1).copy()
    k1 = I.deriv_barred(y, acc_fn, Omega)
    t = np.zeros(y.shape[1])
    dt = np.full(y.shape[1], dt0)
    for _ in range(N):
        t, y, dt, k1, _ = I.rk_step(t, y, dt, k1, acc_fn, Omega, atol, rtol, dt_min, dt_max)
    return t, y


def test_exact_for_constant_acceleration():
    """BS3 and its embedded estimate are exact when x(t) is quadratic in t."""
    g = np.array([0.3, -1.2, 0.7])

    def acc(x, y, z):
        return (np.full_like(x, g[0]), np.full_like(x, g[1]), np.full_like(x, g[2]))
    y0 = np.array([1.0, -2.0, 0.5, 0.4, 0.1, -0.3])
    t, y = _run(y0, acc, 0.0, 40, 0.01)
    t = t[0]
    ref_x = y0[:3] + y0[3:] * t + 0.5 * g * t * t
    ref_v = y0[3:] + g * t
    np.testing.assert_allclose(y[:3, 0], ref_x, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(y[3:, 0], ref_v, rtol=1e-13, atol=1e-13)


def test_global_order_three_on_harmonic_oscillator():
    """Fixed steps (tolerances never reject): error ratio 8 when dt halves."""
    w2 = 4.0

    def acc(x, y, z):
        return -w2 * x, -w2 * y, -w2 * z
    y0 = np.array([1.0, 0.0, 0.0, 0.0, 0.0, 0.0])
    errs = []
    for dt in (0.02, 0.01, 0.005):
        N = int(round(2.0 / dt))
        t, y = _run(y0, acc, 0.0, N, dt, atol=1e6, rtol=1e6, dt_min=dt, dt_max=dt)
        errs.append(abs(y[0, 0] - np.cos(2.0 * t[0])))
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 7.0 < r1 < 9.0 and 7.0 < r2 < 9.0, (r1, r2)


def test_rotating_frame_free_motion_closed_form():
    """No force: X(t) = e^{-i Omega t}(X0 + V0 t), V(t) = V0 e^{-i Omega t}."""
    Om = 0.7

    def acc(x, y, z):
        return 0 * x, 0 * y, 0 * z
    y0 = np.array([1.0, 0.5, 0.2, -0.3, 0.8, 0.1])
    t, y = _run(y0, acc, Om, 3000, 1e-3, atol=1e-12, rtol=1e-12, dt_min=1e-6, dt_max=1e-3)
    t = t[0]
    X0, V0 = y0[0] + 1j * y0[1], y0[3] + 1j * y0[4]
    X = np.exp(-1j * Om * t) * (X0 + V0 * t)
    V = V0 * np.exp(-1j * Om * t)
    assert abs(y[0, 0] + 1j * y[1, 0] - X) < 1e-9
    assert abs(y[3, 0] + 1j * y[4, 0] - V) < 1e-9
    assert abs(y[2, 0] - (y0[2] + y0[5] * t)) < 1e-12


def test_jacobi_integral_conserved(canon_params):
    from oracle.model import baryon_params
    _, ph, pb = canon_params
    acc_fn, pot_fn = I.make_acc_pot(baryon_params(pb), ph)
    y0 = np.array([[4.0, 0.0, 0.1, 0.0, 190.0, 5.0], [-2.0, 3.0, -0.2, -150.0, -90.0, 10.0]]).T
    Om = -pb['Omega_bar']
    EJ0 = I.compute_EJ(y0, pot_fn, Om)
    t, y = _run(y0, acc_fn, Om, 1500, 1e-4, atol=1e-10, rtol=1e-9, dt_min=1e-6, dt_max=0.01)
    EJ = I.compute_EJ(y, pot_fn, Om)
    assert np.all(np.abs(EJ / EJ0 - 1) < 1e-7)


def test_grid_cell_centres_map_in_order():
    """Cell centres built as the reference's loader builds them (utils.py:627-657)
    fall in cells 0..P-1 in order; edges follow the reference rule."""
    from schwarmax_amd.synth import rzphi_cells
    nR, nz, nphi = 7, 5, 6
    Rg, zg, pg, *_ = rzphi_cells(nR, nz, nphi)
    idx = I.assign_regular_grid(np.stack([Rg, zg, pg], -1), [0, -3, -np.pi], [10, 3, np.pi],
                                [nR, nz, nphi])
    assert np.array_equal(idx, np.arange(nR * nz * nphi))
    # lower edges belong to the cell, upper limit is out of bounds, phi = +pi dropped
    pts = np.array([[0.0, -3.0, -np.pi], [10.0, 0.0, 0.0], [5.0, 3.0, 0.0], [5.0, 0.0, np.pi],
                    [-1e-300, 0.0, 0.0], [9.999999, 2.999999, 3.14159]])
    got = I.assign_regular_grid(pts, [0, -3, -np.pi], [10, 3, np.pi], [nR, nz, nphi])
    P = nR * nz * nphi
    assert got[0] == 0 and got[1] == P and got[2] == P and got[3] == P and got[4] == P
    assert got[5] == (nR - 1) + (nz - 1) * nR + (nphi - 1) * nR * nz


def test_gauss_hermite_basis():
    y = np.linspace(-3, 3, 13)
    H = I.hermite(y)
    np.testing.assert_allclose(H[2], (2 * y * y - 1) / np.sqrt(2), rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(H[3], y * (2 * y * y - 3) / np.sqrt(3), rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(H[4], (4 * y ** 4 - 12 * y * y + 3) / np.sqrt(24), rtol=1e-13, atol=1e-13)
    x, w = np.polynomial.hermite.hermgauss(40)
    Hq = I.hermite(x)
    G = np.array([[np.sum(w * Hq[a] * Hq[b]) for b in range(5)] for a in range(5)])
    np.testing.assert_allclose(G, np.sqrt(np.pi) * np.eye(5), atol=1e-12)
    yy, g = I.gauss(np.array([10.0]), 10.0, 2.0)
    assert yy[0] == 0.0 and abs(g[0] - 1 / (2 * np.sqrt(2 * np.pi))) < 1e-16


def test_h_to_V_sigma_special_cases():
    v, s = I.h_to_V_sigma(0.0, 0.0, 12.0, 80.0)
    assert v == 12.0 and abs(s - 80.0) < 1e-12
    v, s = I.h_to_V_sigma(0.1, 0.0, 0.0, 100.0, h3=0.0, h4=0.0)
    assert abs(v - 100 * np.sqrt(2) * 0.1) < 1e-12
    assert abs(s - 100 * np.sqrt(1 - 2 * 0.01)) < 1e-12
