"""Pins of oracle functions by closed forms, hand-derived values, an
independent quadrature and invariants (CPU).  Each test names the reference
lines the oracle function restates.
"""
import numpy as np
import pytest

from oracle import model as OM, integrate as OI, jeans as J
from oracle.constants import KPCGYR_TO_KMS


# --------------------------------------------------------------------------
# density target (model.py:192-213): Sobol cell masses vs Gauss-Legendre
# --------------------------------------------------------------------------
def test_density_target_matches_gauss_legendre_cell_masses(canon_params):
    """The 1024-point Sobol estimate of each cell's mass against a 12^3-point
    tensor Gauss-Legendre rule of the integral of rho R over the cell."""
    from schwarmax_amd import synth
    from oracle import potential as P
    theta, dd, _ = synth.make_galaxy('plain-2k', seed=2, n=4, grid=(20, 10, 16))
    _, pb = OM.theta_to_params(theta)
    got = OM.density_target(dd, pb)
    xg, wg = np.polynomial.legendre.leggauss(12)
    rng = np.random.default_rng(0)
    cells = rng.choice(got.size, size=25, replace=False)
    dR, dz, dphi = float(dd['dR']), float(dd['dz']), float(dd['dphi'])
    for c in cells:
        R = dd['R_grid'][c] + 0.5 * dR * xg[:, None, None]
        z = dd['z_grid'][c] + 0.5 * dz * xg[None, :, None]
        ph = dd['phi_grid'][c] + 0.5 * dphi * xg[None, None, :]
        w = wg[:, None, None] * wg[None, :, None] * wg[None, None, :]
        rho = P.density_func(R * np.cos(ph), R * np.sin(ph), z + 0 * R, pb)
        ref = np.sum(w * rho * R) * (0.5 * dR) * (0.5 * dz) * (0.5 * dphi)
        assert abs(got[c] - ref) <= 3e-3 * abs(ref), (c, got[c], ref)
    # the Jacobian R matters at this precision: dropping it is an O(1) error
    assert np.all(got > 0)


# --------------------------------------------------------------------------
# data normalisation, mass per orbit and gamma_kin (model.py:217-258)
# --------------------------------------------------------------------------
def test_targets_on_a_two_aperture_example(monkeypatch):
    """Hand-derived: M/L = 2 (so the data are divided by 1/2), two orbits and
    cell masses (1, 3), so mass per orbit = 2; pixels of 1 kpc^2 = 1e6 pc^2,
    s = 1 / (2 sqrt(pi)) so the Gaussian normalisation 2 sqrt(pi) s is 1, and
    h4 = (0, 4/sqrt(6)) so that 1 + h4 sqrt(6)/4 = (1, 2)."""
    monkeypatch.setattr(OM, 'density_target', lambda dd, p: np.array([1.0, 3.0]))
    s = 1.0 / (2.0 * np.sqrt(np.pi))
    dd = {
        'XY_density_data': np.array([2.0, 4.0]), 'XY_density_data_err': np.array([0.2, 0.4]),
        'h1_data': np.zeros(2), 'h2_data': np.zeros(2), 'h3_data': np.zeros(2),
        'h4_data': np.array([0.0, 4.0 / np.sqrt(6.0)]),
        'h1_data_err': np.full(2, 0.1), 'h2_data_err': np.full(2, 0.1),
        'h3_data_err': np.full(2, 0.1), 'h4_data_err': np.full(2, 0.1),
        'X_minmax': np.array([-1.0, 1.0]), 'Y_minmax': np.array([-1.0, 1.0]),
        'nX_nY': np.array([2, 2]), 'num_per_bin': np.array([1, 3]),
        's': np.array([s, s]), 'v0': np.zeros(2),
    }
    t = OM._targets(dd, {'mass_to_light_ratio': 2.0}, 2)
    assert t['lm'] == 0.5 and t['mmpo'] == 2.0
    assert np.allclose(t['y_Rzphi'], [0.5, 1.5], rtol=1e-15)
    assert np.allclose(t['sig_Rzphi'], [0.01 + 5e-11, 0.03 + 5e-11], rtol=1e-14)
    assert np.allclose(t['y_xy'], [2.0, 4.0], rtol=1e-15)
    assert np.allclose(t['sig_xy'], [0.2, 0.4], rtol=1e-10)
    # M_per_bin = y_xy * num_per_bin * 1e6 = (2e6, 12e6); over (1, 2)
    assert np.allclose(t['gamma_kin'], [2e6, 6e6], rtol=1e-12)


# --------------------------------------------------------------------------
# bootstrap chi^2 and log-mean-exp (likelihood.py:184-244)
# --------------------------------------------------------------------------
def _loglike_case(K, amp=1.0):
    md = {
        'density_3d': np.array([1.0, 2.0]),
        'density_3d_model': np.array([[1.1, 2.0], [1.0, 2.2]] + [[1.0, 2.0]] * (K - 2)),
        'surface_density': np.array([[1.0, 1.0], [1.1, 1.1]] + [[1.0, 1.0]] * (K - 2)),
    }
    dd = {'XY_density_data': np.array([1.0, 1.0]), 'XY_density_data_err': np.array([0.05, 0.05]),
          'XY_standard_normal': np.zeros((K, 2))}
    for k in (1, 2, 3, 4):
        md['h%d' % k] = np.zeros((K, 2))
        dd['h%d_data' % k] = np.zeros(2)
        dd['h%d_data_err' % k] = np.full(2, 0.1)
        dd['h%d_standard_normal' % k] = np.zeros((K, 2))
    return md, dd


def test_bootstrap_loglike_hand_values():
    """sig = 0.1 y: the density residuals are (0.1/0.1)^2 = 1 in row 0 and
    (0.2/0.2)^2 = 1 in row 1, summed over all rows (no axis, likelihood.py:221)
    and added to every row: res_R = 2.  Surface density: row 1 is 10% high in
    both apertures, res_xy = (0, 2).  logl = -(1, 2); log-mean-exp."""
    md, dd = _loglike_case(2)
    val, rows = OM.bootstrap_loglike(md, dd)
    assert np.allclose(rows, [-1.0, -2.0], rtol=1e-12)
    assert abs(val - (-1.0 + np.log((1.0 + np.exp(-1.0)) / 2.0))) < 1e-12
    # sigma_amplify 2: every residual divided by 4
    val2, rows2 = OM.bootstrap_loglike(md, dd, sigma_amplify=2.0)
    assert np.allclose(rows2, [-0.25, -0.5], rtol=1e-12)
    # a NaN model entry is dropped by nansum
    md['h2'][1, 0] = np.nan
    assert np.allclose(OM.bootstrap_loglike(md, dd)[1], rows, rtol=1e-12)


def test_bootstrap_loglike_equal_rows():
    """Equal bootstrap rows: log-mean-exp returns that row's value."""
    md, dd = _loglike_case(3)
    md['density_3d_model'] = np.tile([[1.0, 2.2]], (3, 1))
    md['surface_density'] = np.tile([[1.05, 1.0]], (3, 1))
    val, rows = OM.bootstrap_loglike(md, dd)
    # res_R = 3 rows x 1 = 3 (broadcast), res_xy = 0.25 per row
    assert np.allclose(rows, -0.5 * (3.0 + 0.25), rtol=1e-12)
    assert abs(val - rows[0]) < 1e-12


# --------------------------------------------------------------------------
# realization average (integrates.py:279-300)
# --------------------------------------------------------------------------
def test_realization_average_divides_by_valid_count_plus_tenth():
    per = {'valid': np.array([1.0, 1.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0])}
    vals = np.array([2.0, 4.0, 0.0, 6.0, 0.0, 0.0, 0.0, 0.0])     # invalid ones already zeroed
    for key in ('Rzphi', 'surface_density', 'h0', 'h1', 'h2', 'h3', 'h4'):
        per[key] = np.stack([vals, 2 * vals], axis=1)
    out = OI.realization_average(per, 2, 4)
    assert np.allclose(out['Rzphi'][0], [12.0 / 3.1, 24.0 / 3.1], rtol=1e-15)
    assert np.array_equal(out['Rzphi'][1], [0.0, 0.0])          # 0 / 0.1
    assert np.array_equal(out['n_valid'], [3.0, 0.0])


# --------------------------------------------------------------------------
# trajectory path (integrates.py:319-512)
# --------------------------------------------------------------------------
def test_withtraj_time_units_and_normalisation():
    """t_traj is the running sum of the accepted steps (brute force by the
    single-step function), velocities are the state's times KPCGYR_TO_KMS, and
    an orbit that never leaves the (R, z, phi) grid has a density row summing
    to 1 (samples weighted dt / T, T = sum dt + 1e-12)."""
    from oracle import potential as P
    from schwarmax_amd import synth
    theta, dd, _ = synth.make_galaxy('plain-2k', seed=1, n=3, grid=(8, 4, 6), n_ap=10, nxy=10)
    ph, pb = OM.theta_to_params(theta)
    pbb = OM.baryon_params(pb)
    acc_fn, pot_fn = OI.make_acc_pot(pbb, ph)
    w0 = np.array([[4.0, 0.0, 0.05, 0.0, -180.0, 5.0], [2.5, 1.0, -0.1, 60.0, -150.0, -3.0]])
    grids = OI.Grids(10, dd['bin_mapping'], dd['num_per_bin'], [[0, 10.], [-3, 3], [-np.pi, np.pi]],
                     [[-5, 5], [-5, 5]], [8, 4, 6], [10, 10], np.zeros(10), np.full(10, 150.0),
                     np.eye(3))
    N = 300
    T_total = np.array([1e-3, 1e-3])              # validity gate passed at once
    dt0 = np.array([1e-3, 2e-3])
    out = OI.integrate_withtraj(w0, acc_fn, pot_fn, N, T_total, dt0, -40.0, 1e-7, 1e-6, 1e-5, 0.3,
                                grids)
    y = w0.T.copy()
    k1 = OI.deriv_barred(y, acc_fn, -40.0)
    t, dt = np.zeros(2), dt0.copy()
    tsum = np.zeros(2)
    for i in range(N):
        t, y, dt, k1, dt_out = OI.rk_step(t, y, dt, k1, acc_fn, -40.0, 1e-7, 1e-6, 1e-5, 0.3)
        tsum = tsum + dt_out
        assert np.array_equal(out['t_traj'][:, i], tsum)
        assert np.array_equal(out['o_traj'][:, i, :3], y[:3].T)
        assert np.array_equal(out['o_traj'][:, i, 3:], y[3:].T * KPCGYR_TO_KMS)
    assert np.array_equal(out['t_final'], t) and np.array_equal(out['t_traj'][:, -1], t)
    assert KPCGYR_TO_KMS == pytest.approx(3.0856775814913673e16 / 3.15576e16, rel=1e-15)
    assert np.all(out['valid'] == 1.0)
    R = np.hypot(out['o_traj'][:, :, 0], out['o_traj'][:, :, 1])
    inside = (R < 10.0) & (np.abs(out['o_traj'][:, :, 2]) < 3.0)
    for i in range(2):
        if inside[i].all():
            assert abs(out['Rzphi'][i].sum() - 1.0) < 1e-10
    assert inside.all()


# --------------------------------------------------------------------------
# rotation curve, orbital time, clockwise conversion (utils.py:144-273) and
# the Jeans mean rotation (jeans.py:9-79)
# --------------------------------------------------------------------------
GM = 1.0e5


def _kepler(x, y, z, pb, ph):
    return -GM / np.sqrt(x * x + y * y + z * z)


def test_rotation_curve_and_orbital_time_kepler():
    R = np.array([0.5, 2.0, 7.0])
    vc = J.rotation_curve(R, 0.0 * R, _kepler, None, None)
    assert np.allclose(vc, np.sqrt(GM / R), rtol=1e-5)
    T = J.orbital_timescale(R, 0.0 * R, _kepler, None, None)
    assert np.allclose(T, 2 * np.pi * np.sqrt(R ** 3 / GM), rtol=1e-5)
    # R_safe = max(|R|, 2 dR): every radius below 2e-3 is evaluated at 2e-3
    small = J.rotation_curve(np.array([0.0, 1e-4, -1e-3]), np.zeros(3), _kepler, None, None)
    assert np.all(small == J.rotation_curve(np.array([2e-3]), np.zeros(1), _kepler, None, None))
    # a repulsive potential: v_c clamps to 0, T_orb uses |dPhi/dR|
    rep = lambda x, y, z, pb, ph: -_kepler(x, y, z, pb, ph)  # noqa: E731
    assert np.all(J.rotation_curve(R, 0.0 * R, rep, None, None) == 0.0)
    assert np.allclose(J.orbital_timescale(R, 0.0 * R, rep, None, None), T, rtol=1e-12)


def test_clockwise_conversion():
    """Orthogonal map (speed preserved), v_R radial, and clockwise sense:
    L_z = x v_y - y v_x = -R v_phi; at phi = 0, (v_x, v_y) = (v_R, -v_phi)."""
    rng = np.random.default_rng(4)
    R, ph = rng.uniform(0.1, 9, 50), rng.uniform(-np.pi, np.pi, 50)
    vR, vp = rng.normal(size=50), rng.normal(size=50)
    x, y, vx, vy = J.cyl_to_cart_clockwise(R, ph, vR, vp)
    assert np.allclose(vx ** 2 + vy ** 2, vR ** 2 + vp ** 2, rtol=1e-13)
    assert np.allclose((x * vx + y * vy) / R, vR, rtol=1e-12, atol=1e-13)
    assert np.allclose(x * vy - y * vx, -R * vp, rtol=1e-12, atol=1e-12)
    x0, y0, vx0, vy0 = J.cyl_to_cart_clockwise(2.0, 0.0, 3.0, 5.0)
    assert (x0, y0, vx0, vy0) == (2.0, 0.0, 3.0, -5.0)


def test_jeans_mean_rotation_separable_closed_form():
    """Separable model: Phi = Omega^2 R^2 / 2 + k z^2 / 2, nu = exp(-R/h)
    exp(-z^2 / 2 s^2).  Then nu sigma_z^2 = int_{|z|}^{10} nu dPhi/dz' dz' =
    k s^2 exp(-R/h) (exp(-z^2/2s^2) - exp(-50/s^2)), so sigma_z^2 = k s^2 (to
    exp(-200)), d(nu sigma_R^2)/dR = -nu sigma_R^2 / h, and the radial Jeans
    equation with sigma_R = sigma_phi gives v_phi^2 = Omega^2 R^2 - (R/h) k s^2."""
    Om, k, h, s = 10.0, 100.0, 3.0, 0.5
    pot = lambda x, y, z, pb, ph: 0.5 * Om ** 2 * (x * x + y * y) + 0.5 * k * z * z  # noqa: E731
    dens = lambda x, y, z, pb: np.exp(-np.sqrt(x * x + y * y) / h - z * z / (2 * s * s))  # noqa: E731
    xs = np.array([3.0, 1.0, -2.0])
    ys = np.array([0.0, 2.0, 1.0])
    zs = np.array([0.2, -0.1, 0.0])
    vmean, sR, sz, sphi = J.jeans_moments(xs, ys, zs, None, None, pot, dens)
    R = np.hypot(xs, ys)
    assert np.allclose(sz ** 2, k * s * s, rtol=2e-4)
    assert np.array_equal(sR, sz) and np.array_equal(sphi, sR)
    assert np.allclose(vmean ** 2, Om ** 2 * R ** 2 - (R / h) * k * s * s, rtol=2e-4)
