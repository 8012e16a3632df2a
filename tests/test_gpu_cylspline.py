"""CylSpline on the GPU (CylindricalSpline.py, CubicSpline.py): the device
Legendre Q against mpmath, the builder's tables and the evaluator against the
oracle on the same inputs, and the potential, force and Poisson density of an
off-centre Plummer sphere against its closed forms at higher resolution."""
import numpy as np
import pytest

from test_oracle_cylspline import plummer_rho, plummer_phi, plummer_acc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def test_legendre_q_device(gpu):
    """smx_legendre_q_half for m = 0..20 from chi - 1 = 1e-10 to 1e3 against
    mpmath (5e-14), and the reference's floor at chi - 1 below 5e-13."""
    mp = pytest.importorskip("mpmath")
    import torch
    from schwarmax_amd import _lib
    from oracle import cylspline as OC
    mp.mp.dps = 30
    delta = np.concatenate([np.logspace(-10, 3, 60), np.linspace(0.02, 0.6, 30)])
    d_t = _lib.dev(delta)
    out = torch.empty_like(d_t)
    L = _lib.lib()
    for m in range(21):
        _lib.check(L.smx_legendre_q_half(m, _lib.ptr(d_t), d_t.numel(), _lib.ptr(out),
                                         _lib.stream_ptr()), 'smx_legendre_q_half')
        got = out.cpu().numpy()
        ref = np.array([float(mp.legenq(m - 0.5, 0, 1 + mp.mpf(d), type=3).real) for d in delta])
        assert np.max(np.abs(got - ref) / np.abs(ref)) < 5e-14, m
    fl = _lib.dev(np.array([0.0, 1e-14, 4e-13]))
    for m in (0, 7, 12):
        _lib.check(L.smx_legendre_q_half(m, _lib.ptr(fl), 3, _lib.ptr(out), _lib.stream_ptr()),
                   'smx_legendre_q_half')
        got = out[:3].cpu().numpy()
        ref = OC.legendre_q_half(m, 1.0 + np.array([0.0, 1e-14, 4e-13]))
        assert np.max(np.abs(got - ref) / np.abs(ref)) < 1e-12, m


@pytest.fixture(scope="module")
def small(gpu):
    from schwarmax_amd import cylspline as CS
    from oracle import cylspline as OC
    args = (12, 9, 0.05, 20.0, 0.05, 15.0, 4, 16, 400)
    return CS.get_phi_m(plummer_rho, None, *args), OC.get_phi_m(plummer_rho, None, *args)


def test_builder_matches_oracle(small):
    """The tables (Phi_m and their spline second derivatives) of the device
    builder against the oracle's on the same density and grids."""
    got, ref = small
    for k in ('Rgrid', 'Zgrid', 'R0', 'm'):
        assert np.array_equal(np.asarray(got[k]), np.asarray(ref[k])), k
    scale = np.abs(ref['Phi_m_real']).max()
    for k in ('Phi_m_real', 'Phi_m_img'):
        assert np.max(np.abs(got[k] - ref[k])) <= 1e-10 * scale, k
    for k in ('Mx_real', 'My_real', 'Mx_img', 'My_img'):
        assert np.max(np.abs(got[k] - ref[k])) <= 1e-9 * np.abs(ref[k]).max(), k


def test_evaluator_matches_oracle(small):
    """evaluate_phi, evaluate_phi_axisymmetric and get_acc on the oracle's own
    tables at points inside, on the axis, at z = 0, below the plane and
    outside the table."""
    from schwarmax_amd import cylspline as CS
    from oracle import cylspline as OC
    _, ref = small
    d = dict(ref)
    rng = np.random.default_rng(2)
    n = 500
    R, ph, z = rng.uniform(0.0, 25.0, n), rng.uniform(0, 2 * np.pi, n), rng.uniform(-18, 18, n)
    R[:5], z[5:10] = 0.0, 0.0
    x, y = R * np.cos(ph), R * np.sin(ph)
    p = CS.evaluate_phi(x, y, z, d).cpu().numpy()
    pe = OC.evaluate_phi(x, y, z, ref)
    assert np.max(np.abs(p - pe)) <= 1e-13 * np.abs(pe).max()
    assert np.all(p[np.abs(pe) == 0] == 0) and np.any(pe == 0)
    p0 = CS.evaluate_phi_axisymmetric(x, y, z, d).cpu().numpy()
    assert np.max(np.abs(p0 - OC.evaluate_phi(x, y, z, ref, modes=1))) <= 1e-13 * np.abs(pe).max()
    a = CS.get_acc(x, y, z, d).cpu().numpy()
    ae = OC.get_acc(x, y, z, ref)
    assert a.shape == ae.shape == (3, n)
    assert np.max(np.abs(a - ae)) <= 1e-11 * np.abs(ae).max()


def test_density_matches_oracle(small):
    """Analytic nabla^2 Phi/(4 pi G) against the oracle's central differences
    of its analytic force (h = 1e-5), away from the axis."""
    from schwarmax_amd import cylspline as CS
    from oracle import cylspline as OC
    _, ref = small
    rng = np.random.default_rng(3)
    R, ph, z = rng.uniform(0.3, 8.0, 200), rng.uniform(0, 2 * np.pi, 200), rng.uniform(-5, 5, 200)
    x, y = R * np.cos(ph), R * np.sin(ph)
    r = CS.get_density(x, y, z, dict(ref)).cpu().numpy()
    re = OC.get_density(x, y, z, ref)
    assert np.max(np.abs(r - re)) <= 1e-4 * np.abs(re).max()


def test_plummer_high_resolution(gpu):
    """A 48 x 40 grid, 10 modes, 101^2 Simpson nodes: Phi to 1e-3, -grad Phi
    to 2e-2 and the Poisson density to 0.1 of their largest values against
    the closed forms (the reference's quadrature converges slowly)."""
    from schwarmax_amd import cylspline as CS
    d = CS.get_phi_m(plummer_rho, None, 48, 40, 0.02, 30.0, 0.02, 20.0, 10, 48, 101 ** 2)
    rng = np.random.default_rng(0)
    R, ph, z = rng.uniform(0.3, 6, 400), rng.uniform(0, 2 * np.pi, 400), rng.uniform(-4, 4, 400)
    x, y = R * np.cos(ph), R * np.sin(ph)
    pe = plummer_phi(x, y, z)
    p = CS.evaluate_phi(x, y, z, d).cpu().numpy()
    a = CS.get_acc(x, y, z, d).cpu().numpy()
    ae = plummer_acc(x, y, z)
    rho = CS.get_density(x, y, z, d).cpu().numpy()
    re = plummer_rho(x, y, z)
    e = (np.max(np.abs(p - pe) / np.abs(pe)), np.max(np.abs(a - ae)) / np.abs(ae).max(),
         np.max(np.abs(rho - re)) / re.max())
    print("plummer errors phi %.2e acc %.2e rho %.2e" % e, flush=True)
    assert e[0] < 1e-3 and e[1] < 2e-2 and e[2] < 0.1
