"""Pins of the oracle's gradient (CPU): central differences of model_likelihood
with the discrete decisions held fixed (oracle.model.model_likelihood_grad_fd)."""
import numpy as np
import pytest

NAP = 10


@pytest.fixture(scope="module")
def tiny():
    from schwarmax_amd import synth, likelihood as L
    from oracle import model as OM, potential as OP
    theta, dd, sched = synth.make_galaxy('plain-2k', seed=2, n=6, n_ap=NAP, grid=(5, 4, 4), K=3,
                                         nxy=10)
    sched['N_max'] = 100
    sched['N_dynamical_time'] = 2
    ph, pb = L.theta_to_params(theta)
    Rl, Rn, xyl, xyn = L.grid_args(dd)
    blank = {'XY_density_data': np.ones(NAP), 'XY_density_data_err': np.full(NAP, 0.01)}
    for k in (1, 2, 3, 4):
        blank['h%d_data' % k] = np.zeros(NAP)
        blank['h%d_data_err' % k] = np.full(NAP, 0.03)
    _, lib = OM.build_model(OP.density_func, OP.potential_func, ph, pb, {**dd, **blank}, NAP,
                            dd['Rzphi_n_tot'], Rn, Rl, xyl, xyn, nnls_maxiter=2, schedule=sched,
                            return_library=True)
    A_xy = lib['surface_density'].T
    A_h = [lib['h%d' % k].T for k in (1, 2, 3, 4)]
    y_R = OM.density_target(dd, pb)
    synth.observe(dd, A_xy, A_h, y_R.sum() / A_xy.shape[1], pb['mass_to_light_ratio'])
    assert lib['n_valid'].sum() > 0
    return theta, dd, sched


def _ll(theta, dd, sched, **kw):
    from oracle import model as OM
    return OM.model_likelihood(theta, dd, NAP, dd['Rzphi_n_tot'], schedule=sched, nnls_maxiter=20,
                               **kw)


def test_replay_reproduces_the_recorded_run(tiny):
    theta, dd, sched = tiny
    tape = {'mode': 'record'}
    a, W = _ll(theta, dd, sched, tape=tape, return_weights=True)
    tape['mode'] = 'replay'
    b = _ll(theta, dd, sched, tape=tape, W_fixed=W)
    assert a == b
    assert len(tape['accept']) == sched['N_max'] and tape['valid'].shape == (6 * 4,)


def _d4(f, h):
    return (8.0 * (f(h) - f(-h)) - (f(2 * h) - f(-2 * h))) / (12.0 * h)


def test_frozen_differences_converge(tiny):
    """With every branch of the base run replayed the map is smooth: the
    fourth-order differences at h and 10 h agree to the map's roundoff
    (~1e-10 relative, from the Jeans moments' nested differences) / h."""
    theta, dd, sched = tiny
    tape = {'mode': 'record'}
    f0, W = _ll(theta, dd, sched, tape=tape, return_weights=True)
    tape['mode'] = 'replay'
    for i in (0, 5, 6, 11):
        e = np.zeros(12)
        e[i] = 1.0
        f = lambda h: _ll(theta + h * e, dd, sched, tape=tape, W_fixed=W)  # noqa: E731
        d1, d2 = _d4(f, 1e-5), _d4(f, 1e-4)
        assert d1 != 0.0
        assert abs(d1 - d2) <= 1e-4 * abs(d1), (i, d1, d2)


def test_step_decisions_replay_exactly():
    """The replayed step equals the free step on its own branch (same
    arithmetic), so recording changes nothing."""
    from oracle import integrate as I
    rng = np.random.default_rng(0)

    def acc(x, y, z):
        return -x, -2 * y, -3 * z
    y = rng.standard_normal((6, 40))
    k1 = I.deriv_barred(y, acc, 0.3)
    for dt in (1e-6, 1e-2, 0.5, 5.0):
        a = I.rk_step_dec(np.zeros(40), y, np.full(40, dt), k1, acc, 0.3, 1e-7, 1e-4, 1e-5, 0.3)
        b = I.rk_step_dec(np.zeros(40), y, np.full(40, dt), k1, acc, 0.3, 1e-7, 1e-4, 1e-5, 0.3,
                          dec_in=a[5])
        for u, v in zip(a[:5], b[:5]):
            assert np.array_equal(u, v)


def test_alpha_enters_through_membership_only(tiny):
    """v_los uses the rotation's third row, R[2] = Rz(gamma) Rx(beta) e_z, which
    does not depend on alpha; alpha moves sky positions, i.e. bin membership,
    which jax.grad holds fixed: d logL / d alpha = 0 exactly."""
    theta, dd, sched = tiny
    tape = {'mode': 'record'}
    _, W = _ll(theta, dd, sched, tape=tape, return_weights=True)
    tape['mode'] = 'replay'
    e = np.zeros(12)
    e[7] = 1e-4
    assert _ll(theta + e, dd, sched, tape=tape, W_fixed=W) == _ll(theta - e, dd, sched, tape=tape,
                                                                   W_fixed=W)
