"""End-to-end build_model / model_likelihood / get_model_with_orbit parity (GPU)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def galaxy():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from schwarmax_amd import synth, likelihood as L
    from oracle import model as OM, potential as OP
    theta, dd, sched = synth.make_galaxy('plain-2k', seed=11, n=40, n_ap=24, grid=(8, 4, 6), K=5,
                                         nxy=16)
    sched['N_max'] = 500
    ph, pb = L.theta_to_params(theta)
    Rl, Rn, xyl, xyn = L.grid_args(dd)
    blank = {'XY_density_data': np.ones(24), 'XY_density_data_err': np.full(24, 0.01)}
    for k in (1, 2, 3, 4):
        blank['h%d_data' % k] = np.zeros(24)
        blank['h%d_data_err' % k] = np.full(24, 0.03)
    _, lib = OM.build_model(OP.density_func, OP.potential_func, ph, pb, {**dd, **blank}, 24,
                            dd['Rzphi_n_tot'], Rn, Rl, xyl, xyn, nnls_maxiter=2, schedule=sched,
                            return_library=True)
    A_xy = lib['surface_density'].T
    A_h = [lib['h%d' % k].T for k in (1, 2, 3, 4)]
    y_R = OM.density_target(dd, pb)
    synth.observe(dd, A_xy, A_h, y_R.sum() / A_xy.shape[1], pb['mass_to_light_ratio'])
    return theta, dd, sched, ph, pb


def test_build_model_outputs(galaxy):
    from schwarmax_amd import model as M, likelihood as L
    from oracle import model as OM, potential as OP
    theta, dd, sched, ph, pb = galaxy
    Rl, Rn, xyl, xyn = L.grid_args(dd)
    ref = OM.build_model(OP.density_func, OP.potential_func, ph, pb, dd, 24, dd['Rzphi_n_tot'], Rn,
                         Rl, xyl, xyn, nnls_maxiter=200, schedule=sched)
    got = M.build_model(L.density_func, L.potential_func, ph, pb, dd, 24, dd['Rzphi_n_tot'], Rn, Rl,
                        xyl, xyn, nnls_maxiter=200, schedule=sched)
    assert set(got) == set(ref)
    d3 = got['density_3d'].cpu().numpy()
    assert np.max(np.abs(d3 - ref['density_3d']) / np.abs(ref['density_3d']).max()) < 1e-12
    for k in ('density_3d_model', 'surface_density', 'h1', 'h2', 'h3', 'h4', 'V', 'sigma'):
        g = got[k].cpu().numpy()
        r = ref[k]
        assert g.shape == r.shape, k
        err = np.linalg.norm(g - r) / max(np.linalg.norm(r), 1e-300)
        assert err < 1e-3, (k, err)


def test_model_likelihood(galaxy):
    from schwarmax_amd import likelihood as L
    from oracle import model as OM
    theta, dd, sched, ph, pb = galaxy
    ref = OM.model_likelihood(theta, dd, 24, dd['Rzphi_n_tot'], schedule=sched, nnls_maxiter=200)
    got = float(L.model_likelihood(theta, dd, 24, dd['Rzphi_n_tot'], schedule=sched,
                                   nnls_maxiter=200).item())
    assert abs(got - ref) / abs(ref) < 1e-4, (got, ref)


def test_get_model_with_orbit(galaxy):
    from schwarmax_amd import model as M, likelihood as L
    from oracle import model as OM, potential as OP
    theta, dd, sched, ph, pb = galaxy
    Rl, Rn, xyl, xyn = L.grid_args(dd)
    ref = OM.get_model_with_orbit(OP.density_func, OP.potential_func, ph, pb, dd, 24,
                                  dd['Rzphi_n_tot'], Rn, Rl, xyl, xyn, nnls_maxiter=100,
                                  schedule=sched)
    got = M.get_model_with_orbit(L.density_func, L.potential_func, ph, pb, dd, 24, dd['Rzphi_n_tot'],
                                 Rn, Rl, xyl, xyn, nnls_maxiter=100, schedule=sched)
    assert set(got) == set(ref)
    yt = got['y_traj'].cpu().numpy()
    assert yt.shape == ref['y_traj'].shape
    # trajectories of the same ICs: per-orbit relative error, most orbits to 1e-8
    e = np.abs(yt - ref['y_traj']).reshape(yt.shape[0], -1).max(axis=1) / \
        np.abs(ref['y_traj']).reshape(yt.shape[0], -1).max(axis=1)
    assert np.quantile(e, 0.9) < 1e-8, np.sort(e)[-5:]
    tt = got['t_traj'].cpu().numpy()
    assert np.quantile(np.abs(tt[:, 0, -1] - ref['t_traj'][:, 0, -1]) / ref['t_traj'][:, 0, -1], 0.9) < 1e-8
