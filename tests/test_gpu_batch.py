"""Batched evaluation (GPU): smx_build_library_batch / model_likelihood_batch
against per-model model_likelihood on the same orbits (uneven orbit counts and
step counts, different theta)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NAP = 24


@pytest.fixture(scope="module")
def setup():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from schwarmax_amd import synth, likelihood as L
    from oracle import model as OM, potential as OP
    theta, dd, sched = synth.make_galaxy('plain-2k', seed=11, n=40, n_ap=NAP, grid=(8, 4, 6), K=5,
                                         nxy=16)
    sched['N_max'] = 500
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
    synth.observe(dd, A_xy, A_h, y_R.sum() / A_xy.shape[1], pb['mass_to_light_ratio'], floor=1e-3)
    rng = np.random.default_rng(8)
    thetas = np.stack([theta] + synth.batched_thetas(2, rng))
    return thetas, dd, sched


def _first(dd, n):
    out = dict(dd)
    for k in ('w0', 'jeans_normals', 'jitter_uniforms'):
        out[k] = np.asarray(dd[k])[:n]
    return out


def test_batch_matches_per_model(setup):
    from schwarmax_amd import batch, likelihood as L, model as M
    thetas, dd, sched = setup
    n_orbits = [40, 25, 33]
    s2 = dict(sched)
    s2['N_max'] = 300                    # uneven step counts
    scheds = [sched, sched, s2]
    got, aux = batch.model_likelihood_batch(thetas, dd, NAP, dd['Rzphi_n_tot'], n_orbits=n_orbits,
                                            schedule=scheds, nnls_maxiter=100, return_aux=True)
    got = got.cpu().numpy()
    Rl, Rn, xyl, xyn = L.grid_args(dd)
    for m in range(3):
        dm = _first(dd, n_orbits[m])
        ref = float(L.model_likelihood(thetas[m], dm, NAP, dd['Rzphi_n_tot'], schedule=scheds[m],
                                       nnls_maxiter=100))
        assert abs(got[m] - ref) <= 1e-12 * abs(ref), (m, got[m], ref)
        ph, pb = L.theta_to_params(thetas[m])
        _, a = M.build_model(L.density_func, L.potential_func, ph, pb, dm, NAP, dd['Rzphi_n_tot'],
                             Rn, Rl, xyl, xyn, nnls_maxiter=2, schedule=scheds[m], return_aux=True)
        lib_ref = a['libd']['lib'].cpu().numpy()
        lib_b = aux[m]['lib'].cpu().numpy()
        assert lib_b.shape == lib_ref.shape
        assert np.max(np.abs(lib_b - lib_ref)) <= 1e-13 * np.max(np.abs(lib_ref)), m


def test_batch_groups_and_budget(setup):
    """A tiny library budget forces one model per group: same values."""
    from schwarmax_amd import batch
    thetas, dd, sched = setup
    a = batch.model_likelihood_batch(thetas, dd, NAP, dd['Rzphi_n_tot'], n_orbits=[40, 25, 33],
                                     schedule=sched, nnls_maxiter=50).cpu().numpy()
    b = batch.model_likelihood_batch(thetas, dd, NAP, dd['Rzphi_n_tot'], n_orbits=[40, 25, 33],
                                     schedule=sched, nnls_maxiter=50,
                                     library_budget_bytes=1).cpu().numpy()
    assert np.array_equal(a, b)
