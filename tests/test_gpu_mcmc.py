"""The adaptive RMH driver on the GPU likelihood: a few steps of several chains
through model_likelihood_batch; every recorded log-density equals the box
prior plus model_likelihood at that position (to 1e-10)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
NAP = 12


@pytest.fixture(scope="module")
def galaxy():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from schwarmax_amd import synth, likelihood as L
    from oracle import model as OM, potential as OP
    theta, dd, sched = synth.make_galaxy('plain-2k', seed=21, n=24, n_ap=NAP, grid=(6, 4, 4), K=3,
                                         nxy=12)
    sched['N_max'] = 200
    sched['N_dynamical_time'] = 5
    ph, pb = L.theta_to_params(theta)
    Rl, Rn, xyl, xyn = L.grid_args(dd)
    blank = {'XY_density_data': np.ones(NAP), 'XY_density_data_err': np.full(NAP, 0.01)}
    for k in (1, 2, 3, 4):
        blank['h%d_data' % k] = np.zeros(NAP)
        blank['h%d_data_err' % k] = np.full(NAP, 0.03)
    _, lib = OM.build_model(OP.density_func, OP.potential_func, ph, pb, {**dd, **blank}, NAP,
                            dd['Rzphi_n_tot'], Rn, Rl, xyl, xyn, nnls_maxiter=2, schedule=sched,
                            return_library=True)
    synth.observe(dd, lib['surface_density'].T, [lib['h%d' % k].T for k in (1, 2, 3, 4)],
                  OM.density_target(dd, pb).sum() / lib['surface_density'].shape[0],
                  pb['mass_to_light_ratio'], floor=1e-3)
    return theta, dd, sched


def test_rmh_on_batched_likelihood(galaxy):
    from schwarmax_amd import mcmc, likelihood as L
    theta, dd, sched = galaxy
    f = mcmc.likelihood_batch_fn(dd, NAP, dd['Rzphi_n_tot'], schedule=sched, nnls_maxiter=50)
    lo, hi = theta - 0.3, theta + 0.3
    lo[7:10], hi[7:10] = theta[7:10] - 0.05, theta[7:10] + 0.05
    out = mcmc.run_adaptive_rmh(f, theta, lo, hi, np.full(12, 0.01), np.full(12, 0.02),
                                n_chains=4, n_steps=3, burnin=0, seed=5)
    assert out['chain'].shape == (3, 4, 12)
    assert 4 <= out['n_likelihood_evals'] <= 16
    log_vol = np.sum(np.log(hi - lo))
    for c in range(4):
        th = out['chain'][-1, c]
        ref = float(L.model_likelihood(th, dd, NAP, dd['Rzphi_n_tot'], schedule=sched,
                                       nnls_maxiter=50).item()) - log_vol
        got = out['logprob'][-1, c]
        print("chain", c, got, ref, flush=True)
        # the batched ICs come from their own kernel; last-bit differences grow
        # along the orbits to ~1e-12 of logL
        assert abs(got - ref) <= 1e-10 * abs(ref), (c, got, ref)
