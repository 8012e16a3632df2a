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
    from galaxies import observed_galaxy
    theta, dd, sched, _, _ = observed_galaxy(seed=21, n=24, n_ap=NAP, grid=(6, 4, 4), K=3, nxy=12,
                                             N_max=200)
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
    # and against the oracle: the last position of chain 0 (a proposal the
    # batched likelihood evaluated, away from theta)
    from galaxies import oracle_likelihood
    th = out['chain'][-1, 0]
    ref_o = oracle_likelihood(th, dd, NAP, sched, 50)[0] - log_vol
    got = out['logprob'][-1, 0]
    print("chain 0 vs oracle", got, ref_o, flush=True)
    assert abs(got - ref_o) <= 1e-4 * abs(ref_o), (got, ref_o)
