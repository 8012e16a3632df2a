"""Adaptive RMH driver (CPU, cheap likelihoods): the notebook's sampler loop
(example_schwarmax.ipynb, MCMC cell) with batched likelihood calls."""
import numpy as np
import pytest

from schwarmax_amd import mcmc, prng

D = 3
MU = np.array([0.3, -1.0, 2.0])
SIG = np.array([[0.04, 0.012, 0.0], [0.012, 0.09, -0.02], [0.0, -0.02, 0.25]])
SIG_INV = np.linalg.inv(SIG)


def gauss(thetas):
    d = np.asarray(thetas) - MU
    return -0.5 * np.einsum('ij,jk,ik->i', d, SIG_INV, d)


def test_gaussian_target():
    """Samples of a correlated Gaussian: mean within 0.1 sigma, covariance
    within 20%, acceptance near the random-walk optimum after adaptation."""
    lo, hi = MU - 10, MU + 10
    out = mcmc.run_adaptive_rmh(gauss, MU + 0.5, lo, hi, np.full(D, 0.2), np.full(D, 0.1),
                                n_chains=64, n_steps=700, adapt_every=50, adapt_after=150,
                                burnin=300, seed=3)
    s = out['flat_samples']
    assert s.shape == (400 * 64, D)
    sd = np.sqrt(np.diag(SIG))
    assert np.all(np.abs(s.mean(0) - MU) < 0.1 * sd), s.mean(0)
    C = np.cov(s.T)
    assert np.all(np.abs(C - SIG) <= 0.2 * np.sqrt(np.outer(np.diag(SIG), np.diag(SIG)))), C
    assert out['adapt_count'] == 12             # steps 150, 200, ..., 700
    assert 0.15 < out['acceptance'][-100:].mean() < 0.6


def test_key_flow_and_first_step():
    """Flat likelihood inside the box: every in-box proposal is accepted, and
    the first step moves chain c by scale * normal(split(split(step_key, C)[c])[0])."""
    C = 5
    lo, hi = np.full(D, -100.0), np.full(D, 100.0)
    th0 = np.zeros(D)
    s_init, s_rmh = np.array([0.1, 0.2, 0.3]), np.array([0.05, 0.5, 1.0])
    out = mcmc.run_adaptive_rmh(lambda t: np.zeros(len(t)), th0, lo, hi, s_init, s_rmh,
                                n_chains=C, n_steps=1, burnin=0, seed=7)
    key = prng.PRNGKey(7)
    key, init_key = prng.split(key)
    start = th0 + prng.normal(init_key, (C, D)).astype(np.float64) * s_init
    key, step_key = prng.split(key)
    keys = prng.split(step_key, C)
    expect = np.stack([start[c] + s_rmh * prng.normal(prng.split(keys[c])[0], (D,)) for c in range(C)])
    assert np.array_equal(out['chain'][0], expect)
    assert np.all(out['acceptance'] == 1.0)


def test_out_of_bounds_not_evaluated():
    """Proposals outside the box get prior -inf, are rejected, and never
    reach the likelihood; the batch holds only in-box chains."""
    seen = []

    def f(t):
        seen.append(np.array(t))
        return gauss(t)
    lo, hi = MU - 0.3, MU + 0.3
    out = mcmc.run_adaptive_rmh(f, MU, lo, hi, np.full(D, 0.1), np.full(D, 0.4), n_chains=16,
                                n_steps=30, burnin=0, seed=1)
    allt = np.concatenate(seen)
    assert np.all((allt >= lo) & (allt <= hi))
    assert out['n_likelihood_evals'] == allt.shape[0] < 16 * 31
    assert np.all((out['chain'] >= lo) & (out['chain'] <= hi))


def test_nonfinite_loglike_is_floor():
    ld = mcmc.LogDensity(lambda t: np.array([np.nan, -np.inf, 1.0]), np.zeros(2), np.ones(2) * 2)
    v = ld(np.full((3, 2), 1.0))
    assert np.array_equal(v, np.array([-1e30, -1e30, 1.0]) - 2 * np.log(2.0))
    assert np.isneginf(ld(np.array([[3.0, 1.0]]))[0])


def test_adapted_factor():
    """L L^T = (2.38^2/d)(0.8 C + 0.2 diag C) over positions from step
    max(adapt_after - 50, step - 200) on; None with <= 2 d samples."""
    rng = np.random.default_rng(0)
    pos = [rng.multivariate_normal(MU, SIG, size=8) for _ in range(400)]
    pos[:200] = [p + 100.0 for p in pos[:200]]          # outside the window: must be ignored
    L, n = mcmc.adapted_factor(pos, 400, 150, D)
    assert n == 200 * 8
    flat = np.concatenate(pos[200:])
    C = np.cov(flat.T)
    ref = (2.38 ** 2 / D) * (0.8 * C + 0.2 * np.diag(np.diag(C)))
    assert np.allclose(L @ L.T, ref, rtol=1e-12, atol=0)
    assert np.allclose(np.triu(L, 1), 0)
    assert mcmc.adapted_factor(pos[:1], 1, 150, D)[0] is None


def test_checkpoint_resume_is_exact(tmp_path):
    """Stopping at a checkpoint and resuming gives the uninterrupted chain."""
    ck = str(tmp_path / "ck.npz")
    kw = dict(n_chains=6, adapt_every=5, adapt_after=5, burnin=0, seed=11)
    args = (gauss, MU, MU - 5, MU + 5, np.full(D, 0.2), np.full(D, 0.1))
    full = mcmc.run_adaptive_rmh(*args, n_steps=20, **kw)
    mcmc.run_adaptive_rmh(*args, n_steps=10, checkpoint_file=ck, checkpoint_every=5, **kw)
    with np.load(ck, allow_pickle=False) as z:
        assert int(z['step']) == 10 and z['all_samples'].shape == (10, 6, D)
    res = mcmc.run_adaptive_rmh(*args, n_steps=20, checkpoint_file=ck, checkpoint_every=5,
                                resume=True, **kw)
    assert np.array_equal(res['chain'], full['chain'])
    assert np.array_equal(res['logprob'], full['logprob'])


def test_bounds_validation():
    with pytest.raises(ValueError):
        mcmc.LogDensity(gauss, np.ones(D), np.ones(D))
