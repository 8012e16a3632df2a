"""Adaptive random-walk Metropolis over many chains, with one batched
likelihood evaluation per step.

The reference's sampler (example/example_schwarmax.ipynb, the "MCMC sampling"
cell) runs blackjax's additive random walk on N_CHAINS chains under jax.vmap
(two half-batches, to bound memory).  It uses a uniform box prior, the
log-likelihood with non-finite values replaced by -1e30, and a proposal
N(0, diag(rmh_sigma_init)^2) that becomes L L^T every ADAPT_EVERY steps once
step >= ADAPT_AFTER.  There L is the Cholesky factor of
(2.38^2 / d) (0.8 C + 0.2 diag C), with C the sample covariance of the
positions from step max(ADAPT_AFTER - 50, step - 200) on.  Chains start at the
best fit plus N(0, initial_guess_sigma^2), clipped to the box.

Here every step proposes for all chains and evaluates the in-bounds
proposals together through ``model_likelihood_batch``; those outside the box
have prior -inf and are rejected without a likelihood call, so a batch holds
only the chains that need one.  The random keys follow the notebook and
blackjax's random-walk kernel:
* ``rng_key, step_key = split(rng_key)``;
* one key per chain from ``split(step_key, n_chains)``;
* each chain splits its key into a proposal key (a normal draw mapped by the
  proposal scale) and an acceptance key (accept iff uniform < min(1, exp(dlogp))).
All draws use the NumPy restatement of JAX's generator (``prng``).
Checkpoints are ``.npz`` archives, not pickles.
"""
import os

import numpy as np

from . import prng

OPTIMAL_SCALE_NUMERATOR = 2.38
PARAM_NAMES = ('logM_10kpc', 'logM_disk', 'logM_bar', 'logC_halo', 'logRs_disk', 'logHs_disk',
               'logL_bar', 'alpha', 'beta', 'gamma', 'log_mass_to_light_ratio', 'log_Omega')


def likelihood_batch_fn(dict_data, num_Vbin, Rzphi_n_tot, sigma_amplify=1.0, **kwargs):
    """thetas (M, 12) -> logL (M,) as a NumPy array, through one
    model_likelihood_batch call (keyword arguments pass through)."""
    from .batch import model_likelihood_batch

    def f(thetas):
        thetas = np.asarray(thetas, dtype=np.float64).reshape(-1, 12)
        if thetas.shape[0] == 0:
            return np.zeros(0)
        return model_likelihood_batch(thetas, dict_data, num_Vbin, Rzphi_n_tot,
                                      sigma_amplify=sigma_amplify, **kwargs).cpu().numpy()
    return f


class LogDensity:
    """Box prior + log-likelihood (non-finite -> -1e30), batched; counts the
    likelihood evaluations it makes."""

    def __init__(self, loglike_batch, lo, hi):
        self.f = loglike_batch
        self.lo = np.asarray(lo, dtype=np.float64)
        self.hi = np.asarray(hi, dtype=np.float64)
        if np.any(self.hi <= self.lo):
            raise ValueError("every upper bound must exceed its lower bound")
        self.log_vol = float(np.sum(np.log(self.hi - self.lo)))
        self.n_evals = 0

    def __call__(self, thetas):
        thetas = np.asarray(thetas, dtype=np.float64)
        inside = np.all((thetas >= self.lo) & (thetas <= self.hi), axis=1)
        out = np.full(thetas.shape[0], -np.inf)
        if inside.any():
            ll = np.asarray(self.f(thetas[inside]), dtype=np.float64)
            self.n_evals += int(inside.sum())
            out[inside] = -self.log_vol + np.where(np.isfinite(ll), ll, -1e30)
        return out


def _scale_apply(scale, z):
    """diag (1-d scale) or dense lower factor (2-d) times the rows of z."""
    scale = np.asarray(scale, dtype=np.float64)
    return z * scale if scale.ndim == 1 else z @ scale.T


def rmh_step(step_key, positions, logdensity, scale, logdensity_fn, partitionable=True):
    """One step of every chain; returns (positions, logdensity, accepted, p_accept)."""
    C, D = positions.shape
    keys = prng.split(step_key, C, partitionable)
    z = np.empty((C, D))
    u = np.empty(C)
    for c in range(C):
        k_prop, k_acc = prng.split(keys[c], 2, partitionable)
        z[c] = prng.normal(k_prop, (D,), partitionable)
        u[c] = prng.uniform(k_acc, (), partitionable=partitionable)
    proposed = positions + _scale_apply(scale, z)
    lp_new = logdensity_fn(proposed)
    with np.errstate(invalid='ignore', over='ignore'):
        p_accept = np.minimum(np.exp(lp_new - logdensity), 1.0)
    p_accept = np.where(np.isnan(p_accept), 0.0, p_accept)
    accept = u < p_accept
    return (np.where(accept[:, None], proposed, positions), np.where(accept, lp_new, logdensity),
            accept, p_accept)


def adapted_factor(all_positions, step, adapt_after, ndim):
    """Cholesky factor of (2.38^2/d)(0.8 C + 0.2 diag C) over the positions
    from step max(adapt_after - 50, step - 200) on; None when there are too
    few samples (<= 2 d) or C + regularisation is not positive definite."""
    start = int(max(adapt_after - 50, step - 200))
    window = all_positions[start:]
    if not window:
        return None, 0
    flat = np.stack(window, axis=0).reshape(-1, ndim)
    if flat.shape[0] <= 2 * ndim:
        return None, flat.shape[0]
    cov = np.cov(flat.T)
    reg = 0.8 * cov + 0.2 * np.diag(np.diag(cov))
    try:
        return np.linalg.cholesky((OPTIMAL_SCALE_NUMERATOR ** 2 / ndim) * reg), flat.shape[0]
    except np.linalg.LinAlgError:
        return None, flat.shape[0]


def save_checkpoint(filename, all_positions, all_logprob, step, rng_key, adapt_count,
                    proposal_L, scale_factor):
    np.savez(filename, all_samples=np.stack(all_positions), all_logprob=np.stack(all_logprob),
             step=np.array(step), rng_key=np.asarray(rng_key, dtype=np.uint32),
             adapt_count=np.array(adapt_count),
             proposal_L=np.zeros((0, 0)) if proposal_L is None else np.asarray(proposal_L),
             scale_factor=np.array(scale_factor))


def load_checkpoint(filename):
    if not os.path.exists(filename):
        return None
    with np.load(filename, allow_pickle=False) as z:
        ck = {k: z[k] for k in z.files}
    L = ck['proposal_L']
    return dict(all_samples=list(ck['all_samples']), all_logprob=list(ck['all_logprob']),
                step=int(ck['step']), rng_key=ck['rng_key'].astype(np.uint32),
                adapt_count=int(ck['adapt_count']), proposal_L=None if L.size == 0 else L,
                scale_factor=float(ck['scale_factor']))


def run_adaptive_rmh(loglike_batch, theta0, bounds_lo, bounds_hi, initial_guess_sigma,
                     rmh_sigma_init, n_chains=32, n_steps=1000, adapt_every=50, adapt_after=150,
                     burnin=300, seed=42, checkpoint_file=None, checkpoint_every=5, resume=True,
                     partitionable=True, log=None):
    """The notebook's adaptive RMH (fresh start or resume from checkpoint_file).

    loglike_batch: thetas (M, d) -> logL (M,), e.g. ``likelihood_batch_fn(...)``.
    Returns dict(chain (S, C, d), logprob (S, C), flat_samples, acceptance (S,),
    n_likelihood_evals, adapt_count).
    """
    theta0 = np.asarray(theta0, dtype=np.float64)
    D = theta0.shape[0]
    logdensity_fn = LogDensity(loglike_batch, bounds_lo, bounds_hi)
    say = log or (lambda s: None)
    ck = load_checkpoint(checkpoint_file) if (resume and checkpoint_file) else None
    if ck is not None:
        all_positions, all_logprob = ck['all_samples'], ck['all_logprob']
        start, rng_key = ck['step'], ck['rng_key']
        adapt_count, proposal_L, scale_factor = ck['adapt_count'], ck['proposal_L'], ck['scale_factor']
        positions = np.asarray(all_positions[-1], dtype=np.float64)
        logd = logdensity_fn(positions)
        scale = rmh_sigma_init if proposal_L is None else proposal_L * scale_factor
        say("resumed from step %d" % start)
    else:
        rng_key = prng.PRNGKey(seed)
        rng_key, init_key = prng.split(rng_key, 2, partitionable)
        noise = prng.normal(init_key, (n_chains, D), partitionable).astype(np.float64)
        positions = np.clip(theta0[None, :] + noise * np.asarray(initial_guess_sigma)[None, :],
                            logdensity_fn.lo, logdensity_fn.hi)
        logd = logdensity_fn(positions)
        all_positions, all_logprob = [], []
        start, adapt_count, proposal_L, scale_factor = 0, 0, None, 1.0
        scale = np.asarray(rmh_sigma_init, dtype=np.float64)
    recent, acceptance = [], []
    for step in range(start + 1, n_steps + 1):
        rng_key, step_key = prng.split(rng_key, 2, partitionable)
        positions, logd, acc, p_acc = rmh_step(step_key, positions, logd, scale, logdensity_fn,
                                               partitionable)
        all_positions.append(positions.copy())
        all_logprob.append(logd.copy())
        acceptance.append(float(np.mean(p_acc)))
        recent = (recent + [acceptance[-1]])[-adapt_every:]
        if step >= adapt_after and step % adapt_every == 0:
            L, n_used = adapted_factor(all_positions, step, adapt_after, D)
            if L is not None:
                proposal_L = L
                scale = L * scale_factor
                adapt_count += 1
                say("adapt #%d at step %d: recent acceptance %.3f, n = %d"
                    % (adapt_count, step, float(np.mean(recent)), n_used))
            else:
                say("adapt at step %d skipped (n = %d)" % (step, n_used))
        if checkpoint_file and step % checkpoint_every == 0:
            save_checkpoint(checkpoint_file, all_positions, all_logprob, step, rng_key,
                            adapt_count, proposal_L, scale_factor)
    chain = np.stack(all_positions) if all_positions else np.zeros((0, n_chains, D))
    logprob = np.stack(all_logprob) if all_logprob else np.zeros((0, n_chains))
    flat = chain[burnin:].reshape(-1, D) if chain.shape[0] > burnin else chain.reshape(-1, D)
    return dict(chain=chain, logprob=logprob, flat_samples=flat,
                acceptance=np.asarray(acceptance), n_likelihood_evals=logdensity_fn.n_evals,
                adapt_count=adapt_count, rng_key=rng_key)
