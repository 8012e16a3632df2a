"""Loader-made dict_data on the GPU (load_data_bootstrap, utils.py:572-743):
without 'jeans_normals'/'jitter_uniforms' the model takes the reference's
own JAX-keyed draws (model.py:112-138), and the value matches the oracle
given those draws explicitly."""
import numpy as np
import pytest

from test_data import make_bin_dict

pytestmark = pytest.mark.gpu
N = 64


@pytest.fixture(scope="module")
def loaded(tmp_path_factory):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from schwarmax_amd import data as D
    d = tmp_path_factory.mktemp("bins")
    bd = make_bin_dict(seed=5)
    D.write_bin_dict(str(d / "galaxy.npz"), bd)
    dd = D.load_data_bootstrap(str(d) + "/", "galaxy.npz", N_BOOTSTRAP=4, n_samples=N)
    return int(bd['total_bins']), dd, d


SCHED = dict(N_max=300, N_dynamical_time=3)


def test_default_draws_match_oracle(loaded):
    from schwarmax_amd import likelihood as L, prng, synth
    from oracle import model as OM
    B, dd, _ = loaded
    theta = synth.base_theta()
    assert 'jeans_normals' not in dd and 'jitter_uniforms' not in dd
    got = float(L.model_likelihood(theta, dd, B, 600, schedule=SCHED, nnls_maxiter=100).item())
    nrm, uni = prng.reference_draws(N, 4)
    ref = OM.model_likelihood(theta, {**dd, 'jeans_normals': nrm, 'jitter_uniforms': uni}, B, 600,
                              schedule=SCHED, nnls_maxiter=100)
    print("loader dict logL", got, ref, flush=True)
    assert np.isfinite(ref) and abs(got - ref) <= 1e-4 * abs(ref), (got, ref)


def test_saved_dict_data_and_prefix_draws(loaded):
    """A dict_data saved and reloaded evaluates identically; a batch model on
    the first 40 orbits equals model_likelihood on a 40-orbit galaxy (the
    draws of the first n orbits are the draws of an n-orbit galaxy)."""
    from schwarmax_amd import data as D, likelihood as L, batch, synth
    B, dd, d = loaded
    theta = synth.base_theta()
    D.save_dict_data(str(d / "dd.npz"), dd)
    dd2 = D.load_dict_data(str(d / "dd.npz"))
    a = float(L.model_likelihood(theta, dd, B, 600, schedule=SCHED, nnls_maxiter=50).item())
    b = float(L.model_likelihood(theta, dd2, B, 600, schedule=SCHED, nnls_maxiter=50).item())
    assert a == b
    got = batch.model_likelihood_batch(np.stack([theta, theta]), dd, B, 600, n_orbits=[N, 40],
                                       schedule=SCHED, nnls_maxiter=50).cpu().numpy()
    sub = dict(dd)
    sub['w0'] = dd['w0'][:40]
    ref40 = float(L.model_likelihood(theta, sub, B, 600, schedule=SCHED, nnls_maxiter=50).item())
    assert abs(got[0] - a) <= 1e-12 * abs(a), (got[0], a)
    assert abs(got[1] - ref40) <= 1e-12 * abs(ref40), (got[1], ref40)
