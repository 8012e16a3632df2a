"""Batched paths at their benchmark shapes, by size-independent properties
(the oracle would take hours at these sizes; the same kernels are checked
against it at small sizes in test_gpu_batch.py and test_gpu_grad.py).

* batched evaluation (1,000 apertures, 32 x 24 x 16 grid, K = 100,
  T = 1,000): three theta of the batched workload with orbit counts 5,000,
  23,000 and 40,000 in one call; every logL finite and within 1e-6 of the
  single-model path on the same orbits.  (The batched path forms the initial
  conditions and integrates in its own kernels; their last-bit differences
  flip a few accept decisions among 20,000+ integrations of 2,000 attempts,
  which moves logL by ~1e-9.)
* batched forward + adjoint (40,000 orbits, 2,000 apertures, 40 x 20 x 32
  grid, K = 200): two theta; logL and gradients finite and equal to the
  single-model gradient to 1e-12.
Peak device memory of each call is printed.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _first(dd, n):
    out = dict(dd)
    for k in ('w0', 'jeans_normals', 'jitter_uniforms'):
        out[k] = np.asarray(dd[k])[:n]
    return out


def test_batched64_shape_matches_single_model():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from schwarmax_amd import synth, batch, likelihood as L
    theta, dd, sched, info = synth.observed_workload('batched-64', 0, n=40000)
    assert info['valid'] >= 0.8
    thetas = np.stack(synth.batched_thetas(3, np.random.default_rng(1)))
    n_orbits = [5000, 23000, 40000]
    B, P = int(dd['total_bins']), int(dd['Rzphi_n_tot'])
    torch.cuda.reset_peak_memory_stats()
    got = batch.model_likelihood_batch(thetas, dd, B, P, n_orbits=n_orbits, schedule=sched,
                                       nnls_maxiter=1000).cpu().numpy()
    peak = torch.cuda.max_memory_allocated() / 2 ** 30
    print(f"batched-64 shape, 3 models: logL {got}, peak device memory {peak:.1f} GiB", flush=True)
    assert np.all(np.isfinite(got))
    for m in range(3):
        one = float(L.model_likelihood(thetas[m], _first(dd, n_orbits[m]), B, P, schedule=sched,
                                       nnls_maxiter=1000))
        print(f"model {m}: n={n_orbits[m]} batched {got[m]:.10g} single {one:.10g} "
              f"rel {abs(got[m] - one) / abs(one):.1e}", flush=True)
        assert abs(got[m] - one) <= 1e-6 * abs(one), (m, got[m], one)


def test_batched_grad16_shape_matches_single_model():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from schwarmax_amd import synth, grad as GR
    theta, dd, sched, info = synth.observed_workload('batched-grad-16', 0, n=40000)
    assert info['valid'] >= 0.8
    thetas = np.stack(synth.batched_thetas(2, np.random.default_rng(2)))
    B, P = int(dd['total_bins']), int(dd['Rzphi_n_tot'])
    torch.cuda.reset_peak_memory_stats()
    v, g = GR.model_likelihood_and_grad_batch(thetas, dd, B, P, schedule=sched, nnls_maxiter=1000)
    peak = torch.cuda.max_memory_allocated() / 2 ** 30
    print(f"batched-grad-16 shape, 2 models: logL {v.tolist()}, peak device memory {peak:.1f} GiB",
          flush=True)
    assert torch.isfinite(v).all() and torch.isfinite(g).all()
    for m in range(2):
        v1, g1 = GR.model_likelihood_and_grad(thetas[m], dd, B, P, schedule=sched,
                                              nnls_maxiter=1000)
        assert abs(float(v[m]) - float(v1)) <= 1e-12 * abs(float(v1)), m
        scale = float(g1.abs().max())
        assert float((g[m] - g1).abs().max()) <= 1e-12 * scale, m
