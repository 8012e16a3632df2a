"""The forward path's model-map epilogue and bootstrap chi^2 kernels
(smx_model_epilogue, smx_bootstrap_loglike) against the same formulas as torch
operations (the gradient's autograd path), on a non-trivial observed galaxy.

* surface density (map * mmpo) * lm and the clamped h1..h4: identical (the
  same IEEE operations); V and sigma (ghmoments.py:53-79) to 1e-15.
* logL and its K rows (likelihood.py:184-244): to 1e-13, the sums being
  reduced in another order; the rows' density term is the broadcast total.
* NaN entries are skipped as nansum skips them; maps that clamp at +-10 and a
  gamma_kin entry under the 1e-8 floor take the reference's branches.
The whole evaluation against the FP64 oracle: tests/test_gpu_data.py and the
other logL parity tests.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _torch_epilogue(maps, P, B, g_kin, v0, s, mmpo, lm):
    import torch
    from schwarmax_amd.ghmoments import h_to_V_sigma
    g = torch.where(g_kin.abs() > 1e-8, g_kin, torch.full_like(g_kin, 1e-8))
    d2 = maps[:, P:P + B] * mmpo * lm
    h = [torch.clamp(maps[:, P + B * (k + 1):P + B * (k + 2)] / g, -10.0, 10.0) for k in range(4)]
    V, sig = h_to_V_sigma(h[0], h[1], v0, s, h3=h[2], h4=h[3])
    return d2, h, V, sig


def test_epilogue_and_chi2_match_torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from schwarmax_amd import model as M, likelihood as L
    from galaxies import observed_galaxy
    theta, dd, sched = observed_galaxy(seed=11, n=40, n_ap=24, N_max=500)[:3]
    ph, pb = L.theta_to_params(theta)
    md, aux = M.build_model(L.density_func, L.potential_func, ph, pb, dd, 24, dd['Rzphi_n_tot'],
                            *(lambda a: (a[1], a[0], a[2], a[3]))(L.grid_args(dd)),
                            nnls_maxiter=50, schedule=sched, return_aux=True)
    lib = aux['libd']['lib']
    P, B = aux['libd']['P'], aux['libd']['B']
    W = md['weights']
    K = W.shape[0]
    rng = np.random.default_rng(2)
    g_kin = torch.as_tensor(rng.uniform(0.5, 2.0, B), device='cuda')
    g_kin[:5] = 1e-9                                  # under the floor: these clamp
    v0 = torch.as_tensor(rng.normal(0, 30, B), device='cuda')
    s = torch.as_tensor(rng.uniform(50, 200, B), device='cuda')
    mmpo, lm = 0.37, 1.7
    d3, d2, h1, h2, h3, h4, V, sig = M.compute_model_single(W, lib, P, B, g_kin, v0, s, mmpo, lm)
    from schwarmax_amd.nnls import model_maps
    maps = model_maps(lib, P + 5 * B, W)
    e2, eh, eV, esig = _torch_epilogue(maps, P, B, g_kin, v0, s, mmpo, lm)
    assert torch.equal(d3, maps[:, :P]) and torch.equal(d2, e2)
    for a, b in zip((h1, h2, h3, h4), eh):
        assert torch.equal(a, b)
    assert float((h1.abs() == 10.0).sum() + (h3.abs() == 10.0).sum()) > 0   # clamping exercised
    assert float((V - eV).abs().max()) <= 1e-15 * float(eV.abs().max())
    assert float((sig - esig).abs().max()) <= 1e-15 * float(esig.abs().max())
    # the chi^2 on the model dict of build_model, with one NaN entry per block
    mdn = dict(md)
    for k in ('density_3d_model', 'surface_density', 'h2'):
        t = mdn[k].clone()
        t[1, 2] = float('nan')
        mdn[k] = t
    for model_dict in (md, mdn):
        got, rows = L.bootstrap_loglike(model_dict, dd, 1.3)
        ref, rrows = L._bootstrap_loglike_graph(model_dict, dd, 1.3)
        assert np.isfinite(float(got))
        assert abs(float(got) - float(ref)) <= 1e-13 * abs(float(ref)), (float(got), float(ref))
        assert float((rows - rrows).abs().max()) <= 1e-13 * float(rrows.abs().max())
        assert rows.shape == (K,)
