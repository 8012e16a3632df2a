"""Jackknife delta-chi^2 (GPU): the shared-operator solve of all replicates
against the oracle's per-replicate re-solve (model.py:290-603).  Tolerance: the
replicates' chi^2 to 1e-4 of the largest, delta-chi^2 to 1e-3 (the ADMM
products are FP32-accurate, the oracle's Cholesky solves FP64)."""
import numpy as np
import pytest

from test_oracle_jackknife import problem  # noqa: F401  (fixture)

pytestmark = pytest.mark.gpu
NAP = 8


@pytest.fixture(scope="module")
def galaxy():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from galaxies import observed_galaxy
    theta, dd, sched, _, _ = observed_galaxy(seed=4, n=16, n_ap=NAP, grid=(5, 4, 4), K=2, nxy=10,
                                             N_max=200)
    return theta, dd, sched


@pytest.mark.parametrize("n_groups,iters", [(NAP, 60), (3, 400)])
def test_jackknife_matches_oracle(galaxy, n_groups, iters):
    from schwarmax_amd import jackknife as J, likelihood as L
    from oracle import model as OM, potential as OP
    theta, dd, sched = galaxy
    ph, pb = L.theta_to_params(theta)
    Rl, Rn, xyl, xyn = L.grid_args(dd)
    got, aux = J.model_deltaChi2_jackknife(L.density_func, L.potential_func, L.chi2_func, ph, pb, dd,
                                           NAP, dd['Rzphi_n_tot'], Rn, Rl, xyl, xyn,
                                           nnls_maxiter=iters, n_groups=n_groups, schedule=sched,
                                           return_aux=True)
    ref, ref_all = OM.model_deltaChi2_jackknife(OP.density_func, OP.potential_func, OM.chi2_func,
                                                ph, pb, dd, NAP, dd['Rzphi_n_tot'], Rn, Rl, xyl,
                                                xyn, nnls_maxiter=iters, n_groups=n_groups,
                                                schedule=sched, return_all=True)
    chi2 = aux['chi2'].cpu().numpy()
    # the Neumann terms' products mix full and delta products; the last
    # iteration's are full
    print(f"operator products {aux['products']}, full {aux['full_products']}")
    assert aux['products'] == iters * aux['neumann_terms']
    assert aux['neumann_terms'] <= aux['full_products'] <= aux['products']
    if iters >= 400:
        assert aux['full_products'] < aux['products'], "no delta products were exercised"
    assert chi2.shape == ref_all.shape
    assert np.max(np.abs(chi2 - ref_all)) <= 1e-4 * np.max(np.abs(ref_all)), (chi2, ref_all)
    assert abs(float(got) - ref) <= 1e-3 * abs(ref), (float(got), ref)


@pytest.mark.parametrize("G,n,d,T", [(7, 1000, 5, 3), (3, 777, 15, 1), (2, 64, 1, 5), (4, 300, 17, 2)])
def test_jackknife_admm_step_matches_torch(G, n, d, T):
    """smx_jackknife_admm_step against its formulas written out in FP64 torch:
    x = sum_k delta^k S_k, x + P^T Cinv (V x), then nnls.py:155-162 and the
    next right-hand sides (FP32), on random inputs with d across several
    passes of the dot products."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from schwarmax_amd import _lib
    gen = torch.Generator(device='cuda')
    gen.manual_seed(G * n + d + T)
    rnd = lambda *s: torch.randn(*s, dtype=torch.float64, device='cuda', generator=gen)  # noqa: E731
    S = rnd(T, G, n).to(torch.float32)
    delta = rnd(G).abs() * 0.1
    V, P, Cinv = rnd(G, d, n), rnd(G, d, n), rnd(G, d, d)
    rho_g, c = rnd(G).abs() + 1.0, rnd(G, n)
    z, u = rnd(G, n).abs(), rnd(G, n) * 0.1
    alpha = 1.6
    x = S[0].double().clone()
    for k in range(1, T):
        x += (delta ** k)[:, None] * S[k].double()
    s = torch.bmm(V, x.unsqueeze(-1)).squeeze(-1)
    w = x + torch.bmm(P.transpose(1, 2).contiguous(), torch.bmm(Cinv, s.unsqueeze(-1))).squeeze(-1)
    w_hat = alpha * w + (1.0 - alpha) * z
    z_ref = torch.clamp(w_hat + u, min=0.0)
    u_ref = u + w_hat - z_ref
    r_ref = rho_g[:, None] * (z_ref - u_ref) - c
    r = torch.full((G, n), float('nan'), dtype=torch.float32, device='cuda')
    _lib.check(_lib.lib().smx_jackknife_admm_step(G, n, d, T, _lib.ptr(S, torch.float32),
                                                  _lib.ptr(delta), _lib.ptr(V), _lib.ptr(P),
                                                  _lib.ptr(Cinv), _lib.ptr(rho_g), _lib.ptr(c), alpha,
                                                  _lib.ptr(z), _lib.ptr(u),
                                                  _lib.ptr(r, torch.float32), _lib.stream_ptr()),
               'smx_jackknife_admm_step')
    # the dot products over n and the rank-d sum round differently from bmm's
    tol = 1e-13 * (float(w.abs().max()) + float(c.abs().max())) * (d + 1) * float(rho_g.max())
    for got, ref in ((z, z_ref), (u, u_ref)):
        assert torch.isfinite(got).all()
        assert float((got - ref).abs().max()) <= tol, float((got - ref).abs().max())
    # r is stored in FP32, the operator products' input precision
    err = float((r.double() - r_ref).abs().max())
    assert torch.isfinite(r).all() and err <= tol + 2.0 ** -23 * float(r_ref.abs().max()), err
