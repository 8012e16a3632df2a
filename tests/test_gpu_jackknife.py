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
