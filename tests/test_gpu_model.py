"""End-to-end build_model / model_likelihood / get_model_with_orbit parity (GPU)
against the FP64 oracle, on a galaxy whose library passes ``library_guard``
(>= 80% valid integrations, non-zero density, surface density and h1..h4).

Bars: aperture moments (h1..h4, V, sigma), surface density, density model and
NNLS weights within 1e-4 relative (row-wise 2-norm over apertures / orbits);
logL within 1e-4; trajectories within 1e-8 relative per orbit, and any orbit
beyond that must follow an accept decision that flips at err_norm = 1 to
within 1e-6 (integrates.py:101), reported with its first-divergence step.
"""
import numpy as np
import pytest

from galaxies import observed_galaxy, rowwise_rel

pytestmark = pytest.mark.gpu
NAP = 24
MOMENT_KEYS = ('density_3d_model', 'surface_density', 'h1', 'h2', 'h3', 'h4', 'V', 'sigma')


@pytest.fixture(scope="module")
def galaxy():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from schwarmax_amd import likelihood as L
    theta, dd, sched, _, frac = observed_galaxy(seed=11, n=40, n_ap=NAP, K=5, N_max=500)
    ph, pb = L.theta_to_params(theta)
    return theta, dd, sched, ph, pb


def test_build_model_outputs(galaxy):
    from schwarmax_amd import model as M, likelihood as L
    from oracle import model as OM, potential as OP
    theta, dd, sched, ph, pb = galaxy
    Rl, Rn, xyl, xyn = L.grid_args(dd)
    ref = OM.build_model(OP.density_func, OP.potential_func, ph, pb, dd, NAP, dd['Rzphi_n_tot'], Rn,
                         Rl, xyl, xyn, nnls_maxiter=200, schedule=sched)
    got, aux = M.build_model(L.density_func, L.potential_func, ph, pb, dd, NAP, dd['Rzphi_n_tot'],
                             Rn, Rl, xyl, xyn, nnls_maxiter=200, schedule=sched, return_aux=True)
    assert set(got) == set(ref)
    assert float(aux['libd']['valid'].mean()) >= 0.8
    d3 = got['density_3d'].cpu().numpy()
    assert np.max(np.abs(d3 - ref['density_3d']) / np.abs(ref['density_3d']).max()) < 1e-12
    errs = {}
    for k in MOMENT_KEYS + ('weights',):
        g = got[k].cpu().numpy()
        assert g.shape == ref[k].shape, k
        errs[k] = rowwise_rel(g, ref[k])
    print("build_model relative errors:", {k: f"{v:.2e}" for k, v in errs.items()})
    for k, e in errs.items():
        assert e < 1e-4, (k, e)
    # the maps are not trivial: the kinematic moments vary across apertures
    for k in ('h1', 'h2', 'h3', 'h4'):
        assert np.std(ref[k]) > 1e-3 * np.max(np.abs(ref[k])) > 0.0, k


def test_model_likelihood(galaxy):
    from schwarmax_amd import likelihood as L
    from oracle import model as OM
    theta, dd, sched, ph, pb = galaxy
    ref = OM.model_likelihood(theta, dd, NAP, dd['Rzphi_n_tot'], schedule=sched, nnls_maxiter=200)
    got = float(L.model_likelihood(theta, dd, NAP, dd['Rzphi_n_tot'], schedule=sched,
                                   nnls_maxiter=200).item())
    print(f"logL gpu={got:.12g} oracle={ref:.12g} rel={abs(got - ref) / abs(ref):.2e}")
    assert np.isfinite(ref)
    assert abs(got - ref) / abs(ref) < 1e-4, (got, ref)


def divergence_report(y_got, t_got, y_ref, rec, tol=1e-8):
    """Per orbit beyond ``tol``: (orbit, first step past tol, first step whose
    accept decision differs, the oracle's err_norm at that step)."""
    n, N = y_ref.shape[0], y_ref.shape[1]
    scale = np.abs(y_ref).reshape(n, -1).max(axis=1)
    dev = np.abs(y_got - y_ref).max(axis=2) / scale[:, None]          # (n, N)
    acc_got = np.diff(np.concatenate([np.zeros((n, 1)), t_got], axis=1), axis=1) > 0.0
    rows = []
    for i in np.nonzero(dev.max(axis=1) > tol)[0]:
        first_div = int(np.argmax(dev[i] > tol))
        flips = np.nonzero(acc_got[i] != rec['accept'][i])[0]
        first_flip = int(flips[0]) if flips.size else -1
        err = float(rec['err'][i, first_flip]) if flips.size else float('nan')
        rows.append((int(i), first_div, first_flip, err))
    return rows


def test_get_model_with_orbit(galaxy):
    from schwarmax_amd import model as M, likelihood as L
    from oracle import model as OM, potential as OP, integrate as OI
    theta, dd, sched, ph, pb = galaxy
    Rl, Rn, xyl, xyn = L.grid_args(dd)
    ref = OM.get_model_with_orbit(OP.density_func, OP.potential_func, ph, pb, dd, NAP,
                                  dd['Rzphi_n_tot'], Rn, Rl, xyl, xyn, nnls_maxiter=100,
                                  schedule=sched)
    got = M.get_model_with_orbit(L.density_func, L.potential_func, ph, pb, dd, NAP, dd['Rzphi_n_tot'],
                                 Rn, Rl, xyl, xyn, nnls_maxiter=100, schedule=sched)
    assert set(got) == set(ref)
    yt = got['y_traj'].cpu().numpy()
    assert yt.shape == ref['y_traj'].shape
    # the oracle's per-attempt decisions for the same initial conditions
    s = OM.schedule_of(sched)
    pbb = OM.baryon_params(pb)
    wb, Tt, dt0 = OM.initial_conditions(dd, pbb, ph, 1, s['N_step_per_orb'], s['N_dynamical_time'])
    acc_fn, pot_fn = OI.make_acc_pot(pbb, ph)
    rot = OM.make_rotation_matrix(pb['alpha'], pb['beta'], pb['gamma'])
    grids = OI.Grids(NAP, dd['bin_mapping'], dd['num_per_bin'], Rl, xyl, Rn, xyn, dd['v0'], dd['s'],
                     rot)
    rec = {}
    per = OI.integrate_withtraj(wb.reshape(-1, 6), acc_fn, pot_fn, s['N_max'], Tt.reshape(-1),
                                dt0.reshape(-1), -pb['Omega_bar'], s['atol'], s['rtol'], s['dt_min'],
                                s['dt_max'], grids, record=rec)
    assert np.array_equal(per['o_traj'], ref['y_traj'][:, 0])
    assert per['valid'].mean() >= 0.8, per['valid'].mean()
    n = yt.shape[0]
    y_got, y_ref = yt[:, 0], ref['y_traj'][:, 0]
    t_got = got['t_traj'].cpu().numpy()[:, 0]
    rows = divergence_report(y_got, t_got, y_ref, rec)
    print(f"trajectories: {n - len(rows)}/{n} orbits within 1e-8; beyond it "
          "(orbit, first step past 1e-8, first flipped decision, oracle err_norm there):", rows)
    assert len(rows) <= n // 10, rows
    for i, first_div, first_flip, err in rows:
        assert 0 <= first_flip <= first_div, (i, first_div, first_flip)
        assert abs(err - 1.0) < 1e-6, (i, err)
    ok = np.setdiff1d(np.arange(n), [r[0] for r in rows])
    tt_err = np.abs(t_got[ok] - ref['t_traj'][ok, 0]) / ref['t_traj'][ok, 0, -1:]
    assert tt_err.max() < 1e-8, tt_err.max()
    # the library of the trajectory path (no symmetry, dt / T, unit channel
    # weights) and everything built on it, when no orbit diverged
    if not rows:
        errs = {k: rowwise_rel(got[k].cpu().numpy(), ref[k]) for k in MOMENT_KEYS}
        errs['weights'] = rowwise_rel(got['weights_all_boot'].cpu().numpy(),
                                      ref['weights_all_boot'])
        print("get_model_with_orbit relative errors:", {k: f"{v:.2e}" for k, v in errs.items()})
        for k, e in errs.items():
            assert e < 1e-4, (k, e)
