"""Jackknife pins (CPU): the oracle's per-replicate re-solve (model.py:290-603)
and the shared-operator algebra the GPU path uses in its place."""
import numpy as np
import pytest

NAP = 8


@pytest.fixture(scope="module")
def problem():
    """Library, targets and the unmasked design of a tiny galaxy."""
    from schwarmax_amd import synth, likelihood as L
    from oracle import model as OM, potential as OP, nnls as NN
    theta, dd, sched = synth.make_galaxy('plain-2k', seed=4, n=16, n_ap=NAP, grid=(5, 4, 4), K=2,
                                         nxy=10)
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
    A_xy = lib['surface_density'].T
    A_h = [lib['h%d' % k].T for k in (1, 2, 3, 4)]
    y_R = OM.density_target(dd, pb)
    synth.observe(dd, A_xy, A_h, y_R.sum() / A_xy.shape[1], pb['mass_to_light_ratio'], floor=1e-3)
    n = A_xy.shape[1]
    t = OM._targets(dd, pb, n)
    return dict(A_R=lib['Rzphi'].T, A_xy=A_xy, A_h=A_h, t=t, n=n)


def _solve(pr, sxy, sh, maxiter=60):
    from oracle import nnls as NN
    t = pr['t']
    return NN.solve_nnls_admm(pr['A_R'], pr['A_xy'], pr['A_h'], t['y_Rzphi'], t['y_xy'],
                              t['gamma_kin'], t['y_h'], t['sig_Rzphi'], sxy, sh, lambda_reg=1.0,
                              maxiter=maxiter, w_rzphi=np.sqrt(5.0), w_xy=np.sqrt(5.0), w_h=1.0)


def test_masking_is_row_deletion(problem):
    """sigma = 1e20 (model.py:486-493) removes the aperture's rows: the masked
    solve equals the solve of the design matrix without those rows."""
    from oracle import nnls as NN
    pr, t = problem, problem['t']
    for a in (0, 3, NAP - 1):
        drop = np.arange(NAP) == a
        sxy = np.where(drop, 1e20, t['sig_xy'])
        sh = [np.where(drop, 1e20, s) for s in t['sig_h']]
        w_mask = _solve(pr, sxy, sh)
        keep = ~drop
        U = NN.design_matrix(pr['A_R'], pr['A_xy'][keep], [h[keep] for h in pr['A_h']],
                             t['sig_Rzphi'], t['sig_xy'][keep], [s[keep] for s in t['sig_h']],
                             t['gamma_kin'][keep], np.sqrt(5.0), np.sqrt(5.0), 1.0)
        Y = NN.rhs_vectors(t['y_Rzphi'], t['sig_Rzphi'], t['y_xy'][keep][None],
                           [y[keep][None] for y in t['y_h']], t['sig_xy'][keep],
                           [s[keep] for s in t['sig_h']], np.sqrt(5.0), np.sqrt(5.0), 1.0)
        _, rho, Lc = NN.gram_factor(U, 1.0)
        w_del = NN.admm(Lc, rho, -(U.T @ Y[0]), np.sum(t['y_Rzphi']) / pr['n'], 60)
        assert np.max(np.abs(w_mask - w_del)) <= 1e-10 * np.max(np.abs(w_del)), a


def test_shared_operator_algebra(problem):
    """One inverse M = (Q + rho I)^-1 of the unmasked Gram, a Neumann series
    for the shift rho - rho_g and Woodbury for the dropped rows reproduce each
    replicate's own Cholesky iterates."""
    from oracle import nnls as NN
    pr, t = problem, problem['t']
    n = pr['n']
    U = NN.design_matrix(pr['A_R'], pr['A_xy'], pr['A_h'], t['sig_Rzphi'], t['sig_xy'],
                         t['sig_h'], t['gamma_kin'], np.sqrt(5.0), np.sqrt(5.0), 1.0)
    Y = NN.rhs_vectors(t['y_Rzphi'], t['sig_Rzphi'], t['y_xy'][None], [y[None] for y in t['y_h']],
                       t['sig_xy'], t['sig_h'], np.sqrt(5.0), np.sqrt(5.0), 1.0)[0]
    Q, rho, _ = NN.gram_factor(U, 1.0)
    M = np.linalg.inv(Q + rho * np.eye(n))
    c = -(U.T @ Y)
    P = pr['A_R'].shape[0]
    for a in range(NAP):
        rows = [P + k * NAP + a for k in range(5)]
        V = U[rows]
        delta = np.sum(V * V) / n
        c_g = c + V.T @ Y[rows]
        rho_g = rho - delta
        q = delta * np.linalg.norm(M, 2)
        T = max(1, int(np.ceil(np.log(1e-15) / np.log(q))))

        def binv(r):
            s = M @ r
            acc = s.copy()
            for k in range(1, T):
                s = M @ s
                acc = acc + delta ** k * s
            return acc
        Pg = binv(V.T)                                  # (n, 5)
        Cinv = np.linalg.inv(np.eye(5) - V @ Pg)
        z = np.full(n, np.sum(t['y_Rzphi']) / n)
        u = np.zeros(n)
        for _ in range(60):
            x = binv(rho_g * (z - u) - c_g)
            w = x + Pg @ (Cinv @ (V @ x))
            w_hat = 1.6 * w - 0.6 * z
            z_new = np.maximum(0.0, w_hat + u)
            u = u + w_hat - z_new
            z = z_new
        drop = np.arange(NAP) == a
        ref = _solve(pr, np.where(drop, 1e20, t['sig_xy']),
                     [np.where(drop, 1e20, s) for s in t['sig_h']])
        assert np.max(np.abs(z - ref)) <= 1e-9 * np.max(np.abs(ref)), (a, T, q)


def test_jackknife_statistic(problem):
    """The oracle's delta-chi^2 is the population std of the replicates'
    chi2_func values, and chi2_func ignores the masked (sigma = 1e20) bins."""
    from oracle import model as OM
    y = np.ones(4)
    assert OM.chi2_func(y, y, y, y, y, y, y, y + 1, np.array([1.0, 1.0, 1e20, 1e20]),
                        y, y, y, y, y, y, y, y, y) == -0.5 * 2.0
