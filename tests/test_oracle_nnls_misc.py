"""Pins of the oracle's NNLS, parameter maps, rotation and Jeans moments (CPU)."""
import numpy as np
import pytest
import scipy.linalg as sla
from scipy import integrate, optimize

from oracle import nnls as N, model as OM, jeans as J, potential as P


def _small(seed=0, m=60, n=25):
    rng = np.random.default_rng(seed)
    U = rng.random((m, n)) * (rng.random((m, n)) < 0.4)
    y = U @ rng.dirichlet(np.ones(n)) * n + 0.05 * rng.normal(size=m)
    return U, y


def test_admm_converges_to_scipy_nnls():
    """The ADMM fixed point is the unique minimiser of ||Uw - y||^2 + (lam/n)||w||^2, w >= 0."""
    U, y = _small()
    n = U.shape[1]
    lam = 1.0
    _, rho, L = N.gram_factor(U, lam)
    w = N.admm(L, rho, -(U.T @ y), 1.0, 20000)
    Ust = np.vstack([U, np.sqrt(lam / n) * np.eye(n)])
    ref, _ = optimize.nnls(Ust, np.concatenate([y, np.zeros(n)]), maxiter=5000)
    np.testing.assert_allclose(w, ref, atol=1e-7)


def test_admm_first_iterate_closed_form():
    U, y = _small(1)
    n = U.shape[1]
    Q = U.T @ U + (1.0 / n) * np.eye(n)
    rho = np.trace(Q) / n
    c = -(U.T @ y)
    w1 = np.linalg.solve(Q + rho * np.eye(n), rho * 0.7 - c)
    z1 = np.maximum(0.0, 1.6 * w1 - 0.6 * 0.7)
    _, rho_o, L = N.gram_factor(U, 1.0)
    assert abs(rho_o - rho) < 1e-14 * rho
    np.testing.assert_allclose(N.admm(L, rho_o, c, 0.7, 1), z1, rtol=1e-12, atol=1e-14)


def test_bootstrap_matches_single():
    """Row k of the bootstrap solve equals the single-RHS solve of that data."""
    rng = np.random.default_rng(3)
    P_, B, n, K = 20, 8, 30, 3
    A = rng.random((P_ + 5 * B, n)) * (rng.random((P_ + 5 * B, n)) < 0.3)
    yR = A[:P_] @ np.ones(n)
    sR, sxy = 0.05 * yR + 0.01, np.full(B, 0.1)
    sh = [np.full(B, 0.03)] * 4
    g = rng.random(B) + 0.5
    yxy_b = rng.random((K, B))
    yh_b = [rng.normal(size=(K, B)) * 0.1 for _ in range(4)]
    Ah = [A[P_ + B * (k + 1):P_ + B * (k + 2)] for k in range(4)]
    W = N.solve_nnls_admm_bootstrap(A[:P_], A[P_:P_ + B], Ah, yR, yxy_b[0], g, yxy_b, yh_b, sR, sxy,
                                    sh, maxiter=50)
    for k in range(K):
        w = N.solve_nnls_admm(A[:P_], A[P_:P_ + B], Ah, yR, yxy_b[k], g, [a[k] for a in yh_b], sR,
                              sxy, sh, maxiter=50)
        np.testing.assert_allclose(W[k], w, rtol=1e-12, atol=1e-14)


def test_theta_map_round_trip():
    for logM, logRs in [(12.0, 1.3), (11.5, 0.9), (12.4, 1.5)]:
        lme, lc = OM.logM_logRs_to_logMenc_logc(logM, logRs)
        a, b = OM.logMenc_logc_to_logM_logRs(lme, lc)
        assert abs(a - logM) < 1e-10 and abs(b - logRs) < 1e-10


def test_rotation_matrix():
    R = OM.make_rotation_matrix(20.0, 65.0, 45.0)
    np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-14)
    assert abs(np.linalg.det(R) - 1) < 1e-14
    R0 = OM.make_rotation_matrix(30.0, 0.0, 15.0)
    t = np.radians(45.0)
    Rz = np.array([[np.cos(t), -np.sin(t), 0], [np.sin(t), np.cos(t), 0], [0, 0, 1]])
    np.testing.assert_allclose(R0, Rz.T, atol=1e-14)


def test_jeans_sigma_z_against_quad(canon_params):
    _, ph, pb = canon_params
    bp = OM.baryon_params(pb)
    xs, ys, zs = np.array([3.0, 6.0]), np.array([1.0, -2.0]), np.array([0.2, -0.5])
    vm, sR, sz, sphi = J.jeans_moments(xs, ys, zs, bp, ph, P.potential_func, P.density_func)
    for i in range(2):
        d = 5e-3

        def f(zp):
            dpz = (P.potential_func(xs[i], ys[i], zp + d, bp, ph) - P.potential_func(xs[i], ys[i], zp - d, bp, ph)) / (2 * d)
            return P.density_func(xs[i], ys[i], zp, bp) * dpz
        I, _ = integrate.quad(f, abs(zs[i]), 10.0, limit=400, epsrel=1e-10)
        nu = P.density_func(xs[i], ys[i], zs[i], bp)
        # 1000-point trapezoid vs adaptive quadrature: discretisation-level agreement
        assert abs(sz[i] ** 2 - I / nu) < 1e-4 * abs(I / nu)
        assert sR[i] == sz[i] and sphi[i] == sz[i]


def test_fp64_torch_admm_equals_oracle():
    """The torch FP64 restatement used at full shape on the GPU reproduces the
    oracle's iterate sequence (nnls.py:101-163) on a small problem."""
    import torch
    from oracle import nnls as ON
    from fp64_admm import admm_fp64
    rng = np.random.default_rng(7)
    m, n, K = 90, 60, 5
    U = rng.random((m, n)) * (rng.random((m, n)) < 0.3)
    Y = rng.random((K, m)) * 3.0
    Q, rho, L = ON.gram_factor(U, 1.0)
    ref = ON.admm(L, rho, -(Y @ U), 0.7, 300)
    got = admm_fp64(torch.tensor(U), torch.tensor(Y), 1.0, 300, 0.7).numpy()
    assert got.shape == ref.shape == (K, n)
    assert np.max(np.abs(got - ref)) <= 1e-10 * np.max(np.abs(ref))
    assert np.any(ref == 0.0) and np.any(ref > 0.0)            # the projection is active
