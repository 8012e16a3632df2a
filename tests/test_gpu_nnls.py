"""ADMM NNLS on the GPU against the FP64 oracle (nnls.py:6-166)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problem(seed, n=300, P=120, B=40, K=6):
    rng = np.random.default_rng(seed)
    # sparse-ish nonnegative library like orbit occupancies
    A = rng.random((P + 5 * B, n)) * (rng.random((P + 5 * B, n)) < 0.15)
    A[P + B:] = rng.normal(scale=0.2, size=(4 * B, n)) * (A[P + B:] != 0)
    w_true = rng.dirichlet(np.ones(n)) * n
    y = A @ w_true
    sig = np.abs(y) * 0.05 + 0.01
    gamma = np.abs(A[P:P + B] @ w_true) + 0.5
    y_R, y_xy = y[:P], y[P:P + B]
    yb_xy = y_xy[None] + rng.normal(size=(K, B)) * sig[P:P + B]
    yb_xy[0] = y_xy
    yh = [y[P + B * (k + 1):P + B * (k + 2)] / gamma for k in range(4)]
    yb_h = [yh[k][None] + 0.03 * rng.normal(size=(K, B)) for k in range(4)]
    sh = [np.full(B, 0.03) for _ in range(4)]
    A_h = [A[P + B * (k + 1):P + B * (k + 2)] for k in range(4)]
    return A[:P], A[P:P + B], A_h, y_R, y_xy, gamma, yb_xy, yb_h, sig[:P], sig[P:P + B], sh


@pytest.mark.parametrize("maxiter", [50, 400])
def test_bootstrap_weights(maxiter):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from schwarmax_amd import nnls as GN
    from oracle import nnls as ON
    A_R, A_xy, A_h, y_R, y_xy, gamma, yb_xy, yb_h, sR, sxy, sh = _problem(0)
    w5 = np.sqrt(5.0)
    ref = ON.solve_nnls_admm_bootstrap(A_R, A_xy, A_h, y_R, y_xy, gamma, yb_xy, yb_h, sR, sxy, sh,
                                       lambda_reg=1.0, maxiter=maxiter, w_rzphi=w5, w_xy=w5, w_h=1.0)
    got = GN.solve_nnls_admm_bootstrap(A_R, A_xy, *A_h, y_R, y_xy, gamma, yb_xy, *yb_h, sR, sxy, *sh,
                                       lambda_reg=1.0, maxiter=maxiter, w_rzphi=w5, w_xy=w5,
                                       w_h=1.0).cpu().numpy()
    # FP32 factor / products against the FP64 oracle: normwise per right-hand side
    err = np.linalg.norm(got - ref, axis=1) / np.linalg.norm(ref, axis=1)
    print("weights: normwise relative error per right-hand side, max %.2e" % err.max())
    assert err.max() < 1e-4, err
    # and the fit quality (the chi^2 the likelihood sees) to 1e-4
    U = ON.design_matrix(A_R, A_xy, A_h, sR, sxy, sh, gamma, w5, w5, 1.0)
    Y = ON.rhs_vectors(y_R, sR, yb_xy, yb_h, sxy, sh, w5, w5, 1.0)
    f_ref = np.sum((ref @ U.T - Y) ** 2, axis=1)
    f_got = np.sum((got @ U.T - Y) ** 2, axis=1)
    assert np.max(np.abs(f_got - f_ref) / f_ref) < 1e-4


def test_delta_products_exercised():
    """The solve mixes full products with delta products (the inverse's hi
    plane times R - R_base) and still meets the oracle's bars; the first
    and the last iteration are full."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from schwarmax_amd import nnls as GN, _lib
    from oracle import nnls as ON
    A_R, A_xy, A_h, y_R, y_xy, gamma, yb_xy, yb_h, sR, sxy, sh = _problem(2, n=300, K=12)
    w5 = np.sqrt(5.0)
    maxiter = 600
    before = dict(GN.STATS)
    _lib.timing_enable(True)
    try:
        got = GN.solve_nnls_admm_bootstrap(A_R, A_xy, *A_h, y_R, y_xy, gamma, yb_xy, *yb_h, sR, sxy,
                                           *sh, lambda_reg=1.0, maxiter=maxiter, w_rzphi=w5,
                                           w_xy=w5, w_h=1.0).cpu().numpy()
    finally:
        _lib.timing_enable(False)
    products = GN.STATS['products'] - before['products']
    full = GN.STATS['full_products'] - before['full_products']
    print(f"ADMM products {products}, full {full}")
    assert products == maxiter
    assert 2 <= full < maxiter // 2, full
    ref = ON.solve_nnls_admm_bootstrap(A_R, A_xy, A_h, y_R, y_xy, gamma, yb_xy, yb_h, sR, sxy, sh,
                                       lambda_reg=1.0, maxiter=maxiter, w_rzphi=w5, w_xy=w5, w_h=1.0)
    err = np.linalg.norm(got - ref, axis=1) / np.linalg.norm(ref, axis=1)
    print("weights: normwise relative error per right-hand side, max %.2e" % err.max())
    assert err.max() < 1e-4, err
    U = ON.design_matrix(A_R, A_xy, A_h, sR, sxy, sh, gamma, w5, w5, 1.0)
    Y = ON.rhs_vectors(y_R, sR, yb_xy, yb_h, sxy, sh, w5, w5, 1.0)
    f_ref = np.sum((ref @ U.T - Y) ** 2, axis=1)
    f_got = np.sum((got @ U.T - Y) ** 2, axis=1)
    assert np.max(np.abs(f_got - f_ref) / f_ref) < 1e-4
