"""Edge cases of the NNLS drop-ins on the GPU against the FP64 oracle
(nnls.py:6-166): the single-right-hand-side solver, zero iterations, orbit
counts that are not multiples of the kernels' vector width or tile, more
right-hand sides than one row block, and a right-hand side with no data."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _fit_err(U, Y, got, ref):
    f_ref = np.sum((ref @ U.T - Y) ** 2, axis=-1)
    f_got = np.sum((got @ U.T - Y) ** 2, axis=-1)
    return np.max(np.abs(f_got - f_ref) / f_ref)


def test_single_rhs_solver(cuda):
    """solve_nnls_admm (nnls.py:6-66), one right-hand side: weights (n,)."""
    from schwarmax_amd import nnls as GN
    from oracle import nnls as ON
    from test_gpu_nnls import _problem
    A_R, A_xy, A_h, y_R, y_xy, gamma, yb_xy, yb_h, sR, sxy, sh = _problem(3, n=257, K=1)
    y_h = [b[0] for b in yb_h]
    w5 = np.sqrt(5.0)
    ref = ON.solve_nnls_admm(A_R, A_xy, A_h, y_R, y_xy, gamma, y_h, sR, sxy, sh, lambda_reg=1.0,
                             maxiter=300, w_rzphi=w5, w_xy=w5, w_h=1.0)
    got = GN.solve_nnls_admm(A_R, A_xy, *A_h, y_R, y_xy, gamma, *y_h, sR, sxy, *sh,
                             lambda_reg=1.0, maxiter=300, w_rzphi=w5, w_xy=w5,
                             w_h=1.0).cpu().numpy()
    assert got.shape == ref.shape == (257,)
    e = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print("weights: normwise relative error %.2e" % e)
    assert e < 1e-4, e
    U = ON.design_matrix(A_R, A_xy, A_h, sR, sxy, sh, gamma, w5, w5, 1.0)
    Y = ON.rhs_vectors(y_R, sR, y_xy[None], [y[None] for y in y_h], sxy, sh, w5, w5, 1.0)
    assert _fit_err(U, Y, got[None], ref[None]) < 1e-4


def test_zero_iterations_return_the_start(cuda):
    """maxiter = 0: z keeps its initial value sum(y_Rzphi) / n (nnls.py:122)."""
    from schwarmax_amd import nnls as GN
    from test_gpu_nnls import _problem
    A_R, A_xy, A_h, y_R, y_xy, gamma, yb_xy, yb_h, sR, sxy, sh = _problem(4, n=130, K=3)
    got = GN.solve_nnls_admm_bootstrap(A_R, A_xy, *A_h, y_R, y_xy, gamma, yb_xy, *yb_h, sR, sxy,
                                       *sh, lambda_reg=1.0, maxiter=0).cpu().numpy()
    assert np.all(got == got[0, 0])
    assert abs(got[0, 0] - np.sum(y_R) / 130) <= 1e-14 * abs(np.sum(y_R) / 130)


@pytest.mark.parametrize("n,K", [(131, 5), (301, 150), (257, 256), (200, 300)])
def test_ragged_shapes(cuda, n, K):
    """n not a multiple of 4 (the update kernel's scalar path) or of 256 (the
    product's column blocks); K above 128: two row blocks of R (the second
    partly and wholly filled) and three."""
    from schwarmax_amd import nnls as GN
    from oracle import nnls as ON
    from test_gpu_nnls import _problem
    A_R, A_xy, A_h, y_R, y_xy, gamma, yb_xy, yb_h, sR, sxy, sh = _problem(5, n=n, K=K)
    w5 = np.sqrt(5.0)
    ref = ON.solve_nnls_admm_bootstrap(A_R, A_xy, A_h, y_R, y_xy, gamma, yb_xy, yb_h, sR, sxy, sh,
                                       lambda_reg=1.0, maxiter=250, w_rzphi=w5, w_xy=w5, w_h=1.0)
    got = GN.solve_nnls_admm_bootstrap(A_R, A_xy, *A_h, y_R, y_xy, gamma, yb_xy, *yb_h, sR, sxy,
                                       *sh, lambda_reg=1.0, maxiter=250, w_rzphi=w5, w_xy=w5,
                                       w_h=1.0).cpu().numpy()
    assert got.shape == (K, n)
    err = np.linalg.norm(got - ref, axis=1) / np.linalg.norm(ref, axis=1)
    print("weights: normwise relative error per right-hand side, max %.2e" % err.max())
    assert err.max() < 1e-4, err
    U = ON.design_matrix(A_R, A_xy, A_h, sR, sxy, sh, gamma, w5, w5, 1.0)
    Y = ON.rhs_vectors(y_R, sR, yb_xy, yb_h, sxy, sh, w5, w5, 1.0)
    assert _fit_err(U, Y, got, ref) < 1e-4


def test_rhs_without_kinematic_data(cuda):
    """A bootstrap row whose kinematic data are all zero: its R row carries
    only the density block, the other rows are unaffected."""
    from schwarmax_amd import nnls as GN
    from oracle import nnls as ON
    from test_gpu_nnls import _problem
    A_R, A_xy, A_h, y_R, y_xy, gamma, yb_xy, yb_h, sR, sxy, sh = _problem(6, n=200, K=4)
    yb_xy = yb_xy.copy()
    yb_xy[2] = 0.0
    yb_h = [b.copy() for b in yb_h]
    for b in yb_h:
        b[2] = 0.0
    w5 = np.sqrt(5.0)
    ref = ON.solve_nnls_admm_bootstrap(A_R, A_xy, A_h, y_R, y_xy, gamma, yb_xy, yb_h, sR, sxy, sh,
                                       lambda_reg=1.0, maxiter=200, w_rzphi=w5, w_xy=w5, w_h=1.0)
    got = GN.solve_nnls_admm_bootstrap(A_R, A_xy, *A_h, y_R, y_xy, gamma, yb_xy, *yb_h, sR, sxy,
                                       *sh, lambda_reg=1.0, maxiter=200, w_rzphi=w5, w_xy=w5,
                                       w_h=1.0).cpu().numpy()
    assert np.all(np.isfinite(got)) and np.all(got >= 0.0)
    err = np.linalg.norm(got - ref, axis=1) / np.linalg.norm(ref, axis=1)
    print("weights: normwise relative error per right-hand side, max %.2e" % err.max())
    assert err.max() < 1e-4, err
