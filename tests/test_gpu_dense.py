"""Dense building blocks on MFMA against FP64 references (GPU)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def T():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0)])
@pytest.mark.parametrize("shape", [(130, 70, 45), (256, 384, 512), (1, 300, 33)])
def test_gemm_f32(T, ta, tb, shape):
    from schwarmax_amd import _lib
    M, N, K = shape
    rng = np.random.default_rng(M + N + K)
    A = rng.normal(size=(K, M) if ta else (M, K)).astype(np.float32)
    B = rng.normal(size=(N, K) if tb else (K, N)).astype(np.float32)
    C0 = rng.normal(size=(M, N)).astype(np.float32)
    ref = 0.7 * (A.T if ta else A).astype(np.float64) @ (B.T if tb else B).astype(np.float64) - 0.5 * C0
    Ad, Bd, Cd = (T.tensor(x, device='cuda') for x in (A, B, C0))
    _lib.check(_lib.lib().smx_gemm_f32(ta, tb, M, N, K, 0.7, _lib.ptr(Ad), A.shape[1], _lib.ptr(Bd),
                                       B.shape[1], -0.5, _lib.ptr(Cd), N, _lib.stream_ptr()), 'gemm')
    got = Cd.cpu().numpy()
    assert np.abs(got - ref).max() < 1e-5 * np.sqrt(K) * np.abs(ref).max()


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0)])
@pytest.mark.parametrize("shape", [(130, 70, 45), (256, 384, 512), (1, 300, 33)])
def test_gemm_f64(T, ta, tb, shape):
    from schwarmax_amd import _lib
    M, N, K = shape
    rng = np.random.default_rng(M * N + K)
    A = rng.normal(size=(K, M) if ta else (M, K))
    B = rng.normal(size=(N, K) if tb else (K, N))
    C0 = rng.normal(size=(M, N))
    ref = 0.7 * (A.T if ta else A) @ (B.T if tb else B) - 0.5 * C0
    Ad, Bd, Cd = (T.tensor(x, device='cuda') for x in (A, B, C0))
    _lib.check(_lib.lib().smx_gemm_f64(ta, tb, M, N, K, 0.7, _lib.ptr(Ad), A.shape[1], _lib.ptr(Bd),
                                       B.shape[1], -0.5, _lib.ptr(Cd), N, _lib.stream_ptr()), 'dgemm')
    assert np.abs(Cd.cpu().numpy() - ref).max() < 1e-13 * np.sqrt(K) * np.abs(ref).max()


@pytest.mark.parametrize("n", [60, 64, 300, 777, 2100])
def test_spd_inverse(T, n):
    """FP64 inverse of an orbit-library-like Gram (+ rho I), ill-coupled blocks."""
    from schwarmax_amd import _lib
    rng = np.random.default_rng(n)
    X = rng.random((2 * n, n)) * (rng.random((2 * n, n)) < 0.2)
    X[:, : n // 3] *= 30.0                       # strongly different column scales
    Q = X.T @ X
    Q += (np.trace(Q) / n) * np.eye(n)          # the ADMM's Q + rho I
    Ad = T.tensor(Q, device='cuda')
    work = T.empty(n * n // 2 + n, dtype=T.float64, device='cuda')
    _lib.check(_lib.lib().smx_spd_inverse_f64(_lib.ptr(Ad), n, _lib.ptr(work), _lib.stream_ptr()), 'inv')
    got = Ad.cpu().numpy()
    ref = np.linalg.inv(Q)
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert err < 1e-13 * np.linalg.cond(Q), err
    assert np.abs(got - got.T).max() <= 1e-12 * np.abs(got).max()
