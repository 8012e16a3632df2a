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


@pytest.mark.parametrize("dtype", ["f32", "f64"])
@pytest.mark.parametrize("shape", [(130, 70, 45), (256, 384, 512), (1, 300, 33), (300, 260, 96),
                                   (129, 257, 31)])
def test_gemm_nt(T, dtype, shape):
    """C = 0.7 A B^T - 0.5 C against an FP64 product (vector and scalar loaders)."""
    from schwarmax_amd import _lib
    M, N, K = shape
    rng = np.random.default_rng(M * N + K)
    A = rng.normal(size=(M, K))
    B = rng.normal(size=(N, K))
    C0 = rng.normal(size=(M, N))
    ref = 0.7 * A @ B.T - 0.5 * C0
    td = T.float32 if dtype == "f32" else T.float64
    Ad, Bd, Cd = (T.tensor(x, device='cuda', dtype=td) for x in (A, B, C0))
    fn = _lib.lib().smx_gemm_nt_f32 if dtype == "f32" else _lib.lib().smx_gemm_nt_f64
    _lib.check(fn(M, N, K, 0.7, _lib.ptr(Ad), K, _lib.ptr(Bd), K, -0.5, _lib.ptr(Cd), N,
                  _lib.stream_ptr()), 'gemm_nt')
    tol = (1e-5 if dtype == "f32" else 1e-13) * np.sqrt(K) * np.abs(ref).max()
    assert np.abs(Cd.cpu().numpy().astype(np.float64) - ref).max() < tol


@pytest.mark.parametrize("kernel", [1, 2])
@pytest.mark.parametrize("shape", [(100, 2000, 1700, 1702), (6, 1100, 1100, 1100), (130, 1201, 516, 518),
                                   (1, 65, 14, 16), (896, 130, 300, 300), (200, 3000, 2998, 3000),
                                   (17, 64, 16, 16), (5, 40, 33, 35)])
def test_gemm_nt_f64_split(T, kernel, shape):
    """C = -1.3 A B^T with a split workspace (the NNLS's c = -Y U and anchor
    products) on the thin kernel (1: row passes, reduction tails, columns past
    a 64-column group) and on the tiled split-K kernel (2), against FP64.  The
    odd leading dimension of the last shape sends it to the tiled kernel in
    both cases."""
    from schwarmax_amd import _lib, tune
    M, N, K, ldb = shape
    rng = np.random.default_rng(M + 7 * N + K)
    A = rng.normal(size=(M, K))
    B = rng.normal(size=(N, ldb))
    ref = -1.3 * A @ B[:, :K].T
    Ad, Bd = T.tensor(A, device='cuda'), T.tensor(B, device='cuda')
    Cd = T.full((M, N), float('nan'), dtype=T.float64, device='cuda')
    work = T.empty(16 * M * N, dtype=T.float64, device='cuda')
    before = tune.get_knob(tune.KNOB_GEMM64_KERNEL)
    try:
        tune.set_knob(tune.KNOB_GEMM64_KERNEL, kernel)
        _lib.check(_lib.lib().smx_gemm_nt_f64_split(M, N, K, -1.3, _lib.ptr(Ad), K, _lib.ptr(Bd), ldb,
                                                    _lib.ptr(Cd), N, _lib.ptr(work), work.numel(),
                                                    _lib.stream_ptr()), 'gemm_nt_f64_split')
    finally:
        tune.set_knob(tune.KNOB_GEMM64_KERNEL, before)
    got = Cd.cpu().numpy()
    assert np.isfinite(got).all()
    assert np.abs(got - ref).max() < 1e-13 * np.sqrt(K) * np.abs(ref).max()


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
    work = T.empty(n * n + n, dtype=T.float64, device='cuda')
    _lib.check(_lib.lib().smx_spd_inverse_f64(_lib.ptr(Ad), n, _lib.ptr(work), _lib.stream_ptr()), 'inv')
    got = Ad.cpu().numpy()
    ref = np.linalg.inv(Q)
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert err < 1e-13 * np.linalg.cond(Q), err
    assert np.abs(got - got.T).max() <= 1e-12 * np.abs(got).max()


def test_spd_inverse_large(T):
    """A 16,500-row inverse, whose top node takes rocBLAS syrkx and whose
    lower nodes take full dgemm updates: residual A X - I and symmetry on the
    device (the CPU inverse would take minutes)."""
    from schwarmax_amd import _lib
    n = 16500
    g = T.Generator(device='cuda')
    g.manual_seed(7)
    X = T.rand((n + n // 4, n), dtype=T.float64, device='cuda', generator=g)
    X *= (T.rand(X.shape, device='cuda', generator=g) < 0.2)
    X[:, : n // 3] *= 30.0
    Q = X.T @ X
    del X
    Q += (Q.diagonal().sum() / n) * T.eye(n, dtype=T.float64, device='cuda')
    A = Q.clone()
    work = T.empty(n * n + n, dtype=T.float64, device='cuda')
    _lib.check(_lib.lib().smx_spd_inverse_f64(_lib.ptr(A), n, _lib.ptr(work), _lib.stream_ptr()), 'inv')
    del work
    R = Q @ A
    R.diagonal().sub_(1.0)
    res = R.abs().max().item()
    asym = (A - A.T).abs().max().item() / A.abs().max().item()
    assert res < 1e-10, res
    assert asym <= 1e-12, asym
