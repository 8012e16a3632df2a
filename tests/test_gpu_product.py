"""The ADMM product kernel (admm_gemm_kernel, through smx_nnls_operator_apply
and smx_nnls_operator_apply_hi) against R M in FP64, at shapes that exercise
its tiling: one, two and three 128-row blocks of R (K = 1, 100, 130, 200,
256, 300; more than one block reads M with plain loads), splits of odd
and even chunk counts, a split of a single chunk and empty splits (n = 1,100:
40 chunks of 32 over 16 splits), zero padding of n to a multiple of 256, and
three occupancy settings (which change the split count).

M is formed in torch FP64 from the same U, row scales and lambda:
M = (U^T U + (lambda/n) I + rho I)^-1, rho = tr(U^T U + (lambda/n) I) / n
(nnls.py:115-121).  Bars (row-wise relative 2-norm): 1e-5 for the full
product (fp16 hi/lo splits carry 2^-22, FP32 accumulation over n / S terms),
3e-3 for the hi-only product (2^-11 per operand).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rowwise(got, ref):
    num = np.linalg.norm(got - ref, axis=1)
    den = np.maximum(np.linalg.norm(ref, axis=1), 1e-300)
    return float((num / den).max())


@pytest.mark.parametrize("n,m,K,occ", [(1100, 700, 1, 2), (1100, 700, 100, 2), (1100, 700, 130, 2),
                                       (1100, 700, 200, 4), (1100, 700, 300, 2), (3000, 1500, 256, 1),
                                       (5000, 3000, 100, 1), (5000, 3000, 100, 2),
                                       (4100, 2500, 60, 4)])
def test_product_matches_fp64(n, m, K, occ):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from schwarmax_amd import jackknife as J, tune
    rng = np.random.default_rng(n + K + occ)
    lam = 1.0
    lib_np = rng.random((n, m + 3))                      # ld_lib = m + 3 > m
    rs_np = rng.uniform(0.5, 2.0, m)
    U = lib_np[:, :m].T * rs_np[:, None]
    Q = U.T @ U + (lam / n) * np.eye(n)
    rho = float(np.trace(Q)) / n
    R_np = rng.standard_normal((K, n)).astype(np.float32)
    R_np[0, : n // 3] *= 1e-3                            # rows of mixed scale
    lib = torch.tensor(lib_np, dtype=torch.float64, device='cuda')
    row_scale = torch.tensor(rs_np, dtype=torch.float64, device='cuda')
    R = torch.tensor(R_np, device='cuda')
    before = tune.get_occupancy()
    try:
        tune.set_occupancy(occ)
        op = J._Operator(lib, m, row_scale, lam, K_max=K)
        full = op.apply(R).cpu().numpy()
        hi = op.apply(R, hi_only=True).cpu().numpy()
    finally:
        tune.set_occupancy(before)
    # the device Gram is an fp16 hi/lo split product with FP32 segments, so
    # its trace (and rho) agrees with FP64 to ~1e-8
    assert abs(op.rho - rho) <= 1e-6 * rho
    Mref = np.linalg.inv(Q + rho * np.eye(n))            # FP64 on the host
    ref = R_np.astype(np.float64) @ Mref
    e_full, e_hi = _rowwise(full, ref), _rowwise(hi, ref)
    print(f"n={n} K={K} occ={occ}: full {e_full:.2e} hi {e_hi:.2e}", flush=True)
    assert e_full <= 1e-5
    assert e_hi <= 3e-3
