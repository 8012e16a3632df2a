"""Grouped bootstrap NNLS (GPU): smx_nnls_admm_group against each model's own
smx_nnls_admm solve and the FP64 oracle.

The grouped loop runs every model's ADMM iteration in one product launch and
one update launch, each model on its own run of blocks with its own split of
the reduction, so each model's weights must be bit-identical to its single
solve.  The cases cover uneven orbit counts (one not a multiple of 4, which
takes the one-element update for the whole group), one and two 128-row blocks
of right-hand sides (K = 6 and K = 130: non-temporal and cached loads of the
inverse's planes), and solves long enough for FP64 anchors (every 100
iterations) and delta products.
"""
import numpy as np
import pytest

from test_gpu_nnls import _problem

pytestmark = pytest.mark.gpu


def _inputs(seed, n, K):
    import torch
    from schwarmax_amd import nnls as GN, _lib
    A_R, A_xy, A_h, y_R, y_xy, gamma, yb_xy, yb_h, sR, sxy, sh = _problem(seed, n=n, K=K)
    d = _lib.dev
    lib = GN._stack_library(A_R, A_xy, A_h)
    data = (d(y_R), d(gamma), d(yb_xy), [d(v) for v in yb_h], d(sR), d(sxy), [d(v) for v in sh])
    return lib, data, (A_R, A_xy, A_h, y_R, y_xy, gamma, yb_xy, yb_h, sR, sxy, sh)


@pytest.mark.parametrize("ns,K,maxiter", [((300, 517, 1100), 6, 230), ((400, 1200), 130, 120)])
def test_group_equals_single_solves(ns, K, maxiter):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from schwarmax_amd import nnls as GN
    from oracle import nnls as ON
    P, B = 120, 40
    w5 = np.sqrt(5.0)
    cases = [_inputs(10 + i, n, K) for i, n in enumerate(ns)]
    group = GN.solve_group([c[0] for c in cases], P, B, [c[1] for c in cases], 1.0, maxiter,
                           w5, w5, 1.0)
    for i, (lib, data, host) in enumerate(cases):
        single, _ = GN.solve_nnls_admm_bootstrap_lib(lib, P, B, *data, lambda_reg=1.0,
                                                     maxiter=maxiter, w_rzphi=w5, w_xy=w5, w_h=1.0)
        same = torch.equal(group[i], single)
        print(f"model {i}: n={lib.shape[0]} K={K} grouped == single: {same}", flush=True)
        assert same, (i, float((group[i] - single).abs().max()))
    # and one model against the FP64 oracle (the single path's bar)
    A_R, A_xy, A_h, y_R, y_xy, gamma, yb_xy, yb_h, sR, sxy, sh = cases[0][2]
    ref = ON.solve_nnls_admm_bootstrap(A_R, A_xy, A_h, y_R, y_xy, gamma, yb_xy, yb_h, sR, sxy, sh,
                                       lambda_reg=1.0, maxiter=maxiter, w_rzphi=w5, w_xy=w5, w_h=1.0)
    got = group[0].cpu().numpy()
    err = np.linalg.norm(got - ref, axis=1) / np.linalg.norm(ref, axis=1)
    print(f"model 0 against the oracle: weights max {err.max():.2e}", flush=True)
    assert err.max() < 1e-4, err
