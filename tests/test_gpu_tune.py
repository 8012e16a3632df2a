"""ADMM product autotuner (GPU): the timing entry point, the occupancy knob,
and that a tuned setting leaves the solve within the oracle's bars."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_autotune_and_solve():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from schwarmax_amd import tune, nnls as GN
    from oracle import nnls as ON
    from test_gpu_nnls import _problem
    before = tune.get_occupancy()
    try:
        res = tune.autotune_admm(n=3000, m=1500, K=20, reps=3)
        print(res)
        assert set(res['ms']) == set(tune.CANDIDATES)
        for occ, (tf, td, tw) in res['ms'].items():
            assert tf > 0 and td > 0 and min(tf, td) <= tw <= max(tf, td)
        assert tune.get_occupancy() == res['occupancy']
        # every setting solves the same problem to the oracle's bars
        A_R, A_xy, A_h, y_R, y_xy, gamma, yb_xy, yb_h, sR, sxy, sh = _problem(2, n=300, K=12)
        w5 = np.sqrt(5.0)
        ref = ON.solve_nnls_admm_bootstrap(A_R, A_xy, A_h, y_R, y_xy, gamma, yb_xy, yb_h, sR, sxy,
                                           sh, lambda_reg=1.0, maxiter=300, w_rzphi=w5, w_xy=w5,
                                           w_h=1.0)
        for occ in tune.CANDIDATES:
            tune.set_occupancy(occ)
            got = GN.solve_nnls_admm_bootstrap(A_R, A_xy, *A_h, y_R, y_xy, gamma, yb_xy, *yb_h, sR,
                                               sxy, *sh, lambda_reg=1.0, maxiter=300, w_rzphi=w5,
                                               w_xy=w5, w_h=1.0).cpu().numpy()
            err = np.linalg.norm(got - ref, axis=1) / np.linalg.norm(ref, axis=1)
            assert err.max() < 1e-3, (occ, err)
        with pytest.raises(RuntimeError):
            tune.set_occupancy(9)
    finally:
        tune.set_occupancy(before)
