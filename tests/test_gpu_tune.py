"""Autotuner (GPU): the ADMM product's timing entry point and occupancy knob,
the Gram / inverse / FP64-GEMM knobs timed on a one-iteration solve, and that
tuned settings leave the solve within the oracle's bars."""
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
        assert res['loads'] in (0, 1, 2) and set(res['loads_ms']) == {0, 1, 2}
        assert tune.get_knob(tune.KNOB_ADMM_LOADS) == res['loads']
        # every setting solves the same problem to the oracle's bars
        A_R, A_xy, A_h, y_R, y_xy, gamma, yb_xy, yb_h, sR, sxy, sh = _problem(2, n=300, K=12)
        w5 = np.sqrt(5.0)
        ref = ON.solve_nnls_admm_bootstrap(A_R, A_xy, A_h, y_R, y_xy, gamma, yb_xy, yb_h, sR, sxy,
                                           sh, lambda_reg=1.0, maxiter=300, w_rzphi=w5, w_xy=w5,
                                           w_h=1.0)
        for occ, loads in [(o, 0) for o in tune.CANDIDATES] + [(2, 1), (2, 2)]:
            tune.set_occupancy(occ)
            tune.set_knob(tune.KNOB_ADMM_LOADS, loads)
            got = GN.solve_nnls_admm_bootstrap(A_R, A_xy, *A_h, y_R, y_xy, gamma, yb_xy, *yb_h, sR,
                                               sxy, *sh, lambda_reg=1.0, maxiter=300, w_rzphi=w5,
                                               w_xy=w5, w_h=1.0).cpu().numpy()
            err = np.linalg.norm(got - ref, axis=1) / np.linalg.norm(ref, axis=1)
            assert err.max() < 1e-3, (occ, err)
        with pytest.raises(RuntimeError):
            tune.set_occupancy(9)
    finally:
        tune.set_occupancy(before)
        tune.set_knob(tune.KNOB_ADMM_LOADS, 0)


def test_autotune_factor_knobs():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from schwarmax_amd import tune, nnls as GN
    from oracle import nnls as ON
    from test_gpu_nnls import _problem
    knobs = (tune.KNOB_GRAM_WAVES, tune.KNOB_INV_SYRKX_MIN, tune.KNOB_GEMM64_SPLIT_TARGET,
             tune.KNOB_GEMM64_KERNEL)
    before = [tune.get_knob(k) for k in knobs]
    try:
        res = tune.autotune_factor(n=3000, m=1500, K=20, repeats=1)
        print(res)
        assert set(res['ms']['gram_waves']) == set(tune.GRAM_WAVES)
        assert set(res['ms']['gemm64_kernel']) == set(tune.GEMM64_KERNELS)
        if res['gemm64_kernel'] == 2:                  # the tiled kernel's split target is tuned
            assert set(res['ms']['split_target']) == set(tune.SPLIT_TARGETS)
        else:
            assert 'split_target' not in res['ms']
        assert 3001 in res['ms']['syrkx_min']          # every node on dgemm
        assert all(t > 0 for r in res['ms'].values() for t in r.values())
        assert tune.get_knob(tune.KNOB_GRAM_WAVES) == res['gram_waves']
        assert tune.get_knob(tune.KNOB_INV_SYRKX_MIN) == res['syrkx_min']
        assert tune.get_knob(tune.KNOB_GEMM64_KERNEL) == res['gemm64_kernel']
        assert tune.get_knob(tune.KNOB_GEMM64_SPLIT_TARGET) == res['split_target']
        A_R, A_xy, A_h, y_R, y_xy, gamma, yb_xy, yb_h, sR, sxy, sh = _problem(3, n=300, K=12)
        w5 = np.sqrt(5.0)
        ref = ON.solve_nnls_admm_bootstrap(A_R, A_xy, A_h, y_R, y_xy, gamma, yb_xy, yb_h, sR, sxy,
                                           sh, lambda_reg=1.0, maxiter=200, w_rzphi=w5, w_xy=w5,
                                           w_h=1.0)
        # the 4-wave Gram and the syrkx updates at every node, with either
        # FP64 GEMM kernel, against the oracle
        for k, v in ((tune.KNOB_GRAM_WAVES, 4), (tune.KNOB_INV_SYRKX_MIN, 256),
                     (tune.KNOB_GEMM64_SPLIT_TARGET, 4096)):
            tune.set_knob(k, v)
        for kern in tune.GEMM64_KERNELS:
            tune.set_knob(tune.KNOB_GEMM64_KERNEL, kern)
            got = GN.solve_nnls_admm_bootstrap(A_R, A_xy, *A_h, y_R, y_xy, gamma, yb_xy, *yb_h, sR,
                                               sxy, *sh, lambda_reg=1.0, maxiter=200, w_rzphi=w5,
                                               w_xy=w5, w_h=1.0).cpu().numpy()
            err = np.linalg.norm(got - ref, axis=1) / np.linalg.norm(ref, axis=1)
            print(f"weights with 4-wave Gram, syrkx everywhere, FP64 GEMM kernel {kern}: "
                  f"{err.max():.2e}")
            assert err.max() < 1e-4, err
    finally:
        for k, v in zip(knobs, before):
            tune.set_knob(k, v)
