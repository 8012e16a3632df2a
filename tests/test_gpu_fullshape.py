"""Parity at the benchmark shape: the plain 20,000-orbit evaluation (n = 20,000
columns, 1,000 apertures, a 32 x 24 x 16 grid, so m = 17,288 design rows;
K = 100 bootstrap right-hand sides; T = 1,000 ADMM iterations).

The device library is built once; from it, the GPU path (fp16-split MFMA Gram,
FP64 inverse, split ADMM products) and an FP64 restatement of the reference's
solve (torch FP64 Cholesky and cho_solve on the device, tests/fp64_admm.py,
pinned against the oracle on the CPU) each produce the weights, and the
oracle's own maps and bootstrap chi^2 (model.py:14-68, likelihood.py:184-244)
turn the FP64 weights into the reference logL.  Bars: logL within 1e-4
relative, the weights within 1e-4 relative (row-wise 2-norm).

A second test shows why the solve must replay the reference's 1,000
iterations rather than stop at convergence: at this shape the T = 1,000
iterate is far from the QP's minimizer.
"""
import os
import time

import numpy as np
import pytest

from galaxies import rowwise_rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def plain20k():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from schwarmax_amd import synth
    t0 = time.time()
    theta, dd, sched, info = synth.observed_workload('plain-20k', seed=0, return_library=True)
    print(f"plain-20k workload: valid {info['valid']:.4f}, {time.time() - t0:.1f} s", flush=True)
    assert info['valid'] >= 0.8
    info['lib'] = info['lib'].cpu().numpy()          # on the host between the tests
    torch.cuda.empty_cache()
    return theta, dd, sched, info


def _columns(libh, P, B):
    cols = {'Rzphi': libh[:, :P], 'surface_density': libh[:, P:P + B]}
    for k in (1, 2, 3, 4):
        cols['h%d' % k] = libh[:, P + B * k:P + B * (k + 1)]
    return cols


def _fp64_admm_on_device(U, Y, lambda_reg, maxiter, w_init, snapshots=None):
    import torch
    from fp64_admm import admm_fp64
    t0 = time.time()
    W = admm_fp64(torch.tensor(U, device='cuda'), torch.tensor(Y, device='cuda'), lambda_reg,
                  maxiter, w_init, snapshots=snapshots,
                  progress=lambda it: (torch.cuda.synchronize(),
                                       print(f"  FP64 ADMM iteration {it}, {time.time() - t0:.1f} s",
                                             flush=True)))
    torch.cuda.synchronize()
    print(f"FP64 reference solve: {time.time() - t0:.1f} s", flush=True)
    return W.cpu().numpy()


def test_plain20k_weights_and_loglike(plain20k):
    import torch
    from schwarmax_amd import model as M, likelihood as L
    from oracle import model as OM
    theta, dd, sched, info = plain20k
    P, B = info['P'], info['B']
    ph, pb = L.theta_to_params(theta)
    w5 = np.sqrt(5.0)
    lib = torch.tensor(info['lib'], device='cuda')
    out, _ = M.model_from_library(lib, P, B, pb, dd, 1000, 1.0, (w5, w5, 1.0))
    logl_gpu = float(L.bootstrap_loglike(out, dd)[0])
    W_gpu = out['weights'].cpu().numpy()
    print(f"GPU logL {logl_gpu:.10g}", flush=True)
    del lib, out
    torch.cuda.empty_cache()
    md, _ = OM.model_from_library(_columns(info['lib'], P, B), dd, pb, 1000, 1.0,
                                  admm=_fp64_admm_on_device)
    logl_ref = float(OM.bootstrap_loglike(md, dd)[0])
    rel = abs(logl_gpu - logl_ref) / abs(logl_ref)
    w_err = rowwise_rel(W_gpu, md['weights'])
    print(f"plain-20k: logL gpu={logl_gpu:.10g} FP64={logl_ref:.10g} rel={rel:.2e}; "
          f"weights {w_err:.2e}", flush=True)
    assert np.isfinite(logl_ref) and np.all(np.isfinite(W_gpu))
    assert rel < 1e-4, rel
    assert w_err < 1e-4, w_err


def test_plain20k_admm_iterate_is_not_converged(plain20k):
    """The reference's T = 1,000 ADMM iterate (nnls.py:146-165) against longer
    runs of the same FP64 iteration.  If T = 1,000 had converged to the
    strictly convex QP's unique minimizer, later iterates would agree with it
    to well within the 1e-4 bars; they do not, so a solver that exits at
    convergence (or reaches the minimizer by another route) would not
    reproduce the reference, and the device solve replays all T iterations.
    Iteration counts: SMX_CONVERGENCE_ITERS (comma-separated, default
    500,1000,2000,4000)."""
    import torch
    from schwarmax_amd import model as M, likelihood as L
    from oracle import model as OM
    theta, dd, sched, info = plain20k
    P, B = info['P'], info['B']
    ph, pb = L.theta_to_params(theta)
    iters = [int(v) for v in os.environ.get('SMX_CONVERGENCE_ITERS', '500,1000,2000,4000').split(',')]
    assert 1000 in iters and max(iters) > 1000
    snaps = {t: None for t in iters}

    def admm(U, Y, lam, maxiter, w_init):
        return _fp64_admm_on_device(U, Y, lam, max(iters), w_init, snapshots=snaps)
    OM.model_from_library(_columns(info['lib'], P, B), dd, pb, 1000, 1.0, admm=admm)
    lib = torch.tensor(info['lib'], device='cuda')
    t = M.prepare_targets(lib.shape[0], pb, dd)
    rows = {}
    for it in iters:
        W, kkt = snaps[it]
        d3, d2, h1, h2, h3, h4, V, sig = M.compute_model_single(W, lib, P, B, t['gamma_kin'],
                                                                t['v0_ap'], t['s_ap'])
        md = {'density_3d': t['y_R'], 'density_3d_model': d3,
              'surface_density': d2 * t['mmpo'] * t['lm'], 'h1': h1, 'h2': h2, 'h3': h3, 'h4': h4,
              'V': V, 'sigma': sig}
        rows[it] = (float(L.bootstrap_loglike(md, dd)[0]), W.cpu().numpy(), kkt)
    ref_l, ref_w, _ = rows[1000]
    last = max(iters)
    print(f"{'T':>6} {'logL':>18} {'vs T=1000':>10} {'vs T=%d' % last:>10} "
          f"{'weights vs T=1000':>18} {'weights vs T=%d' % last:>16} {'KKT':>9}")
    for it in iters:
        lv, w, kkt = rows[it]
        print(f"{it:>6} {lv:>18.8f} {abs(lv - ref_l) / abs(ref_l):>10.2e} "
              f"{abs(lv - rows[last][0]) / abs(ref_l):>10.2e} {rowwise_rel(w, ref_w):>18.2e} "
              f"{rowwise_rel(w, rows[last][1]):>16.2e} {kkt:>9.2e}", flush=True)
    nxt = min(v for v in iters if v > 1000)
    assert abs(rows[nxt][0] - ref_l) / abs(ref_l) > 1e-3
    assert rowwise_rel(rows[nxt][1], ref_w) > 1e-2
