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
"""
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
    return theta, dd, sched, info


def _fp64_admm_on_device(U, Y, lambda_reg, maxiter, w_init):
    import torch
    from fp64_admm import admm_fp64
    t0 = time.time()
    W = admm_fp64(torch.tensor(U, device='cuda'), torch.tensor(Y, device='cuda'), lambda_reg,
                  maxiter, w_init,
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
    lib, P, B = info['lib'], info['P'], info['B']
    ph, pb = L.theta_to_params(theta)
    w5 = np.sqrt(5.0)
    out, _ = M.model_from_library(lib, P, B, pb, dd, 1000, 1.0, (w5, w5, 1.0))
    logl_gpu = float(L.bootstrap_loglike(out, dd)[0])
    W_gpu = out['weights'].cpu().numpy()
    print(f"GPU logL {logl_gpu:.10g}", flush=True)
    libh = lib.cpu().numpy()
    del lib, out
    info.pop('lib')
    torch.cuda.empty_cache()
    cols = {'Rzphi': libh[:, :P], 'surface_density': libh[:, P:P + B]}
    for k in (1, 2, 3, 4):
        cols['h%d' % k] = libh[:, P + B * k:P + B * (k + 1)]
    md, _ = OM.model_from_library(cols, dd, pb, 1000, 1.0, admm=_fp64_admm_on_device)
    logl_ref = float(OM.bootstrap_loglike(md, dd)[0])
    rel = abs(logl_gpu - logl_ref) / abs(logl_ref)
    w_err = rowwise_rel(W_gpu, md['weights'])
    print(f"plain-20k: logL gpu={logl_gpu:.10g} FP64={logl_ref:.10g} rel={rel:.2e}; "
          f"weights {w_err:.2e}", flush=True)
    assert np.isfinite(logl_ref) and np.all(np.isfinite(W_gpu))
    assert rel < 1e-4, rel
    assert w_err < 1e-4, w_err
