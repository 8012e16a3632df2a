"""FP64 restatement of the reference's bootstrap NNLS solve in torch, for the
full-shape parity tests on the GPU (where the NumPy oracle would take hours).

Follows nnls.py:101-163 step by step, as ``oracle.nnls`` does: Q = U^T U +
(lambda/n) I, rho = tr(Q)/n, L = chol(Q + rho I), c = -Y U, z = w = init,
u = 0, then ``maxiter`` times w = cho_solve(L, rho (z - u) - c),
w_hat = alpha w + (1 - alpha) z, z = max(0, w_hat + u), u += w_hat - z;
returns z.  Everything in float64 on the tensors' device.  Pinned against
``oracle.nnls.admm`` on the CPU (tests/test_oracle_nnls_misc.py).
"""
import torch


def admm_fp64(U, Y, lambda_reg, maxiter, w_init, alpha=1.6, progress=None, snapshots=None):
    """U (m, n), Y (K, m) float64 tensors.  Returns z (K, n) float64.
    ``progress``: called with the iteration count every 100 iterations.
    ``snapshots``: a dict {iteration: None}; filled with (z (K, n), KKT residual)
    after those iterations, the residual being max over right-hand sides of
    max_i |g_i| on z_i > 0 and max(-g_i, 0) on z_i = 0, g = Q z + c (the
    optimality conditions of min z^T Q z / 2 + c^T z over z >= 0), relative to
    max |c|."""
    n = U.shape[1]
    Q = U.T @ U
    Q.diagonal().add_(lambda_reg / n)
    rho = torch.trace(Q) / n
    Q.diagonal().add_(rho)
    L = torch.linalg.cholesky(Q)
    if snapshots:
        Q.diagonal().sub_(rho)                       # back to U^T U + lambda/n I
    else:
        del Q
    C = -(Y @ U)                                     # (K, n)
    z = torch.full_like(C, float(w_init))
    u = torch.zeros_like(C)
    CT = C.T.contiguous()
    zT, uT = z.T.contiguous(), u.T.contiguous()
    for it in range(maxiter):
        if progress is not None and it % 100 == 0:
            progress(it)
        w = torch.cholesky_solve(rho * (zT - uT) - CT, L)      # (n, K)
        wh = alpha * w + (1.0 - alpha) * zT
        zn = torch.clamp(wh + uT, min=0.0)
        uT = uT + wh - zn
        zT = zn
        if snapshots and (it + 1) in snapshots:
            g = Q @ zT + CT
            viol = torch.where(zT > 0, g.abs(), torch.clamp(-g, min=0.0))
            kkt = (viol.amax(dim=0) / CT.abs().amax(dim=0)).max().item()
            snapshots[it + 1] = (zT.T.contiguous(), kkt)
    return zT.T.contiguous()
