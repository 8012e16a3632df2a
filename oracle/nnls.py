"""ADMM NNLS with a shared Cholesky factor, FP64.

Restates nnls.py:6-66 (single right-hand side) and nnls.py:68-166 (bootstrap:
one Gram / Cholesky, K right-hand sides).  The iterate sequence is reproduced
exactly (same alpha = 1.6, rho = tr(Q)/n, initial value sum(y_Rzphi)/n, fixed
``maxiter``); scipy's cho_solve plays jax.scipy.linalg.cho_solve.
"""
import numpy as np
import scipy.linalg as sla

EPS = 1e-8


def gamma_safe(gamma_kin):
    """nnls.py:15 / 102."""
    return np.where(np.abs(gamma_kin) > EPS, gamma_kin, EPS)


def design_matrix(A_Rzphi, A_xy, A_h, sig_Rzphi, sig_xy, sig_h, gamma_kin,
                  w_rzphi=1.0, w_xy=1.0, w_h=1.0):
    """U of nnls.py:107-113.  A_h, sig_h: lists of the four kinematic blocks."""
    g = gamma_safe(gamma_kin)
    blocks = [w_rzphi * A_Rzphi / (sig_Rzphi[:, None] + EPS),
              w_xy * A_xy / (sig_xy[:, None] + EPS)]
    for Ak, sk in zip(A_h, sig_h):
        blocks.append(w_h * Ak / (g[:, None] * (sk[:, None] + EPS)))
    return np.vstack(blocks)


def rhs_vectors(y_Rzphi, sig_Rzphi, y_xy_boot, y_h_boot, sig_xy, sig_h,
                w_rzphi=1.0, w_xy=1.0, w_h=1.0):
    """Stacked data vectors y_i (nnls.py:126-140); y_xy_boot (K, B)."""
    K = y_xy_boot.shape[0]
    y_rz = w_rzphi * y_Rzphi / (sig_Rzphi + EPS)
    parts = [np.broadcast_to(y_rz, (K, y_rz.size)), w_xy * y_xy_boot / (sig_xy + EPS)]
    for yk, sk in zip(y_h_boot, sig_h):
        parts.append(w_h * yk / (sk + EPS))
    return np.concatenate(parts, axis=1)


def gram_factor(U, lambda_reg):
    """Q, rho, Cholesky of Q + rho I (nnls.py:115-121)."""
    n = U.shape[1]
    Q = U.T @ U + (lambda_reg / n) * np.eye(n)
    rho = np.trace(Q) / n
    L = np.linalg.cholesky(Q + rho * np.eye(n))
    return Q, rho, L


def admm(L, rho, c, w_init_val, maxiter, alpha=1.6, return_state=False):
    """ADMM iterations (nnls.py:146-163); c has shape (n,) or (K, n)."""
    C = np.atleast_2d(c).T                      # (n, K)
    n, K = C.shape
    w = np.full((n, K), w_init_val)
    z = w.copy()
    u = np.zeros((n, K))
    for _ in range(maxiter):
        rhs = rho * (z - u) - C
        w = sla.cho_solve((L, True), rhs)
        w_hat = alpha * w + (1.0 - alpha) * z
        z_new = np.maximum(0.0, w_hat + u)
        u = u + w_hat - z_new
        z = z_new
    out = z.T if np.ndim(c) == 2 else z[:, 0]
    if return_state:
        return out, (w, z, u)
    return out


def solve_nnls_admm_bootstrap(A_Rzphi, A_xy, A_h, y_Rzphi, y_xy, gamma_kin,
                              y_xy_boot, y_h_boot, sig_Rzphi, sig_xy, sig_h,
                              lambda_reg=1.0, maxiter=200, w_rzphi=1.0, w_xy=1.0, w_h=1.0):
    """nnls.py:68-166.  Returns weights (K, n)."""
    U = design_matrix(A_Rzphi, A_xy, A_h, sig_Rzphi, sig_xy, sig_h, gamma_kin,
                      w_rzphi, w_xy, w_h)
    n = U.shape[1]
    _, rho, L = gram_factor(U, lambda_reg)
    Y = rhs_vectors(y_Rzphi, sig_Rzphi, y_xy_boot, y_h_boot, sig_xy, sig_h, w_rzphi, w_xy, w_h)
    c_all = -(Y @ U)
    return admm(L, rho, c_all, np.sum(y_Rzphi) / n, maxiter)


def solve_nnls_admm(A_Rzphi, A_xy, A_h, y_Rzphi, y_xy, gamma_kin, y_h,
                    sig_Rzphi, sig_xy, sig_h, lambda_reg=1.0, maxiter=200,
                    w_rzphi=1.0, w_xy=1.0, w_h=1.0):
    """nnls.py:6-66.  Returns weights (n,)."""
    U = design_matrix(A_Rzphi, A_xy, A_h, sig_Rzphi, sig_xy, sig_h, gamma_kin,
                      w_rzphi, w_xy, w_h)
    n = U.shape[1]
    _, rho, L = gram_factor(U, lambda_reg)
    Y = rhs_vectors(y_Rzphi, sig_Rzphi, np.asarray(y_xy)[None], [np.asarray(a)[None] for a in y_h],
                    sig_xy, sig_h, w_rzphi, w_xy, w_h)
    c = -(U.T @ Y[0])
    return admm(L, rho, c, np.sum(y_Rzphi) / n, maxiter)
