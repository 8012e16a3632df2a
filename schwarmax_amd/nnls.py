"""ADMM NNLS drop-ins (nnls.py:6-166) backed by smx_nnls_admm.

The GPU solver forms U^T U in FP32, inverts Q + rho I once in FP64 and runs
each ADMM iteration as one product with FP32-level accuracy (fp16 hi/lo splits
on MFMA; while R moves little, the product of the inverse's hi plane with the
change of R since the last full product) followed by the projection / dual
update (iterates in FP64).  Same iterate definition as the reference:
alpha = 1.6, rho = tr(Q)/n, start w = z = sum(y_Rzphi)/n, u = 0, ``maxiter``
iterations, result z.  ``tol > 0`` lets a right-hand side stop once its iterate
moves by less than tol * max|z| over 25 iterations.
"""
import ctypes

import torch

from . import _lib

EPS = 1e-8

# While stage timing is on: ADMM products taken and how many were full (the
# others multiply the inverse's hi plane by R - R_base).
STATS = {'products': 0, 'full_products': 0}


def gamma_safe(gamma_kin):
    """nnls.py:15 / 102."""
    return torch.where(gamma_kin.abs() > EPS, gamma_kin, torch.full_like(gamma_kin, EPS))


def model_maps(lib, m, W):
    """(K, m) = W (K, n) @ lib[:, :m] in FP64 on the GPU (model.py:49-56)."""
    L = _lib.lib()
    W = W.contiguous()
    K, n = W.shape
    out = torch.empty((K, m), dtype=torch.float64, device='cuda')
    _lib.check(L.smx_model_maps(_lib.ptr(lib), lib.shape[1], n, m, _lib.ptr(W), K, _lib.ptr(out),
                                _lib.stream_ptr()), 'smx_model_maps')
    return out


def admm_solve(lib, m, row_scale, Y, lambda_reg, maxiter, w_init, tol=0.0):
    """Core call: weights (K, n) for data rows Y (K, m)."""
    L = _lib.lib()
    n = lib.shape[0]
    K = Y.shape[0]
    ws = torch.empty(int(L.smx_nnls_workspace_bytes(n, m, K)), dtype=torch.uint8, device='cuda')
    W = torch.empty((K, n), dtype=torch.float64, device='cuda')
    iters = ctypes.c_int32(0)
    _lib.check(L.smx_nnls_admm(_lib.ptr(lib), lib.shape[1], n, m, _lib.ptr(row_scale.contiguous()),
                               _lib.ptr(Y.contiguous()), K, float(lambda_reg), int(maxiter),
                               float(w_init), float(tol), _lib.ptr(ws), _lib.ptr(W),
                               ctypes.byref(iters), _lib.stream_ptr()), 'smx_nnls_admm')
    if _lib.TIMING['on']:
        full = ctypes.c_int32(0)
        _lib.check(L.smx_nnls_admm_full_products(n, m, K, _lib.ptr(ws), ctypes.byref(full),
                                                 _lib.stream_ptr()), 'smx_nnls_admm_full_products')
        STATS['products'] += int(iters.value)
        STATS['full_products'] += int(full.value)
    return W, int(iters.value)


def _scales_and_rhs(P, B, y_R, gamma_kin, y_xy_boot, y_h_boot, sig_R, sig_xy, sig_h,
                    w_rzphi, w_xy, w_h):
    g = gamma_safe(gamma_kin)
    row_scale = torch.cat([w_rzphi / (sig_R + EPS), w_xy / (sig_xy + EPS)] +
                          [w_h / (g * (sk + EPS)) for sk in sig_h])
    K = y_xy_boot.shape[0]
    y_rz = w_rzphi * y_R / (sig_R + EPS)
    Y = torch.cat([y_rz[None, :].expand(K, -1), w_xy * y_xy_boot / (sig_xy + EPS)] +
                  [w_h * yk / (sk + EPS) for yk, sk in zip(y_h_boot, sig_h)], dim=1)
    return row_scale, Y.contiguous()


def solve_nnls_admm_bootstrap_lib(lib, P, B, y_R, gamma_kin, y_xy_boot, y_h_boot, sig_R, sig_xy,
                                  sig_h, lambda_reg=1.0, maxiter=200, w_rzphi=1.0, w_xy=1.0, w_h=1.0,
                                  tol=0.0):
    """Bootstrap solve on library columns lib (n, >= P+5B) (nnls.py:68-166)."""
    row_scale, Y = _scales_and_rhs(P, B, y_R, gamma_kin, y_xy_boot, y_h_boot, sig_R, sig_xy,
                                   sig_h, w_rzphi, w_xy, w_h)
    n = lib.shape[0]
    W, it = admm_solve(lib, P + 5 * B, row_scale, Y, lambda_reg, maxiter,
                       float(torch.sum(y_R).item()) / n, tol)
    return W, {'iters': it}


def group_bytes(ns, m, K):
    """Device bytes of a grouped solve of models with orbit counts ns: the
    models' states and the shared scratch, each rounded up to 256 bytes (a
    size query, no GPU)."""
    L = _lib.load_library()
    r = lambda b: (int(b) + 255) // 256 * 256  # noqa: E731
    return (sum(r(L.smx_nnls_group_state_bytes(int(n), m, K)) for n in ns) +
            r(L.smx_nnls_group_scratch_bytes(max(int(n) for n in ns), m, K)))


def solve_group(libs, P, B, data, lambda_reg=1.0, maxiter=200, w_rzphi=1.0, w_xy=1.0, w_h=1.0,
                arena=None):
    """Bootstrap solves (nnls.py:68-166) of several models in grouped launches
    (``smx_nnls_admm_group``: per ADMM iteration, one product launch and one
    update launch for all models).  libs[g] (n_g, >= P+5B) are the models'
    library columns; data[g] = (y_R, gamma_kin, y_xy_boot, y_h_boot, sig_R,
    sig_xy, sig_h) as for solve_nnls_admm_bootstrap_lib, with one K for all.
    ``arena``: optional uint8 CUDA tensor of at least ``group_bytes`` bytes
    from which the states and the scratch are taken (a caller solving many
    groups keeps one arena instead of allocating per group).
    Returns each model's weights (K, n_g), equal to its single solve (tol 0)."""
    L = _lib.lib()
    if len(libs) != len(data):
        raise ValueError("solve_group: one data tuple per library")
    if len(libs) > _lib.NNLS_GROUP_MAX:
        h = _lib.NNLS_GROUP_MAX
        return (solve_group(libs[:h], P, B, data[:h], lambda_reg, maxiter, w_rzphi, w_xy, w_h,
                            arena) +
                solve_group(libs[h:], P, B, data[h:], lambda_reg, maxiter, w_rzphi, w_xy, w_h,
                            arena))
    G = len(libs)
    if G == 0:
        return []
    m = P + 5 * B
    f64 = torch.float64
    probs = (_lib.SmxNnlsProblem * G)()
    keep, Ws, K = [], [], None
    Ks = {int(d[2].shape[0]) for d in data}
    if len(Ks) != 1:
        raise ValueError("solve_group: every model needs the same number of right-hand sides")
    K0 = Ks.pop()
    need = group_bytes([lib.shape[0] for lib in libs], m, K0)
    if arena is None:
        arena = torch.empty(need, dtype=torch.uint8, device='cuda')
    elif arena.dtype != torch.uint8 or not arena.is_cuda or arena.numel() < need:
        raise ValueError(f"solve_group: the arena must be a uint8 CUDA tensor of >= {need} bytes")
    off = 0

    def take(nbytes):
        nonlocal off
        t = arena[off:off + int(nbytes)]
        off += (int(nbytes) + 255) // 256 * 256
        return t
    for g, (lib, d) in enumerate(zip(libs, data)):
        y_R, gamma_kin, y_xy_boot, y_h_boot, sig_R, sig_xy, sig_h = d
        row_scale, Y = _scales_and_rhs(P, B, y_R, gamma_kin, y_xy_boot, y_h_boot, sig_R, sig_xy,
                                       sig_h, w_rzphi, w_xy, w_h)
        row_scale = row_scale.contiguous()
        n = lib.shape[0]
        if K is None:
            K = Y.shape[0]
        elif Y.shape[0] != K:
            raise ValueError("solve_group: every model needs the same number of right-hand sides")
        if lib.shape[1] < m:
            raise ValueError("solve_group: library rows shorter than P + 5B")
        state = take(L.smx_nnls_group_state_bytes(n, m, K))
        W = torch.empty((K, n), dtype=f64, device='cuda')
        probs[g] = _lib.SmxNnlsProblem(_lib.ptr(lib, f64).value, lib.shape[1], n,
                                       _lib.ptr(row_scale, f64).value, _lib.ptr(Y, f64).value,
                                       float(torch.sum(y_R).item()) / n, state.data_ptr(),
                                       W.data_ptr())
        keep += [row_scale, Y, state]
        Ws.append(W)
    n_max = max(lib.shape[0] for lib in libs)
    scratch = take(L.smx_nnls_group_scratch_bytes(n_max, m, K))
    _lib.check(L.smx_nnls_admm_group(G, probs, m, K, float(lambda_reg), int(maxiter),
                                     _lib.ptr(scratch), _lib.stream_ptr()), 'smx_nnls_admm_group')
    if _lib.TIMING['on']:
        for g, lib in enumerate(libs):
            full = ctypes.c_int32(0)
            _lib.check(L.smx_nnls_group_full_products(lib.shape[0], m, K, ctypes.c_void_p(probs[g].state),
                                                      ctypes.byref(full), _lib.stream_ptr()),
                       'smx_nnls_group_full_products')
            STATS['products'] += int(maxiter)
            STATS['full_products'] += int(full.value)
    del keep, scratch
    return Ws


def _stack_library(A_Rzphi, A_xy, A_h):
    A = [_lib.dev(a) for a in [A_Rzphi, A_xy] + list(A_h)]
    return torch.cat(A, dim=0).t().contiguous()      # (n, P + 5B)


def solve_nnls_admm_bootstrap(A_Rzphi, A_xy, A_h1, A_h2, A_h3, A_h4, y_Rzphi, y_xy, gamma_kin,
                              y_xy_boot, y_h1_boot, y_h2_boot, y_h3_boot, y_h4_boot,
                              sig_Rzphi, sig_xy, sig_A1, sig_A2, sig_A3, sig_A4,
                              lambda_reg=1, maxiter=200, w_rzphi=1.0, w_xy=1.0, w_h=1.0, tol=0.0):
    """nnls.py:68-166.  A_* are (bins, n); returns weights (K, n)."""
    lib = _stack_library(A_Rzphi, A_xy, (A_h1, A_h2, A_h3, A_h4))
    P, B = _lib.dev(A_Rzphi).shape[0], _lib.dev(A_xy).shape[0]
    d = _lib.dev
    W, _ = solve_nnls_admm_bootstrap_lib(
        lib, P, B, d(y_Rzphi), d(gamma_kin), d(y_xy_boot),
        [d(y_h1_boot), d(y_h2_boot), d(y_h3_boot), d(y_h4_boot)], d(sig_Rzphi), d(sig_xy),
        [d(sig_A1), d(sig_A2), d(sig_A3), d(sig_A4)], lambda_reg, maxiter, w_rzphi, w_xy, w_h, tol)
    return W


def solve_nnls_admm(A_Rzphi, A_xy, A_h1, A_h2, A_h3, A_h4, y_Rzphi, y_xy, gamma_kin,
                    y_h1, y_h2, y_h3, y_h4, sig_Rzphi, sig_xy, sig_A1, sig_A2, sig_A3, sig_A4,
                    lambda_reg=1, maxiter=200, w_rzphi=1.0, w_xy=1.0, w_h=1.0, tol=0.0):
    """nnls.py:6-66: one right-hand side; returns weights (n,)."""
    d = _lib.dev
    W = solve_nnls_admm_bootstrap(A_Rzphi, A_xy, A_h1, A_h2, A_h3, A_h4, y_Rzphi, y_xy, gamma_kin,
                                  d(y_xy)[None], d(y_h1)[None], d(y_h2)[None], d(y_h3)[None],
                                  d(y_h4)[None], sig_Rzphi, sig_xy, sig_A1, sig_A2, sig_A3, sig_A4,
                                  lambda_reg, maxiter, w_rzphi, w_xy, w_h, tol)
    return W[0]
