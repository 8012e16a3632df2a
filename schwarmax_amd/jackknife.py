"""Delete-d jackknife delta-chi^2 (model.py:290-603, likelihood.py:248-335) on the GPU.

The reference re-solves the single-right-hand-side NNLS once per replicate:
replicate g raises the sigma of the apertures b with b % n_groups == g to
1e20, which zeroes their rows of U (to 1e-20 relative), and rebuilds Gram,
rho and Cholesky.  Written with the shared design matrix U of the unmasked
problem and V_g (the d = 5 |g| rows of U that replicate g drops):

    Q_g + rho_g I = (Q + rho I) - delta_g I - V_g^T V_g,   delta_g = |V_g|_F^2 / n,
    c_g = c + V_g^T y_g,

so every replicate's solve is one shared inverse M = (Q + rho I)^-1 with
* a diagonal shift, by the Neumann series (A - delta I)^-1 = sum_k delta^k M^(k+1)
  (delta |M| <= delta / rho = |V_g|^2 / tr Q, about d / m, so a handful of
  terms reach 1e-13), and
* a rank-d removal, by Woodbury: (B - V^T V)^-1 = B^-1 + B^-1 V^T C^-1 V B^-1,
  C = I - V B^-1 V^T (d x d).
All n_groups replicates then run as n_groups right-hand sides of the ADMM
product kernel (a few products per iteration instead of one Cholesky solve
per replicate), with the reference's iterates: alpha = 1.6, rho_g, z0 =
sum(y_Rzphi)/n, maxiter iterations, result z.
"""
import math

import numpy as np
import torch

from . import _lib
from . import model as MD
from . import potentials as POT
from .likelihood import theta_to_params, grid_args, density_func, potential_func, chi2_func
from .nnls import gamma_safe

BIG = 1e20
EPS = 1e-8


class _Operator:
    """M = (U^T U + (lambda/n) I + rho I)^-1 in libsmx.so's workspace."""

    def __init__(self, lib, m, row_scale, lambda_reg, K_max):
        L = _lib.lib()
        self.n, self.m, self.K_max = lib.shape[0], m, int(K_max)
        self.ws = torch.empty(int(L.smx_nnls_operator_bytes(self.n, m, self.K_max)),
                              dtype=torch.uint8, device='cuda')
        rho = torch.empty(1, dtype=torch.float64, device='cuda')
        _lib.check(L.smx_nnls_operator(_lib.ptr(lib), lib.shape[1], self.n, m,
                                       _lib.ptr(row_scale.contiguous()), float(lambda_reg),
                                       self.K_max, _lib.ptr(self.ws), _lib.ptr(rho),
                                       _lib.stream_ptr()), 'smx_nnls_operator')
        self.rho = float(rho.item())

    def apply_into(self, R32, out, hi_only=False):
        """out (K, n) <- R32 M, both FP32 and contiguous, K <= K_max per call;
        hi_only: R_hi M_hi (the delta product)."""
        L = _lib.lib()
        fn = L.smx_nnls_operator_apply_hi if hi_only else L.smx_nnls_operator_apply
        for k0 in range(0, R32.shape[0], self.K_max):
            k1 = min(k0 + self.K_max, R32.shape[0])
            _lib.check(fn(self.n, self.m, self.K_max, _lib.ptr(self.ws),
                          _lib.ptr(R32[k0:k1], torch.float32), k1 - k0,
                          _lib.ptr(out[k0:k1], torch.float32), _lib.stream_ptr()),
                       'smx_nnls_operator_apply')
        return out

    def apply(self, R, hi_only=False):
        """R (K, n) -> R M (FP32 in, FP64 out)."""
        R = R.to(torch.float32).contiguous()
        return self.apply_into(R, torch.empty_like(R), hi_only).double()


class _DeltaApply:
    """Products R_t M of a slowly changing input sequence R_t (one per ADMM
    iteration), into a fixed FP32 output: a full product when the input has
    moved by more than TAU of some row's size since the last full one,
    otherwise that product plus the delta product (R_t - R_b)_hi M_hi, as the
    solver's own ADMM loop does (one host check per product)."""

    TAU = 2.0 ** -10

    def __init__(self, op, out):
        self.op = op
        self.out = out          # (K, n) FP32
        self.base_R = None      # the input of the last full product, and that product
        self.base_Y = None
        self.full = 0
        self.calls = 0

    def __call__(self, R32, force_full=False):
        """self.out <- R32 M (R32: FP32, contiguous, left unchanged)."""
        self.calls += 1
        if self.base_R is not None and not force_full:
            D = R32 - self.base_R
            dmax = torch.linalg.vector_norm(D, ord=float('inf'), dim=1)
            rmax = torch.linalg.vector_norm(R32, ord=float('inf'), dim=1)
            if not bool((dmax > self.TAU * rmax).any()):
                self.op.apply_into(D, self.out, hi_only=True)
                self.out += self.base_Y
                return self.out
        self.op.apply_into(R32, self.out)
        if self.base_R is None:
            self.base_R, self.base_Y = R32.clone(), self.out.clone()
        else:
            self.base_R.copy_(R32)
            self.base_Y.copy_(self.out)
        self.full += 1
        return self.out


def _neumann_terms(q, tol=1e-13, cap=40):
    if not q > 0.0:
        return 1
    if q >= 0.5:
        raise NotImplementedError(
            "a jackknife group holds %.2f of the Gram's trace; the shared-operator solve "
            "needs < 0.5" % q)
    return min(cap, max(1, int(math.ceil(math.log(tol) / math.log(q)))))


def model_deltaChi2_jackknife(density_func_, potential_func_, chi2_func_, params_halo_pot,
                              params_disk_rho, dict_data, num_Vbin, Rzphi_n_tot=360,
                              Rzphi_n_grid=(10, 6, 6),
                              Rzphi_lim_grid=((0, 10.), (-3, 3), (-np.pi, np.pi)),
                              xy_lim_grid=((-10., 10.), (-3., 3.)), xy_n_grid=(60, 40),
                              nnls_maxiter=200, regularization=1.0, n_groups=100, batch_size=50,
                              schedule=None, return_aux=False):
    """model.py:290-603 on the GPU; returns std over replicates of chi2_func.

    chi2_func_ takes the reference's 18 arguments as torch tensors with a
    leading replicate axis and returns one value per replicate
    (schwarmax_amd.likelihood.chi2_func restates likelihood.py:73-94).
    batch_size only bounds the reference's memory; here all replicates run
    together.
    """
    MD._canonical(density_func_, potential_func_)
    sched = MD.schedule_of(schedule)
    pb = MD.baryon_params(params_disk_rho)
    pot = POT.pack_potential(pb, params_halo_pot)
    rot = MD.makeRotationMatrix(float(params_disk_rho['alpha']), float(params_disk_rho['beta']),
                                float(params_disk_rho['gamma']))
    grids = MD.make_grids(num_Vbin, Rzphi_n_grid, Rzphi_lim_grid, xy_lim_grid, xy_n_grid, rot,
                          symmetry=True)
    if grids.nR * grids.nz * grids.nphi != int(Rzphi_n_tot):
        raise ValueError("Rzphi_n_tot must equal prod(Rzphi_n_grid)")
    wb, Tt, dt0 = MD.initial_conditions(pot, dict_data, int(sched['n_realizations']), sched)
    libd = MD.build_library(pot, grids, MD.make_schedule(sched, -float(params_disk_rho['Omega_bar'])),
                            wb, Tt, dt0, dict_data)
    del wb, Tt, dt0
    lib, P, B = libd['lib'], libd['P'], libd['B']
    n = lib.shape[0]
    m = P + 5 * B
    t = MD.prepare_targets(n, params_disk_rho, dict_data)
    y_R, sig_R, y_xy, sig_xy = t['y_R'], t['sig_R'], t['y_xy'], t['sig_xy']
    y_h, sig_h, gk = t['y_h'], t['sig_h'], t['gamma_kin']
    lum = t['mmpo'] * t['lm']

    # the unmasked single-right-hand-side system (nnls.py:15-48), channel weights sqrt5, sqrt5, 1
    w_rz, w_xy, w_h = math.sqrt(5.0), math.sqrt(5.0), 1.0
    g = gamma_safe(gk)
    row_scale = torch.cat([w_rz / (sig_R + EPS), w_xy / (sig_xy + EPS)] +
                          [w_h / (g * (sk + EPS)) for sk in sig_h])
    yv = torch.cat([w_rz * y_R / (sig_R + EPS), w_xy * y_xy / (sig_xy + EPS)] +
                   [w_h * yk / (sk + EPS) for yk, sk in zip(y_h, sig_h)])
    c = -(lib[:, :m] @ (row_scale * yv))                        # -U^T y

    # the rows each replicate drops: apertures a with a % n_groups == g, 5 rows each
    G = int(n_groups)
    aps = [np.nonzero(np.arange(B) % G == gi)[0] for gi in range(G)]
    d = 5 * max(len(a) for a in aps)
    rows = np.full((G, d), -1, dtype=np.int64)
    for gi, a in enumerate(aps):
        r = np.concatenate([P + k * B + a for k in range(5)])
        rows[gi, :r.size] = r
    rows_t = torch.as_tensor(rows, device='cuda')
    valid = rows_t >= 0
    ridx = rows_t.clamp(min=0)
    V = (lib[:, ridx.reshape(-1)].T.reshape(G, d, n) * row_scale[ridx].unsqueeze(-1)
         * valid.unsqueeze(-1))                                 # (G, d, n) rows of U
    yV = yv[ridx] * valid
    delta = (V ** 2).sum(dim=(1, 2)) / n                        # tr(V^T V) / n
    c_g = c[None, :] + torch.einsum('gdn,gd->gn', V, yV)

    op = _Operator(lib, m, row_scale, regularization, K_max=max(G, G * d))
    rho = op.rho
    rho_g = rho - delta
    T = _neumann_terms(float(delta.max()) / rho)

    def shifted_inverse(R, dl):
        """sum_k dl^k M^(k+1) R, rows scaled by their own shift."""
        s = op.apply(R)
        acc = s.clone()
        for k in range(1, T):
            s = op.apply(s)
            acc += (dl ** k)[:, None] * s
        return acc

    Pg = shifted_inverse(V.reshape(G * d, n), delta.repeat_interleave(d)).reshape(G, d, n)
    C = torch.eye(d, dtype=torch.float64, device='cuda') - torch.einsum('gdn,gen->gde', V, Pg)
    Cinv = torch.linalg.inv(C)

    # ADMM (nnls.py:51-66), all replicates at once; each Neumann term's input
    # changes slowly from one iteration to the next, so its products are
    # delta products between full ones (the last iteration's are full)
    z = torch.full((G, n), float(torch.sum(y_R)) / n, dtype=torch.float64, device='cuda')
    u = torch.zeros_like(z)
    alpha = 1.6
    # the T Neumann terms' products of one iteration, S[k] = r M^(k+1) (FP32);
    # term k + 1 multiplies term k's output
    S = torch.empty((T, G, n), dtype=torch.float32, device='cuda')
    terms = [_DeltaApply(op, S[k]) for k in range(T)]

    # x = sum_k delta^k S[k] (FP64), the Woodbury correction and the z / u
    # update of each iteration, with the next right-hand sides (FP32, the
    # products' input), in one kernel (one workgroup per replicate)
    L = _lib.lib()
    V = V.contiguous()
    Pg = Pg.contiguous()
    Cinv = Cinv.contiguous()
    rho_g = rho_g.contiguous()
    c_g = c_g.contiguous()
    delta = delta.contiguous()
    r32 = (rho_g[:, None] * (z - u) - c_g).to(torch.float32).contiguous()
    for it in range(int(nnls_maxiter)):
        last = it == int(nnls_maxiter) - 1          # the last iteration's products are full
        R = r32
        for tm in terms:
            R = tm(R, force_full=last)
        _lib.check(L.smx_jackknife_admm_step(G, n, d, T, _lib.ptr(S, torch.float32),
                                             _lib.ptr(delta, torch.float64),
                                             _lib.ptr(V, torch.float64), _lib.ptr(Pg, torch.float64),
                                             _lib.ptr(Cinv, torch.float64),
                                             _lib.ptr(rho_g, torch.float64),
                                             _lib.ptr(c_g, torch.float64), alpha,
                                             _lib.ptr(z, torch.float64), _lib.ptr(u, torch.float64),
                                             _lib.ptr(r32, torch.float32), _lib.stream_ptr()),
                   'smx_jackknife_admm_step')

    # per-replicate model and chi^2 (model.py:531-577)
    d3, d2, h1, h2, h3, h4, _, _ = MD.compute_model_single(z, lib, P, B, gk, t['v0_ap'], t['s_ap'])
    d2 = d2 * lum
    dropped = (torch.arange(B, device='cuda')[None, :] % G) == torch.arange(G, device='cuda')[:, None]
    sxy = torch.where(dropped, torch.full_like(d2, BIG), sig_xy[None, :].expand(G, -1))
    sh = [torch.where(dropped, torch.full_like(d2, BIG), sk[None, :].expand(G, -1)) for sk in sig_h]
    e = lambda v: v[None, :].expand(G, -1)  # noqa: E731
    chi2 = chi2_func_(e(y_R), d3, e(sig_R), e(y_xy * lum), d2, sxy * lum,
                      e(y_h[0]), h1, sh[0], e(y_h[1]), h2, sh[1], e(y_h[2]), h3, sh[2],
                      e(y_h[3]), h4, sh[3])
    out = torch.std(chi2, unbiased=False)
    if return_aux:
        return out, dict(chi2=chi2, weights=z, neumann_terms=T, rho=rho, delta=delta,
                         products=sum(tm.calls for tm in terms),
                         full_products=sum(tm.full for tm in terms))
    return out


def calculate_delta_chi2(params, dict_data, num_Vbin, Rzphi_n_tot, schedule=None,
                         nnls_maxiter=200, return_aux=False):
    """likelihood.py:248-335: the jackknife with one group per aperture."""
    ph, pb = theta_to_params(params)
    Rzphi_lim, Rzphi_n, xy_lim, xy_n = grid_args(dict_data)
    return model_deltaChi2_jackknife(density_func, potential_func, chi2_func, ph, pb, dict_data,
                                     num_Vbin, Rzphi_n_tot, Rzphi_n, Rzphi_lim, xy_lim, xy_n,
                                     nnls_maxiter=nnls_maxiter, regularization=1.0,
                                     n_groups=num_Vbin, batch_size=25, schedule=schedule,
                                     return_aux=return_aux)
