"""Gradient of model_likelihood with respect to theta: jax.grad of the
reference's likelihood.py:96-244 (through model.py:70-258) on the GPU.

The derivative is the one jax.grad takes: bin membership, accept / reject
decisions and validity flags are held fixed, the bootstrap NNLS weights are a
constant (nnls.py:66 and 165 wrap them in stop_gradient), and everything else
is differentiated: the Jeans initial conditions and their jitter scales, the
orbit integration including its step-size controller, the Gauss-Hermite
weights, the T normalisation, the realization average, the density target,
the data normalisation and the bootstrap chi^2.

Three torch.autograd.Function stages call the HIP kernels of libsmx.so:

* ``_InitialConditions``: forward ``smx_initial_conditions``; backward
  contracts the field Jacobians of ``smx_initial_conditions_jac``.
* ``_Library``: forward ``smx_build_library_ckpt`` (the library plus state
  checkpoints); backward ``smx_library_vjp``, the hand-written adjoint of the
  integrator and binning, replaying each checkpoint segment.
* ``_DensityTarget``: forward ``smx_density_target``; backward the field
  Jacobian of ``smx_density_target_jac``.

The map theta -> potential fields, the model maps (W @ library, W constant)
and the chi^2 are torch operations, so autograd joins the pieces.
"""
import math

import numpy as np
import torch

from . import _lib
from . import potentials as POT
from . import model as MD
from .ghmoments import h_to_V_sigma
from .likelihood import theta_to_params, grid_args, bootstrap_loglike
from .nnls import solve_nnls_admm_bootstrap_lib

EPSILON = 1e-12
G = POT.G
N_FIELDS = 11
FIELD_NAMES = ('nfw_mGM', 'nfw_Rs', 'mn_mGM', 'mn_rho0', 'mn_Rs', 'mn_Hs', 't3_M', 't3_L', 't3_a',
               't3_b', 'v4_M')


# ---------------------------------------------------------------------------
# theta -> potential fields, differentiable (likelihood.py:105-165)
# ---------------------------------------------------------------------------
def _logMenc_logc_to_logM_logRs(logM_enc, log_c, r_enc=10.0, Delta=200., rho_crit=277.54):
    """utils.py:526-569 in torch: six Newton steps, differentiated through."""
    M_enc = 10.0 ** logM_enc
    c = 10.0 ** log_c
    coeff = 4.0 * math.pi * Delta * rho_crit / 3.0
    ln10 = math.log(10.0)
    logM = logM_enc + 0.5
    for _ in range(6):
        M = 10.0 ** logM
        Rs = (M / coeff) ** (1.0 / 3.0) / c
        x = r_enc / Rs
        f_nfw = torch.log(1.0 + x) - x / (1.0 + x)
        f = M * f_nfw - M_enc
        df = (f_nfw - x * x / (3.0 * (1.0 + x) ** 2)) * M * ln10
        logM = logM - f / df
    M_sol = 10.0 ** logM
    return logM, torch.log10((M_sol / coeff) ** (1.0 / 3.0) / c)


def _rotation_matrix(alpha_deg, beta_deg, gamma_deg):
    """utils.py:164-179 in torch: (Rz(gamma) Rx(beta) Rz(alpha))^T, degrees."""
    a, b, g = (t * (math.pi / 180.0) for t in (alpha_deg, beta_deg, gamma_deg))
    one, zero = torch.ones_like(a), torch.zeros_like(a)

    def Rz(t):
        c, s = torch.cos(t), torch.sin(t)
        return torch.stack([torch.stack([c, -s, zero]), torch.stack([s, c, zero]),
                            torch.stack([zero, zero, one])])

    def Rx(t):
        c, s = torch.cos(t), torch.sin(t)
        return torch.stack([torch.stack([one, zero, zero]), torch.stack([zero, c, -s]),
                            torch.stack([zero, s, c])])
    return (Rz(g) @ Rx(b) @ Rz(a)).T


def theta_fields(theta):
    """theta (12,) float64 tensor -> (fields (11,), Omega_bar, rot (3, 3), lm).

    The fields are the smx_potential entries pack_potential fills from the
    parameter dicts theta_to_params builds, formed with the same operations.
    """
    logM_h, logRs_h = _logMenc_logc_to_logM_logRs(theta[0], theta[3])
    M_disc = 10.0 ** theta[1]
    Hs = 10.0 ** theta[5]
    M_bar = 10.0 ** theta[2]
    L_bar = 10.0 ** theta[6]
    fields = torch.stack([
        -G * 10.0 ** logM_h, 10.0 ** logRs_h,
        -G * M_disc, Hs ** 2 * M_disc / (4 * math.pi), 10.0 ** theta[4], Hs,
        M_bar, L_bar, L_bar / 5.0, Hs,
        M_bar])
    deg = 180.0 / math.pi
    rot = _rotation_matrix(theta[7] * deg, theta[8] * deg, theta[9] * deg)
    lm = 1.0 / 10.0 ** theta[10]
    Omega_bar = 10.0 ** theta[11]
    return fields, Omega_bar, rot, lm


# ---------------------------------------------------------------------------
# autograd stages
# ---------------------------------------------------------------------------
class _InitialConditions(torch.autograd.Function):
    @staticmethod
    def forward(ctx, fields, pot, dict_data, r, sched):
        wb, Tt, dt0 = MD.initial_conditions(pot, dict_data, r, sched)
        ctx.pot, ctx.dd, ctx.r, ctx.sched = pot, dict_data, r, sched
        ctx.field_device = fields.device
        ctx.mark_non_differentiable(Tt)
        return wb, Tt, dt0

    @staticmethod
    def backward(ctx, wb_bar, Tt_bar, dt0_bar):
        L = _lib.lib()
        dd, r = ctx.dd, ctx.r
        w0 = _lib.dev(np.asarray(dd['w0'])[:, :3])
        n = w0.shape[0]
        normals, uni = (_lib.dev(a) for a in MD.ic_draws(dd, r))
        jw = torch.empty((n, r, 6, N_FIELDS), dtype=torch.float64, device='cuda')
        jd = torch.empty((n, r, N_FIELDS), dtype=torch.float64, device='cuda')
        _lib.check(L.smx_initial_conditions_jac(ctx.pot, _lib.ptr(w0), _lib.ptr(normals),
                                                _lib.ptr(uni), n, r, uni.shape[1],
                                                float(ctx.sched['N_step_per_orb']), _lib.ptr(jw),
                                                _lib.ptr(jd), _lib.stream_ptr()),
                   'smx_initial_conditions_jac')
        g = torch.zeros(N_FIELDS, dtype=torch.float64, device='cuda')
        if wb_bar is not None:
            g = g + torch.einsum('nrj,nrjf->f', wb_bar, jw)
        if dt0_bar is not None:
            g = g + torch.einsum('nr,nrf->f', dt0_bar, jd)
        return g.to(ctx.field_device), None, None, None, None


class _Library(torch.autograd.Function):
    @staticmethod
    def forward(ctx, fields, omega_ode, rot_row2, wb, dt0, Tt, pot, grids, sched_struct, dict_data,
                ckpt_every):
        L = _lib.lib()
        n, r = wb.shape[0], wb.shape[1]
        n_int = n * r
        P = grids.nR * grids.nz * grids.nphi
        B = grids.num_vbin
        ws = torch.empty(int(L.smx_library_ckpt_bytes(grids, sched_struct, n_int)),
                         dtype=torch.uint8, device='cuda')
        lib = torch.empty((n, P + 6 * B), dtype=torch.float64, device='cuda')
        valid = torch.empty(n_int, dtype=torch.float64, device='cuda')
        T_int = torch.empty(n_int, dtype=torch.float64, device='cuda')
        t_final = torch.empty(n_int, dtype=torch.float64, device='cuda')
        nck = (sched_struct.N_max + ckpt_every - 1) // ckpt_every
        ckpt = torch.empty((max(nck, 1), n_int, 8), dtype=torch.float64, device='cuda')
        bm = _lib.bin_map(dict_data['bin_mapping'], grids.nX, grids.nY)
        npb = _lib.dev(dict_data['num_per_bin'])
        v0 = _lib.velocities(dict_data['v0'])
        sd = _lib.dispersions(dict_data['s'])
        wb = wb.contiguous()
        dt0 = dt0.contiguous()
        _lib.check(L.smx_build_library_ckpt(pot, grids, sched_struct, _lib.ptr(wb), _lib.ptr(Tt),
                                            _lib.ptr(dt0), n, r, _lib.ptr(bm), _lib.ptr(npb),
                                            _lib.ptr(v0), _lib.ptr(sd), _lib.ptr(ws), ws.numel(),
                                            _lib.ptr(lib),
                                            _lib.ptr(valid), _lib.ptr(T_int), _lib.ptr(t_final),
                                            _lib.ptr(ckpt), int(ckpt_every), _lib.stream_ptr()),
                   'smx_build_library_ckpt')
        ctx.keep = (pot, grids, sched_struct, bm, npb, v0, sd, ws, valid, T_int, ckpt, int(ckpt_every),
                    wb, dt0)
        ctx.field_device = fields.device
        ctx.om_device = omega_ode.device
        ctx.rot_device = rot_row2.device
        ctx.valid = valid
        return lib

    @staticmethod
    def backward(ctx, lib_bar):
        L = _lib.lib()
        pot, grids, sched_struct, bm, npb, v0, sd, ws, valid, T_int, ckpt, every, wb, dt0 = ctx.keep
        n, r = wb.shape[0], wb.shape[1]
        n_int = n * r
        lib_bar = lib_bar.contiguous()
        vws = torch.empty(int(L.smx_library_vjp_workspace_bytes(n_int, every)), dtype=torch.uint8,
                          device='cuda')
        wb_bar = torch.empty((n, r, 6), dtype=torch.float64, device='cuda')
        dt0_bar = torch.empty((n, r), dtype=torch.float64, device='cuda')
        f_bar = torch.empty((n_int, N_FIELDS), dtype=torch.float64, device='cuda')
        om_bar = torch.empty(n_int, dtype=torch.float64, device='cuda')
        rot_bar = torch.empty((n_int, 3), dtype=torch.float64, device='cuda')
        _lib.check(L.smx_library_vjp(pot, grids, sched_struct, _lib.ptr(wb), _lib.ptr(dt0), n, r,
                                     _lib.ptr(bm), _lib.ptr(npb), _lib.ptr(v0), _lib.ptr(sd),
                                     _lib.ptr(ws), _lib.ptr(valid), _lib.ptr(T_int), _lib.ptr(ckpt),
                                     every, _lib.ptr(lib_bar), _lib.ptr(vws), _lib.ptr(wb_bar),
                                     _lib.ptr(dt0_bar), _lib.ptr(f_bar), _lib.ptr(om_bar),
                                     _lib.ptr(rot_bar), _lib.stream_ptr()), 'smx_library_vjp')
        return (f_bar.sum(0).to(ctx.field_device), om_bar.sum().to(ctx.om_device),
                rot_bar.sum(0).to(ctx.rot_device), wb_bar, dt0_bar,
                None, None, None, None, None, None)


class _DensityTarget(torch.autograd.Function):
    @staticmethod
    def forward(ctx, fields, pot, dict_data):
        L = _lib.lib()
        Rg = _lib.dev(dict_data['R_grid'])
        zg = _lib.dev(dict_data['z_grid'])
        pg = _lib.dev(dict_data['phi_grid'])
        smp = _lib.dev(dict_data['sample_for_integration'])
        P = Rg.numel()
        out = torch.empty(P, dtype=torch.float64, device='cuda')
        args = (_lib.ptr(Rg), _lib.ptr(zg), _lib.ptr(pg), P, float(dict_data['dR']),
                float(dict_data['dz']), float(dict_data['dphi']), _lib.ptr(smp), smp.shape[0])
        _lib.check(L.smx_density_target(pot, *args, _lib.ptr(out), _lib.stream_ptr()),
                   'smx_density_target')
        ctx.keep = (pot, args, Rg, zg, pg, smp)
        ctx.field_device = fields.device
        return out

    @staticmethod
    def backward(ctx, y_bar):
        L = _lib.lib()
        pot, args, *_ = ctx.keep
        P = args[3]
        jac = torch.empty((P, N_FIELDS), dtype=torch.float64, device='cuda')
        _lib.check(L.smx_density_target_jac(pot, *args, _lib.ptr(jac), _lib.stream_ptr()),
                   'smx_density_target_jac')
        return (y_bar @ jac).to(ctx.field_device), None, None


# ---------------------------------------------------------------------------
# forward with the autograd graph, and the gradient
# ---------------------------------------------------------------------------
def loglike_graph(theta, dict_data, num_Vbin, Rzphi_n_tot, sigma_amplify=1.0, schedule=None,
                  nnls_maxiter=200, nnls_tol=0.0, ckpt_every=64, W=None):
    """logL(theta) as a torch graph whose backward runs the adjoint kernels.

    theta: float64 CPU tensor (12,), usually with requires_grad.  Returns
    (logL, aux) with aux['W'] the NNLS weights used (or the ``W`` passed in).
    """
    ph, pb = theta_to_params(theta.detach().cpu().numpy())
    sched = MD.schedule_of(schedule)
    r = int(sched['n_realizations'])
    Rzphi_lim, Rzphi_n, xy_lim, xy_n = grid_args(dict_data)
    pot = POT.pack_potential(MD.baryon_params(pb), ph)
    rot_np = MD.makeRotationMatrix(pb['alpha'], pb['beta'], pb['gamma'])
    grids = MD.make_grids(num_Vbin, Rzphi_n, Rzphi_lim, xy_lim, xy_n, rot_np, symmetry=True)
    if grids.nR * grids.nz * grids.nphi != int(Rzphi_n_tot):
        raise ValueError("Rzphi_n_tot must equal prod(Rzphi_n_grid)")
    sched_struct = MD.make_schedule(sched, -float(pb['Omega_bar']))

    fields, Omega_bar, rot, lm = theta_fields(theta)
    wb, Tt, dt0 = _InitialConditions.apply(fields, pot, dict_data, r, sched)
    lib = _Library.apply(fields, -Omega_bar, rot[2], wb, dt0, Tt, pot, grids, sched_struct,
                         dict_data, ckpt_every)
    P, B = grids.nR * grids.nz * grids.nphi, grids.num_vbin
    n = lib.shape[0]
    y_R = _DensityTarget.apply(fields, POT.pack_potential(pb), dict_data)

    # model.py:215-258 (normalisation, gamma_kin, bootstrap data)
    dev = _lib.dev
    lm = lm.to('cuda')
    sig_R = 0.02 * y_R + 1e-10
    y_xy = dev(dict_data['XY_density_data']) / lm
    sig_xy = (dev(dict_data['XY_density_data_err']) + EPSILON) / lm
    y_h = [dev(dict_data['h%d_data' % k]) for k in (1, 2, 3, 4)]
    sig_h = [dev(dict_data['h%d_data_err' % k]) + EPSILON for k in (1, 2, 3, 4)]
    mmpo = torch.sum(y_R) / n
    y_xy = y_xy / mmpo
    sig_xy = sig_xy / mmpo
    y_R = y_R / mmpo
    sig_R = sig_R / mmpo
    Xm = np.asarray(dict_data['X_minmax'], dtype=np.float64)
    Ym = np.asarray(dict_data['Y_minmax'], dtype=np.float64)
    nXY = np.asarray(dict_data['nX_nY'])
    area = ((Xm[1] - Xm[0]) / nXY[0]) * ((Ym[1] - Ym[0]) / nXY[1]) * 1e6
    npb = dev(dict_data['num_per_bin'])
    s_ap, v0_ap = dev(dict_data['s']), dev(dict_data['v0'])
    M_per_bin = y_xy * npb * area
    norm_GH = 1.0 + y_h[3] * np.sqrt(6.0) / 4.0
    gamma_kin = M_per_bin / (norm_GH * 2.0 * np.sqrt(np.pi) * s_ap + EPSILON)
    y_xy_boot = y_xy[None, :] + dev(dict_data['XY_standard_normal']) * sig_xy[None, :]
    y_h_boot = [y_h[k][None, :] + dev(dict_data['h%d_standard_normal' % (k + 1)]) * sig_h[k][None, :]
                for k in range(4)]
    if W is None:
        with torch.no_grad():
            W, _ = solve_nnls_admm_bootstrap_lib(
                lib.detach(), P, B, y_R.detach(), gamma_kin.detach(), y_xy_boot.detach(),
                [t.detach() for t in y_h_boot], sig_R.detach(), sig_xy.detach(), sig_h,
                lambda_reg=1.0, maxiter=nnls_maxiter, w_rzphi=np.sqrt(5.0), w_xy=np.sqrt(5.0),
                w_h=1.0, tol=nnls_tol)
    W = W.detach()
    # model.py:14-66 with W constant
    m = P + 5 * B
    maps = W @ lib[:, :m]
    g = torch.where(gamma_kin.abs() > 1e-8, gamma_kin, torch.full_like(gamma_kin, 1e-8))
    d3 = maps[:, :P]
    d2 = maps[:, P:P + B] * mmpo * lm
    h = [torch.clamp(maps[:, P + B * (k + 1):P + B * (k + 2)] / g, -10.0, 10.0) for k in range(4)]
    V, sig = h_to_V_sigma(h[0], h[1], v0_ap, s_ap, h3=h[2], h4=h[3])
    md = {'density_3d': y_R, 'density_3d_model': d3, 'surface_density': d2, 'h1': h[0],
          'h2': h[1], 'h3': h[2], 'h4': h[3], 'V': V, 'sigma': sig, 'weights': W}
    logl, _ = bootstrap_loglike(md, dict_data, sigma_amplify)
    return logl, {'W': W, 'lib': lib.detach(), 'valid': None}


def model_likelihood_and_grad(params, dict_data, num_Vbin, Rzphi_n_tot, sigma_amplify=1.0,
                              schedule=None, nnls_maxiter=200, nnls_tol=0.0, ckpt_every=64):
    """(logL, dlogL/dtheta): model_likelihood and its jax.grad (likelihood.py:96-244).

    Returns a 0-d CUDA tensor and a float64 CPU tensor of shape (12,).
    """
    theta = torch.as_tensor(np.asarray(params, dtype=np.float64)).clone().requires_grad_(True)
    logl, _ = loglike_graph(theta, dict_data, num_Vbin, Rzphi_n_tot, sigma_amplify, schedule,
                            nnls_maxiter, nnls_tol, ckpt_every)
    logl.backward()
    return logl.detach(), theta.grad.detach().clone()


def model_likelihood_and_grad_batch(thetas, dict_data, num_Vbin, Rzphi_n_tot, n_orbits=None,
                                    sigma_amplify=1.0, schedule=None, nnls_maxiter=200,
                                    nnls_tol=0.0, ckpt_every=64, autotune=True):
    """(logL (M,), dlogL/dtheta (M, 12)) for the rows of thetas (M, 12).

    Model m uses the first n_orbits[m] orbits of dict_data (default: all).
    The models run one after another through model_likelihood_and_grad, each
    freeing its slabs, checkpoints and NNLS workspace before the next (a
    40,000-orbit model holds about 110 GB at its peak).
    autotune: pick the NNLS stages' launch-shape knobs for the batch's median
    shape first (``tune.ensure_tuned``, once per process and shape bucket).
    """
    thetas = np.asarray(thetas, dtype=np.float64).reshape(-1, 12)
    M = thetas.shape[0]
    n_all = np.asarray(dict_data['w0']).shape[0]
    n_orbits = [n_all] * M if n_orbits is None else [int(v) for v in n_orbits]
    if len(n_orbits) != M or min(n_orbits) < 1 or max(n_orbits) > n_all:
        raise ValueError("n_orbits must hold M counts in [1, number of orbits in dict_data]")
    if autotune:
        from . import tune
        K = np.asarray(dict_data['XY_standard_normal']).shape[0]
        tune.ensure_tuned(int(np.median(n_orbits)), int(Rzphi_n_tot) + 5 * int(num_Vbin), K)
    logls = torch.empty(M, dtype=torch.float64, device='cuda')
    grads = torch.empty((M, 12), dtype=torch.float64)
    for m in range(M):
        dm = dict(dict_data)
        if n_orbits[m] < n_all:
            for k in ('w0', 'jeans_normals', 'jitter_uniforms'):
                if k in dm:
                    dm[k] = np.asarray(dict_data[k])[:n_orbits[m]]
        v, g = model_likelihood_and_grad(thetas[m], dm, num_Vbin, Rzphi_n_tot, sigma_amplify,
                                         schedule, nnls_maxiter, nnls_tol, ckpt_every)
        logls[m] = v
        grads[m] = g
        torch.cuda.empty_cache()
    return logls, grads
