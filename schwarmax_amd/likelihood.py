"""model_likelihood drop-in (likelihood.py:96-244) and the canonical model.

``density_func`` and ``potential_func`` are the canonical stellar density and
potential of likelihood.py:13-71 (MN disc + T3 bar + V4 bulge; NFW + MN + T3 +
V4), evaluated by the HIP kernels.  build_model accepts exactly these two.
"""
import numpy as np
import torch

from . import potentials as POT
from .model import build_model
from .utils import logMenc_logc_to_logM_logRs

EPSILON = 1e-12


def density_func(x, y, z, params):
    """Stellar density nu(x, y, z) on the GPU (likelihood.py:13-43)."""
    return POT.density_eval(x, y, z, params)


def potential_func(x, y, z, params_baryon, params_halo):
    """Phi(x, y, z) on the GPU (likelihood.py:46-71)."""
    return POT.potential_eval(x, y, z, params_baryon, params_halo)


def theta_to_params(params):
    """theta (12,) -> (params_halo_pot, params_baryon_rho) (likelihood.py:105-165)."""
    p = [float(v) for v in np.asarray(params, dtype=np.float64)]
    logM_halo, logRs_halo = logMenc_logc_to_logM_logRs(p[0], p[3], r_enc=10.0, Delta=200.,
                                                        rho_crit=277.54)
    L_bar = 10.0 ** p[6]
    Hs = 10.0 ** p[5]
    ph = {'logM': logM_halo, 'Rs': 10 ** logRs_halo, 'a': 1.0, 'b': 1.0, 'c': 1.0,
          'x_origin': 0.0, 'y_origin': 0.0, 'z_origin': 0.0, 'dirx': 0.0, 'diry': 0.0, 'dirz': 1.0}
    pb = {'logM_disc': p[1], 'Rs_disc': 10 ** p[4], 'Hs_disc': Hs, 'logM_bar': p[2],
          'L_bar': L_bar, 'a_bar': L_bar / 5.0, 'b_bar': Hs,
          'mass_to_light_ratio': 10 ** p[10], 'Omega_bar': 10 ** p[11],
          'x_origin': 0.0, 'y_origin': 0.0, 'z_origin': 0.0, 'dirx': 0.0, 'diry': 0.0, 'dirz': 1.0,
          'alpha': p[7] * 180 / np.pi, 'beta': p[8] * 180 / np.pi, 'gamma': p[9] * 180 / np.pi}
    return ph, pb


def grid_args(dict_data):
    """Grid arguments model_likelihood hands build_model (likelihood.py:167-182)."""
    Rzphi_lim = np.array([dict_data['R_minmax'], dict_data['z_minmax'], dict_data['phi_minmax']],
                         dtype=np.float64)
    xy_lim = np.array([np.asarray(dict_data['X_minmax']), np.asarray(dict_data['Y_minmax'])],
                      dtype=np.float64)
    return Rzphi_lim, np.asarray(dict_data['Rzphi_n_grid']), xy_lim, np.asarray(dict_data['nX_nY'])


def chi2_func(y_Rzphi, y_Rzphi_model, sigma_Rzphi,
              y_xy, y_xy_model, sigma_xy,
              y_h1, y_h1_model, sigma_h1,
              y_h2, y_h2_model, sigma_h2,
              y_h3, y_h3_model, sigma_h3,
              y_h4, y_h4_model, sigma_h4):
    """The jackknife's chi^2 (likelihood.py:73-94) on torch tensors; batched
    over a leading replicate axis like the reference's vmap.  Only the
    kinematic terms enter the returned value, as in the reference."""
    def res(y, ym, sg):
        return torch.nansum(((y - ym) / sg) ** 2, dim=-1)
    return -0.5 * (res(y_h1, y_h1_model, sigma_h1) + res(y_h2, y_h2_model, sigma_h2) +
                   res(y_h3, y_h3_model, sigma_h3) + res(y_h4, y_h4_model, sigma_h4))


def bootstrap_loglike(model_dict, dict_data, sigma_amplify=1.0):
    """Bootstrap chi^2 rows and their log-mean-exp (likelihood.py:184-244).

    Forward evaluations run smx_bootstrap_loglike; when the model maps carry
    an autograd graph (the gradient), the same formula runs as torch
    operations so that autograd can differentiate it."""
    if not (torch.is_grad_enabled() and any(
            isinstance(model_dict[k], torch.Tensor) and model_dict[k].requires_grad
            for k in ('density_3d', 'density_3d_model', 'surface_density', 'h1', 'h2', 'h3', 'h4'))):
        return _bootstrap_loglike_hip(model_dict, dict_data, sigma_amplify)
    return _bootstrap_loglike_graph(model_dict, dict_data, sigma_amplify)


def _bootstrap_loglike_hip(model_dict, dict_data, sigma_amplify):
    from . import _lib
    L = _lib.lib()
    f64 = torch.float64
    keep = []

    def dptr(t):
        t = _lib.dev(t)
        keep.append(t)
        return _lib.ptr(t, f64).value

    d3 = _lib.dev(model_dict['density_3d_model'])
    if d3.dim() != 2 or d3.stride(1) != 1:
        d3 = d3.contiguous()
    keep.append(d3)
    K, P = d3.shape
    d2 = _lib.dev(model_dict['surface_density'])
    B = d2.shape[1]
    model = _lib.SmxBootstrapModel(d3=d3.data_ptr(), ld3=d3.stride(0),
                                   y_R=dptr(model_dict['density_3d']), surface_density=dptr(d2))
    data = _lib.SmxBootstrapData(xy=dptr(dict_data['XY_density_data']),
                                 xy_err=dptr(dict_data['XY_density_data_err']),
                                 xy_normal=dptr(dict_data['XY_standard_normal']))
    for j in range(4):
        model.h[j] = dptr(model_dict['h%d' % (j + 1)])
        data.h[j] = dptr(dict_data['h%d_data' % (j + 1)])
        data.h_err[j] = dptr(dict_data['h%d_data_err' % (j + 1)])
        data.h_normal[j] = dptr(dict_data['h%d_standard_normal' % (j + 1)])
    work = torch.empty(2 * K, dtype=f64, device='cuda')
    rows = torch.empty(K, dtype=f64, device='cuda')
    out = torch.empty((), dtype=f64, device='cuda')
    _lib.check(L.smx_bootstrap_loglike(model, K, P, B, data, float(sigma_amplify), _lib.ptr(work, f64),
                                       _lib.ptr(rows, f64), _lib.ptr(out, f64), _lib.stream_ptr()),
               'smx_bootstrap_loglike')
    return out, rows


def _bootstrap_loglike_graph(model_dict, dict_data, sigma_amplify):
    from ._lib import dev
    y_R = model_dict['density_3d']
    y_xy = dev(dict_data['XY_density_data'])
    sig_xy_raw = dev(dict_data['XY_density_data_err']) + EPSILON
    y_xy_boot = y_xy[None, :] + dev(dict_data['XY_standard_normal']) * sig_xy_raw[None, :]
    sig_R = 0.1 * y_R * sigma_amplify
    sig_xy = 0.1 * y_xy * sigma_amplify
    res_R = torch.nansum(((model_dict['density_3d_model'] - y_R) / sig_R) ** 2)
    res_xy = torch.nansum(((model_dict['surface_density'] - y_xy_boot) / sig_xy) ** 2, dim=1)
    res_h = []
    for k in (1, 2, 3, 4):
        sraw = dev(dict_data['h%d_data_err' % k]) + EPSILON
        yb = dev(dict_data['h%d_data' % k])[None, :] + dev(dict_data['h%d_standard_normal' % k]) * sraw[None, :]
        res_h.append(torch.nansum(((model_dict['h%d' % k] - yb) / (sraw * sigma_amplify)) ** 2, dim=1))
    logl = -0.5 * (res_h[0] + res_h[1] + res_h[2] + res_h[3] + res_xy + res_R)
    lmax = torch.max(logl)
    return lmax + torch.log(torch.mean(torch.exp(logl - lmax))), logl


def model_likelihood(params, dict_data, num_Vbin, Rzphi_n_tot, sigma_amplify=1.0, schedule=None,
                     nnls_maxiter=200, nnls_tol=0.0):
    """likelihood.py:96-244 on the GPU; returns a 0-d tensor."""
    ph, pb = theta_to_params(params)
    Rzphi_lim, Rzphi_n, xy_lim, xy_n = grid_args(dict_data)
    md = build_model(density_func, potential_func, ph, pb, dict_data, num_Vbin, Rzphi_n_tot,
                     Rzphi_n, Rzphi_lim, xy_lim, xy_n, nnls_maxiter=nnls_maxiter, regularization=1.0,
                     schedule=schedule, nnls_tol=nnls_tol)
    return bootstrap_loglike(md, dict_data, sigma_amplify)[0]
