"""Forward model and likelihood, FP64 CPU restatement.

Restates model.py:14-68 (compute_model_single), model.py:70-286
(build_model), model.py:606-850 (get_model_with_orbit), likelihood.py:96-244
(model_likelihood) and utils.py:526-569 (NFW parameter map).

The reference draws its Jeans normals and jitter uniforms from fixed JAX PRNG
keys (model.py:112, 128-138).  Those streams cannot be reproduced without JAX,
so they are inputs here: ``dict_data['jeans_normals']`` (n, 3) and
``dict_data['jitter_uniforms']`` (n, r, 6), exactly as the device path takes
them.  The integration schedule constants that build_model hard-codes
(model.py:163-169) are the defaults of :data:`DEFAULT_SCHEDULE`.
"""
import numpy as np

from .constants import EPSILON
from . import potential as P
from . import integrate as I
from . import jeans as J
from . import nnls as NN

DEFAULT_SCHEDULE = dict(N_step_per_orb=100, N_dynamical_time=50, N_max=None,
                        atol=1e-7, rtol=1e-4, dt_min=1e-5, dt_max=0.3, chunk=100,
                        n_realizations=4)


def schedule_of(sched):
    s = dict(DEFAULT_SCHEDULE)
    if sched:
        s.update(sched)
    if s['N_max'] is None:
        s['N_max'] = s['N_step_per_orb'] * s['N_dynamical_time']
    return s


def make_rotation_matrix(alpha, beta, gamma):
    """(Rz(g) Rx(b) Rz(a))^T with angles in degrees (utils.py:164-179)."""
    a, b, g = np.radians(alpha), np.radians(beta), np.radians(gamma)

    def Rz(t):
        return np.array([[np.cos(t), -np.sin(t), 0.0], [np.sin(t), np.cos(t), 0.0], [0.0, 0.0, 1.0]])

    def Rx(t):
        return np.array([[1.0, 0.0, 0.0], [0.0, np.cos(t), -np.sin(t)], [0.0, np.sin(t), np.cos(t)]])
    return (Rz(g) @ Rx(b) @ Rz(a)).T


def baryon_params(params_disk_rho):
    """params_baryon of model.py:87-102."""
    keys = ('logM_disc', 'Rs_disc', 'Hs_disc', 'logM_bar', 'L_bar', 'a_bar', 'b_bar',
            'x_origin', 'y_origin', 'z_origin', 'dirx', 'diry', 'dirz')
    return {k: params_disk_rho[k] for k in keys}


def initial_conditions(dict_data, pb, ph, r, N_step, N_dyn):
    """Jeans ICs, rotation curve, jitter and time scales (model.py:112-169).

    Returns w0 batch (n, r, 6), T_total (n, r), dt_init (n, r).
    """
    w0 = np.asarray(dict_data['w0'], dtype=np.float64)
    w0n = J.get_w0_new(w0, np.asarray(dict_data['jeans_normals'], dtype=np.float64), pb, ph,
                       P.potential_func, P.density_func)
    R = np.sqrt(w0n[:, 0] ** 2 + w0n[:, 1] ** 2)
    z = w0n[:, 2]
    Vc = J.rotation_curve(R, z, P.potential_func, pb, ph)
    U = np.asarray(dict_data['jitter_uniforms'], dtype=np.float64)[:, :r, :]
    d_scale = 0.1 * np.ones(R.shape)
    v_scale = np.clip(0.1 * Vc, 1, 15)
    scales = [d_scale, d_scale, d_scale, v_scale, v_scale, v_scale]
    noise = np.stack([(U[:, :, k] - 0.5) * scales[k][:, None] for k in range(6)], axis=-1)
    wb = w0n[:, None, :] + noise
    T_orb = J.orbital_timescale(R, z, P.potential_func, pb, ph)
    Tb = np.repeat(T_orb[:, None], r, axis=1)
    return wb, Tb * N_dyn, Tb / N_step


def density_target(dict_data, params_disk_rho):
    """Sobol cell-mass quadrature of the stellar density (model.py:192-213)."""
    S = np.asarray(dict_data['sample_for_integration'], dtype=np.float64)
    Rg = np.asarray(dict_data['R_grid'], dtype=np.float64)[:, None]
    zg = np.asarray(dict_data['z_grid'], dtype=np.float64)[:, None]
    pg = np.asarray(dict_data['phi_grid'], dtype=np.float64)[:, None]
    dR, dz, dphi = float(dict_data['dR']), float(dict_data['dz']), float(dict_data['dphi'])
    Rs = Rg + (S[:, 0] - 0.5) * dR
    zs = zg + (S[:, 1] - 0.5) * dz
    ps = pg + (S[:, 2] - 0.5) * dphi
    rho = P.density_func(Rs * np.cos(ps), Rs * np.sin(ps), zs, params_disk_rho)
    return (np.sum(rho * Rs, axis=1) / S.shape[0]) * dR * dz * dphi


def compute_model_single(w, A_Rzphi, A_xy, A_h, gamma_kin, v0, s):
    """model.py:14-66 for one weight vector (or (K, n) rows)."""
    g = NN.gamma_safe(gamma_kin)
    d3 = w @ A_Rzphi.T
    d2 = w @ A_xy.T
    h = [np.clip((w @ Ak.T) / g, -10.0, 10.0) for Ak in A_h]
    V, sig = I.h_to_V_sigma(h[0], h[1], v0, s, h3=h[2], h4=h[3])
    return d3, d2, h[0], h[1], h[2], h[3], V, sig


def _library(dict_data, pb, ph, Omega_bar, rot, num_Vbin, Rzphi_n_grid, Rzphi_lim_grid,
             xy_lim_grid, xy_n_grid, sched, traj, tape=None):
    s = sched
    r = 1 if traj else s['n_realizations']
    wb, Tt, dt0 = initial_conditions(dict_data, pb, ph, r, s['N_step_per_orb'], s['N_dynamical_time'])
    n = wb.shape[0]
    acc_fn, pot_fn = I.make_acc_pot(pb, ph)
    grids = I.Grids(num_Vbin, dict_data['bin_mapping'], dict_data['num_per_bin'],
                    Rzphi_lim_grid, xy_lim_grid, Rzphi_n_grid, xy_n_grid,
                    dict_data['v0'], dict_data['s'], rot)
    args = (wb.reshape(-1, 6), acc_fn, pot_fn, s['N_max'], Tt.reshape(-1), dt0.reshape(-1),
            -Omega_bar, s['atol'], s['rtol'], s['dt_min'], s['dt_max'], grids)
    if traj:
        per = I.integrate_withtraj(*args)
    else:
        per = I.integrate_chunked(*args, chunk_size=s['chunk'], tape=tape)
    lib = I.realization_average(per, n, r)
    lib['per'] = per
    return lib


def library_guard(lib, min_valid=0.8):
    """Check that an oracle library can make a parity comparison bite.

    ``lib``: the library dict of build_model(..., return_library=True).  The
    reference zeroes every integration that fails its validity gate
    (|E_J/E_J0 - 1| < 0.1 and t_final > T_total / 10, integrates.py:209-236);
    a library of such integrations is identically zero and a model built on
    it does not depend on the integrator, the binning or the NNLS.  Raises
    AssertionError unless at least ``min_valid`` of the integrations are valid
    and the density, surface-density and h1..h4 rows all have non-zero entries.
    Returns the valid fraction."""
    frac = float(np.mean(lib['per']['valid']))
    if frac < min_valid:
        raise AssertionError(f"only {frac:.3f} of the integrations are valid (need {min_valid})")
    for k in ('Rzphi', 'surface_density', 'h1', 'h2', 'h3', 'h4'):
        if not np.any(np.asarray(lib[k]) != 0.0):
            raise AssertionError(f"library rows {k} are all zero")
    return frac


def _targets(dict_data, params_disk_rho, n):
    """Density target and normalised data (model.py:192-258)."""
    lm = 1 / params_disk_rho['mass_to_light_ratio']
    y_Rzphi = density_target(dict_data, params_disk_rho)
    sig_Rzphi = 0.02 * y_Rzphi + 1e-10
    dd = {k: np.asarray(v, dtype=np.float64) for k, v in dict_data.items()
          if k.endswith('_data') or k.endswith('_err') or k.endswith('_standard_normal')}
    y_xy = dd['XY_density_data'] / lm
    sig_xy = (dd['XY_density_data_err'] + EPSILON) / lm
    y_h = [dd['h%d_data' % k] for k in (1, 2, 3, 4)]
    sig_h = [dd['h%d_data_err' % k] + EPSILON for k in (1, 2, 3, 4)]
    mmpo = np.sum(y_Rzphi) / n                              # model.py:231-235
    y_xy = y_xy / mmpo
    sig_xy = sig_xy / mmpo
    y_Rzphi = y_Rzphi / mmpo
    sig_Rzphi = sig_Rzphi / mmpo
    Xm = np.asarray(dict_data['X_minmax'], dtype=np.float64)
    Ym = np.asarray(dict_data['Y_minmax'], dtype=np.float64)
    nXY = np.asarray(dict_data['nX_nY'])
    area = ((Xm[1] - Xm[0]) / nXY[0]) * ((Ym[1] - Ym[0]) / nXY[1]) * 1e6
    num_per_bin = np.asarray(dict_data['num_per_bin'], dtype=np.float64)
    s_ap = np.asarray(dict_data['s'], dtype=np.float64)
    v0_ap = np.asarray(dict_data['v0'], dtype=np.float64)
    M_per_bin = y_xy * num_per_bin * area
    norm_GH = 1.0 + y_h[3] * np.sqrt(6.0) / 4.0
    gamma_kin = M_per_bin / (norm_GH * 2.0 * np.sqrt(np.pi) * s_ap + EPSILON)
    return dict(lm=lm, y_Rzphi=y_Rzphi, sig_Rzphi=sig_Rzphi, y_xy=y_xy, sig_xy=sig_xy, y_h=y_h,
                sig_h=sig_h, mmpo=mmpo, gamma_kin=gamma_kin, s_ap=s_ap, v0_ap=v0_ap, dd=dd)


def _forward(density_func, potential_func, params_halo_pot, params_disk_rho, dict_data, num_Vbin,
             Rzphi_n_tot, Rzphi_n_grid, Rzphi_lim_grid, xy_lim_grid, xy_n_grid,
             nnls_maxiter, regularization, schedule, traj, channel_w, tape=None, W_fixed=None):
    if density_func is not P.density_func or potential_func is not P.potential_func:
        raise NotImplementedError("only the canonical density/potential are restated")
    sched = schedule_of(schedule)
    pb = baryon_params(params_disk_rho)
    ph = params_halo_pot
    Omega_bar = params_disk_rho['Omega_bar']
    rot = make_rotation_matrix(params_disk_rho['alpha'], params_disk_rho['beta'], params_disk_rho['gamma'])
    lib = _library(dict_data, pb, ph, Omega_bar, rot, num_Vbin, Rzphi_n_grid, Rzphi_lim_grid,
                   xy_lim_grid, xy_n_grid, sched, traj, tape)
    out, aux = model_from_library(lib, dict_data, params_disk_rho, nnls_maxiter, regularization,
                                  channel_w, W_fixed=W_fixed)
    aux.update(rot=rot, Omega_bar=Omega_bar)
    return out, lib, aux


def model_from_library(lib, dict_data, params_disk_rho, nnls_maxiter, regularization=1.0,
                       channel_w=(np.sqrt(5.0), np.sqrt(5.0), 1.0), W_fixed=None, admm=None):
    """Everything build_model does after the library (model.py:182-286).

    ``lib``: dict of library columns (n, .) 'Rzphi', 'surface_density',
    'h1'..'h4'.  ``admm``: None for the oracle's own ADMM; or a callable
    admm(U, Y, lambda_reg, maxiter, w_init) -> W (K, n) that solves the same
    iterate sequence (nnls.py:101-163) from the design matrix and right-hand
    sides built here.  Returns (model dict, aux)."""
    A_Rzphi = lib['Rzphi'].T
    A_xy = lib['surface_density'].T
    A_h = [lib['h%d' % k].T for k in (1, 2, 3, 4)]
    n = A_Rzphi.shape[1]

    t = _targets(dict_data, params_disk_rho, n)
    y_Rzphi, sig_Rzphi, y_xy, sig_xy = t['y_Rzphi'], t['sig_Rzphi'], t['y_xy'], t['sig_xy']
    y_h, sig_h, mmpo, gamma_kin = t['y_h'], t['sig_h'], t['mmpo'], t['gamma_kin']
    v0_ap, s_ap = t['v0_ap'], t['s_ap']
    lm = t['lm']
    dd = t['dd']
    y_xy_boot = y_xy[None, :] + dd['XY_standard_normal'] * sig_xy[None, :]
    y_h_boot = [y_h[k][None, :] + dd['h%d_standard_normal' % (k + 1)] * sig_h[k][None, :]
                for k in range(4)]
    if W_fixed is None and admm is not None:
        U = NN.design_matrix(A_Rzphi, A_xy, A_h, sig_Rzphi, sig_xy, sig_h, gamma_kin, *channel_w)
        Y = NN.rhs_vectors(y_Rzphi, sig_Rzphi, y_xy_boot, y_h_boot, sig_xy, sig_h, *channel_w)
        W = np.asarray(admm(U, Y, regularization, nnls_maxiter, np.sum(y_Rzphi) / n),
                       dtype=np.float64)
    elif W_fixed is None:
        W = NN.solve_nnls_admm_bootstrap(A_Rzphi, A_xy, A_h, y_Rzphi, y_xy, gamma_kin,
                                         y_xy_boot, y_h_boot, sig_Rzphi, sig_xy, sig_h,
                                         lambda_reg=regularization, maxiter=nnls_maxiter,
                                         w_rzphi=channel_w[0], w_xy=channel_w[1], w_h=channel_w[2])
    else:
        W = np.asarray(W_fixed, dtype=np.float64)   # stop_gradient'ed weights (nnls.py:165)
    d3, d2, h1, h2, h3, h4, V, sig = compute_model_single(W, A_Rzphi, A_xy, A_h, gamma_kin, v0_ap, s_ap)
    d2 = d2 * mmpo * lm
    out = {'density_3d': y_Rzphi, 'density_3d_model': d3, 'surface_density': d2,
           'h1': h1, 'h2': h2, 'h3': h3, 'h4': h4, 'V': V, 'sigma': sig, 'weights': W}
    return out, dict(mmpo=mmpo, n=n, gamma_kin=gamma_kin, sig_Rzphi=sig_Rzphi, sig_xy=sig_xy,
                     sig_h=sig_h)


def build_model(density_func, potential_func, params_halo_pot, params_disk_rho, dict_data, num_Vbin,
                Rzphi_n_tot=360, Rzphi_n_grid=(10, 6, 6),
                Rzphi_lim_grid=((0, 10.), (-3, 3), (-np.pi, np.pi)),
                xy_lim_grid=((-10., 10.), (-3., 3.)), xy_n_grid=(60, 40),
                nnls_maxiter=200, regularization=1.0, schedule=None, return_library=False,
                tape=None, W_fixed=None):
    """model.py:70-286.  ``tape`` / ``W_fixed``: see integrate_chunked and
    model_likelihood_frozen."""
    out, lib, _ = _forward(density_func, potential_func, params_halo_pot, params_disk_rho, dict_data,
                           num_Vbin, Rzphi_n_tot, Rzphi_n_grid, Rzphi_lim_grid, xy_lim_grid,
                           xy_n_grid, nnls_maxiter, regularization, schedule, False,
                           (np.sqrt(5.0), np.sqrt(5.0), np.sqrt(1.0)), tape, W_fixed)
    if return_library:
        return out, lib
    return out


def get_model_with_orbit(density_func, potential_func, params_halo_pot, params_disk_rho, dict_data,
                         num_Vbin, Rzphi_n_tot=360, Rzphi_n_grid=(10, 6, 6),
                         Rzphi_lim_grid=((0, 10.), (-3, 3), (-np.pi, np.pi)),
                         xy_lim_grid=((-10., 10.), (-3., 3.)), xy_n_grid=(60, 40),
                         nnls_maxiter=200, regularization=1.0, schedule=None):
    """model.py:606-850 (one realization, trajectories kept, unit channel weights)."""
    out, lib, aux = _forward(density_func, potential_func, params_halo_pot, params_disk_rho,
                             dict_data, num_Vbin, Rzphi_n_tot, Rzphi_n_grid, Rzphi_lim_grid,
                             xy_lim_grid, xy_n_grid, nnls_maxiter, regularization, schedule, True,
                             (1.0, 1.0, 1.0))
    per = lib['per']
    n = aux['n']
    N = per['o_traj'].shape[1]
    res = {
        'weights': out['weights'][0], 'weights_all_boot': out['weights'],
        'y_traj': per['o_traj'].reshape(n, 1, N, 6), 't_traj': per['t_traj'].reshape(n, 1, N),
        'rotation_matrix': aux['rot'], 'Omega_bar': aux['Omega_bar'], 'n_particles': n,
        'n_realizations': 1, 'mass_per_orbit': aux['mmpo'],
    }
    for k in ('density_3d', 'density_3d_model', 'surface_density', 'h1', 'h2', 'h3', 'h4', 'V', 'sigma'):
        res[k] = out[k]
    return res


def model_deltaChi2_jackknife(density_func, potential_func, chi2_func, params_halo_pot,
                              params_disk_rho, dict_data, num_Vbin, Rzphi_n_tot=360,
                              Rzphi_n_grid=(10, 6, 6),
                              Rzphi_lim_grid=((0, 10.), (-3, 3), (-np.pi, np.pi)),
                              xy_lim_grid=((-10., 10.), (-3., 3.)), xy_n_grid=(60, 40),
                              nnls_maxiter=200, regularization=1.0, n_groups=100, batch_size=50,
                              schedule=None, return_all=False):
    """Delete-d jackknife (model.py:290-603): replicate g raises the sigma of
    the bins b with b % n_groups == g to 1e20 and re-solves the single-RHS
    NNLS with its own Gram and Cholesky; chi2_func per replicate; std over
    replicates.  batch_size only bounds the reference's memory."""
    if density_func is not P.density_func or potential_func is not P.potential_func:
        raise NotImplementedError("only the canonical density/potential are restated")
    sched = schedule_of(schedule)
    pb = baryon_params(params_disk_rho)
    rot = make_rotation_matrix(params_disk_rho['alpha'], params_disk_rho['beta'],
                               params_disk_rho['gamma'])
    lib = _library(dict_data, pb, params_halo_pot, params_disk_rho['Omega_bar'], rot, num_Vbin,
                   Rzphi_n_grid, Rzphi_lim_grid, xy_lim_grid, xy_n_grid, sched, False)
    A_Rzphi = lib['Rzphi'].T
    A_xy = lib['surface_density'].T
    A_h = [lib['h%d' % k].T for k in (1, 2, 3, 4)]
    n = A_Rzphi.shape[1]
    t = _targets(dict_data, params_disk_rho, n)
    lum = t['mmpo'] * t['lm']
    B = t['sig_xy'].shape[0]
    BIG = 1e20
    chi2 = []
    for g in range(n_groups):
        dropped = (np.arange(B) % n_groups) == g
        sxy = np.where(dropped, BIG, t['sig_xy'])
        sh = [np.where(dropped, BIG, sk) for sk in t['sig_h']]
        w = NN.solve_nnls_admm(A_Rzphi, A_xy, A_h, t['y_Rzphi'], t['y_xy'], t['gamma_kin'],
                               t['y_h'], t['sig_Rzphi'], sxy, sh, lambda_reg=regularization,
                               maxiter=nnls_maxiter, w_rzphi=np.sqrt(5.0), w_xy=np.sqrt(5.0),
                               w_h=np.sqrt(1.0))
        d3, d2, h1, h2, h3, h4, _, _ = compute_model_single(w, A_Rzphi, A_xy, A_h, t['gamma_kin'],
                                                            t['v0_ap'], t['s_ap'])
        chi2.append(chi2_func(t['y_Rzphi'], d3, t['sig_Rzphi'], t['y_xy'] * lum, d2 * lum,
                              sxy * lum, t['y_h'][0], h1, sh[0], t['y_h'][1], h2, sh[1],
                              t['y_h'][2], h3, sh[2], t['y_h'][3], h4, sh[3]))
    chi2 = np.asarray(chi2)
    out = np.std(chi2)
    return (out, chi2) if return_all else out


# --------------------------------------------------------------------------
# likelihood.py
# --------------------------------------------------------------------------
def logMenc_logc_to_logM_logRs(logM_enc, log_c, r_enc=10.0, Delta=200., rho_crit=277.54):
    """Six Newton steps (utils.py:526-569)."""
    M_enc = 10.0 ** logM_enc
    c = 10.0 ** log_c
    coeff = 4.0 * np.pi * Delta * rho_crit / 3.0
    ln10 = np.log(10.0)
    logM = logM_enc + 0.5
    for _ in range(6):
        M = 10.0 ** logM
        Rs = (M / coeff) ** (1.0 / 3.0) / c
        x = r_enc / Rs
        f_nfw = np.log(1.0 + x) - x / (1.0 + x)
        f = M * f_nfw - M_enc
        df = (f_nfw - x * x / (3.0 * (1.0 + x) ** 2)) * M * ln10
        logM = logM - f / df
    M_sol = 10.0 ** logM
    return logM, np.log10((M_sol / coeff) ** (1.0 / 3.0) / c)


def logM_logRs_to_logMenc_logc(logM_halo, logRs_halo, r_enc=10.0, Delta=200., rho_crit=277.54):
    """Forward map (utils.py:514-523); used to pin the Newton inverse."""
    M = 10.0 ** logM_halo
    Rs = 10.0 ** logRs_halo
    x = r_enc / Rs
    M_enc = M * (np.log(1.0 + x) - x / (1.0 + x))
    R_vir = (3.0 * M / (4.0 * np.pi * Delta * rho_crit)) ** (1.0 / 3.0)
    return np.log10(M_enc), np.log10(R_vir / Rs)


def theta_to_params(params):
    """theta(12) -> (params_halo_pot, params_baryon_rho) (likelihood.py:105-165)."""
    p = [float(v) for v in params]
    logM_halo, logRs_halo = logMenc_logc_to_logM_logRs(p[0], p[3])
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
    """Grid arguments model_likelihood passes to build_model (likelihood.py:167-182)."""
    Rzphi_lim = np.array([dict_data['R_minmax'], dict_data['z_minmax'], dict_data['phi_minmax']],
                         dtype=np.float64)
    xy_lim = np.array([dict_data['X_minmax'], dict_data['Y_minmax']], dtype=np.float64)
    return Rzphi_lim, np.asarray(dict_data['Rzphi_n_grid']), xy_lim, np.asarray(dict_data['nX_nY'])


def chi2_func(y_Rzphi, y_Rzphi_model, sigma_Rzphi, y_xy, y_xy_model, sigma_xy,
              y_h1, y_h1_model, sigma_h1, y_h2, y_h2_model, sigma_h2,
              y_h3, y_h3_model, sigma_h3, y_h4, y_h4_model, sigma_h4):
    """likelihood.py:73-94 (the density terms are formed but, as there, not
    returned)."""
    def res(y, ym, sg):
        return np.nansum(((y - ym) / sg) ** 2)
    return -0.5 * (res(y_h1, y_h1_model, sigma_h1) + res(y_h2, y_h2_model, sigma_h2) +
                   res(y_h3, y_h3_model, sigma_h3) + res(y_h4, y_h4_model, sigma_h4))


def calculate_delta_chi2(params, dict_data, num_Vbin, Rzphi_n_tot, schedule=None,
                         nnls_maxiter=200):
    """likelihood.py:248-335: jackknife with one group per aperture."""
    ph, pb = theta_to_params(params)
    Rzphi_lim, Rzphi_n, xy_lim, xy_n = grid_args(dict_data)
    return model_deltaChi2_jackknife(P.density_func, P.potential_func, chi2_func, ph, pb,
                                     dict_data, num_Vbin, Rzphi_n_tot, Rzphi_n, Rzphi_lim, xy_lim,
                                     xy_n, nnls_maxiter=nnls_maxiter, regularization=1.0,
                                     n_groups=num_Vbin, batch_size=25, schedule=schedule)


def bootstrap_loglike(model_dict, dict_data, sigma_amplify=1.0):
    """Bootstrap chi^2 and log-mean-exp (likelihood.py:184-244)."""
    dd = {k: np.asarray(v, dtype=np.float64) for k, v in dict_data.items()
          if k.endswith('_data') or k.endswith('_err') or k.endswith('_standard_normal')}
    y_R = np.asarray(model_dict['density_3d'])
    y_xy = dd['XY_density_data']
    sig_xy_raw = dd['XY_density_data_err'] + EPSILON
    y_xy_boot = y_xy[None, :] + dd['XY_standard_normal'] * sig_xy_raw[None, :]
    sig_R = 0.1 * y_R * sigma_amplify
    sig_xy = 0.1 * y_xy * sigma_amplify
    res_R = np.nansum(((model_dict['density_3d_model'] - y_R) / sig_R) ** 2)
    res_xy = np.nansum(((model_dict['surface_density'] - y_xy_boot) / sig_xy) ** 2, axis=1)
    res_h = []
    for k in (1, 2, 3, 4):
        sraw = dd['h%d_data_err' % k] + EPSILON
        yb = dd['h%d_data' % k][None, :] + dd['h%d_standard_normal' % k] * sraw[None, :]
        res_h.append(np.nansum(((model_dict['h%d' % k] - yb) / (sraw * sigma_amplify)) ** 2, axis=1))
    logl = -0.5 * (res_h[0] + res_h[1] + res_h[2] + res_h[3] + res_xy + res_R)
    lmax = np.max(logl)
    return lmax + np.log(np.mean(np.exp(logl - lmax))), logl


def model_likelihood(params, dict_data, num_Vbin, Rzphi_n_tot, sigma_amplify=1.0,
                     schedule=None, nnls_maxiter=200, tape=None, W_fixed=None, return_weights=False):
    """likelihood.py:96-244."""
    ph, pb = theta_to_params(params)
    Rzphi_lim, Rzphi_n, xy_lim, xy_n = grid_args(dict_data)
    md = build_model(P.density_func, P.potential_func, ph, pb, dict_data, num_Vbin, Rzphi_n_tot,
                     Rzphi_n, Rzphi_lim, xy_lim, xy_n, nnls_maxiter=nnls_maxiter,
                     regularization=1.0, schedule=schedule, tape=tape, W_fixed=W_fixed)
    val = bootstrap_loglike(md, dict_data, sigma_amplify)[0]
    return (val, md['weights']) if return_weights else val


def model_likelihood_grad_fd(params, dict_data, num_Vbin, Rzphi_n_tot, sigma_amplify=1.0,
                             schedule=None, nnls_maxiter=200, W_fixed=None, h=5e-5):
    """The gradient jax.grad takes of model_likelihood, by central differences.

    jax.grad of likelihood.py:96-244 differentiates the piece of the
    piecewise-smooth map on which theta lies: accept / reject decisions, bin
    membership and validity flags are constants (integrates.py:103-111,
    149-159, 210-236), and so are the branches of the error scale's max and
    of the controller's clips; the NNLS weights are stop_gradient'ed
    (nnls.py:165).  So the base run records those decisions and every
    perturbed run replays them with the base (or the given) weights; the
    fourth-order central difference of that smooth map is the same derivative
    up to O(h^4) and the map's roundoff (~1e-10 relative, from the nested
    finite differences of the Jeans moments, jeans.py:37-60) / h.

    Returns (logL at theta, gradient (12,), W).
    """
    theta = np.asarray(params, dtype=np.float64)
    tape = {'mode': 'record'}
    f0, W = model_likelihood(theta, dict_data, num_Vbin, Rzphi_n_tot, sigma_amplify, schedule,
                             nnls_maxiter, tape=tape, W_fixed=W_fixed, return_weights=True)
    tape['mode'] = 'replay'
    g = np.zeros(theta.size)
    for i in range(theta.size):
        e = np.zeros(theta.size)
        e[i] = h
        f = {k: model_likelihood(theta + k * e, dict_data, num_Vbin, Rzphi_n_tot, sigma_amplify,
                                 schedule, nnls_maxiter, tape=tape, W_fixed=W) for k in (-2, -1, 1, 2)}
        g[i] = (8.0 * (f[1] - f[-1]) - (f[2] - f[-2])) / (12.0 * h)
    return f0, g, W
