"""Seeded synthetic galaxies with the shapes of the benchmark workloads.

Builds ``dict_data`` in the layout the reference's loader
produces (utils.py:572-743) — Voronoi apertures on a pixel grid with a ``-1``
sentinel after the last pixel, the (R, z, phi) cell-centre arrays sorted by
cell index, the unscrambled 1024-point Sobol set, and bootstrap normals drawn
exactly as utils.py:604-618 draws them (numpy default_rng(42), row 0 zeroed).
The JAX-keyed draws the reference makes inside build_model (Jeans normals,
jitter uniforms; model.py:112, 128-138) are drawn here from the galaxy's
seeded NumPy generator and carried in ``dict_data['jeans_normals']`` and
``dict_data['jitter_uniforms']``, so that every seed gives other draws;
without those keys the model takes the reference's own streams.

Observations (surface density and h1..h4) are made by ``observe`` from an orbit
library and a seeded Dirichlet weight vector, plus Gaussian noise.  Everything
here is host NumPy except ``observed_workload``, which builds the library it
observes on the GPU.
"""
import math

import numpy as np
from scipy.stats import qmc

from .data import rzphi_cells, bootstrap_normals  # noqa: F401  (shared with the loader)

KMS_TO_KPCGYR = 1.022712165045695
ARCSEC_RAD = math.pi / (180.0 * 3600.0)

# Workload shapes.  Sizes follow the benchmark configurations; the integration
# schedule maps "P bar periods, S samples per orbit" onto the reference's
# N_dynamical_time (validity horizon, in orbital times) and N_max (number of
# BS3 attempts), keeping its dt_init = T_orb / 100.
#
# plain-2k (1,000 samples at rtol 1e-6) takes its horizon from the reference's
# own rule N_max = N_step_per_orb * N_dynamical_time (model.py:163-165), so
# N_dynamical_time = 10.  At rtol 1e-6 the BS3 controller takes about 385
# attempts per orbital time, so 1,000 attempts cover a median 2.6 orbital times
# and pass the validity gate t_final > T_total / 10 (integrates.py:210) in 99% of
# integrations; a 50-period horizon would fail it in all of them.
CONFIGS = {
    'plain-2k': dict(n=2000, N_max=1000, N_dyn=10, rtol=1e-6, n_ap=300, field_arcsec=40,
                     grid=(20, 10, 16), K=1, T=500, seeded_params=False),
    'plain-20k': dict(n=20000, N_max=2000, N_dyn=100, rtol=1e-4, n_ap=1000, field_arcsec=60,
                      grid=(32, 16, 24), K=100, T=1000, seeded_params=True),
    'batched-64': dict(n=(5000, 40000), N_max=2000, N_dyn=100, rtol=1e-4, n_ap=1000,
                       field_arcsec=60, grid=(32, 16, 24), K=100, T=1000, seeded_params=True,
                       n_models=64),
    'grad-20k': dict(n=20000, N_max=2000, N_dyn=100, rtol=1e-4, n_ap=1000, field_arcsec=60,
                     grid=(32, 16, 24), K=100, T=1000, seeded_params=True),
    'jackknife-20k': dict(n=20000, N_max=2000, N_dyn=100, rtol=1e-4, n_ap=1000, field_arcsec=60,
                          grid=(32, 16, 24), K=100, T=200, seeded_params=True),
    'batched-grad-16': dict(n=40000, N_max=2000, N_dyn=100, rtol=1e-4, n_ap=2000,
                            field_arcsec=80, grid=(40, 20, 32), K=200, T=1000,
                            seeded_params=True, n_models=16),
}


def logM_logRs_to_logMenc_logc(logM_halo, logRs_halo, r_enc=10.0, Delta=200., rho_crit=277.54):
    """NFW (logM, logRs) -> (logM_enc(<r_enc), log c) (utils.py:514-523)."""
    M = 10.0 ** logM_halo
    Rs = 10.0 ** logRs_halo
    x = r_enc / Rs
    M_enc = M * (np.log(1.0 + x) - x / (1.0 + x))
    R_vir = (3.0 * M / (4.0 * np.pi * Delta * rho_crit)) ** (1.0 / 3.0)
    return float(np.log10(M_enc)), float(np.log10(R_vir / Rs))


def base_theta(rho_s=1e7, r_s=20.0, M_disc=5e10, a_disc=3.0, b_disc=0.3, M_bar=1e10, L_bar=4.0,
               Omega_kms_kpc=40.0, alpha=0.0, beta=65.0, gamma=45.0, M_to_L=1.0):
    """The 12-vector theta (likelihood.py:105-120) of a physical parameter set.

    NFW mass parameter: the reference's NFW_potential uses M = 4 pi rho_s r_s^3.
    Angles in degrees here, radians in theta.
    """
    logM_halo = math.log10(4 * math.pi * rho_s * r_s ** 3)
    logMenc, logc = logM_logRs_to_logMenc_logc(logM_halo, math.log10(r_s))
    return np.array([logMenc, math.log10(M_disc), math.log10(M_bar), logc, math.log10(a_disc),
                     math.log10(b_disc), math.log10(L_bar), math.radians(alpha),
                     math.radians(beta), math.radians(gamma), math.log10(M_to_L),
                     math.log10(Omega_kms_kpc * KMS_TO_KPCGYR)])


def seeded_theta(rng):
    """Parameters drawn as the seeded configurations describe: Omega_b in
    [30, 50] km/s/kpc, M/L in [0.5, 1.5], r_s in [10, 30] kpc."""
    return base_theta(r_s=rng.uniform(10.0, 30.0), Omega_kms_kpc=rng.uniform(30.0, 50.0),
                      M_to_L=rng.uniform(0.5, 1.5))


def batched_thetas(n_models, rng):
    """theta vectors uniform over the 12 continuous parameters (batched workload)."""
    out = []
    for _ in range(n_models):
        out.append(base_theta(rho_s=10 ** rng.uniform(6.7, 7.3), r_s=rng.uniform(10, 30),
                              M_disc=10 ** rng.uniform(10.5, 10.9), a_disc=rng.uniform(2.5, 3.5),
                              b_disc=rng.uniform(0.25, 0.4), M_bar=10 ** rng.uniform(9.8, 10.2),
                              L_bar=rng.uniform(3.5, 4.5), Omega_kms_kpc=rng.uniform(30, 50),
                              beta=rng.uniform(55, 75), gamma=rng.uniform(35, 55),
                              M_to_L=rng.uniform(0.5, 1.5)))
    return out


def voronoi_apertures(nX, nY, n_ap, rng):
    """Pixel -> aperture map from n_ap seeded generator pixels, plus sentinel -1."""
    gen = rng.choice(nX * nY, size=n_ap, replace=False)
    gx, gy = (gen % nX).astype(np.float64), (gen // nX).astype(np.float64)
    px = np.arange(nX * nY)
    X, Y = (px % nX).astype(np.float64), (px // nX).astype(np.float64)
    d2 = (X[:, None] - gx[None, :]) ** 2 + (Y[:, None] - gy[None, :]) ** 2
    owner = np.argmin(d2, axis=1).astype(np.int32)
    num = np.bincount(owner, minlength=n_ap).astype(np.int32)
    return np.concatenate([owner, np.array([-1], dtype=np.int32)]), num


def disc_positions(n, rng, R_d=3.0, h_z=0.3, R_max=12.0):
    """Double-exponential disc: R ~ R exp(-R/R_d) (truncated), |z| ~ exp(-|z|/h_z)."""
    R = np.empty(0)
    while R.size < n:
        r = rng.gamma(2.0, R_d, size=2 * n)
        R = np.concatenate([R, r[r < R_max]])
    R = R[:n]
    phi = rng.uniform(0.0, 2 * np.pi, size=n)
    z = rng.exponential(h_z, size=n) * rng.choice([-1.0, 1.0], size=n)
    return np.stack([R * np.cos(phi), R * np.sin(phi), z], axis=-1)


def make_galaxy(config='plain-2k', seed=0, n=None, n_ap=None, grid=None, K=None, nxy=None,
                distance_mpc=50.0, r=4, N_max=None):
    """Geometry, positions and noise draws for one galaxy (observations empty).

    ``N_max``: a shorter (or longer) schedule than the configuration's, with
    the horizon set by the reference's rule N_dynamical_time = N_max / 100
    (model.py:163-165).

    Returns (theta, dict_data, schedule).
    """
    cfg = dict(CONFIGS[config])
    rng = np.random.default_rng(seed)
    if n is None:
        n = cfg['n'] if not isinstance(cfg['n'], tuple) else int(rng.integers(*cfg['n']))
    n_ap = n_ap or cfg['n_ap']
    nR, nz, nphi = grid or cfg['grid']
    K = K or cfg['K']
    field = cfg['field_arcsec']
    nX = nY = nxy or field
    half = 0.5 * field * ARCSEC_RAD * distance_mpc * 1e3
    theta = seeded_theta(rng) if cfg['seeded_params'] else base_theta()
    bm, num = voronoi_apertures(nX, nY, n_ap, rng)
    Rg, zg, pg, dR, dz, dphi = rzphi_cells(nR, nz, nphi)
    sob = qmc.Sobol(d=3, scramble=False).random_base2(m=10)
    w0 = disc_positions(n, rng)
    dd = {
        'w0': w0,
        'jeans_normals': rng.standard_normal((n, 3)),
        'jitter_uniforms': rng.uniform(size=(n, r, 6)),
        'bin_mapping': bm, 'num_per_bin': num, 'total_bins': n_ap,
        'X_minmax': np.array([-half, half]), 'Y_minmax': np.array([-half, half]),
        'nX_nY': np.array([nX, nY]),
        'R_grid': Rg, 'z_grid': zg, 'phi_grid': pg, 'dR': dR, 'dz': dz, 'dphi': dphi,
        'R_minmax': [0.0, 10.0], 'z_minmax': [-3.0, 3.0], 'phi_minmax': [-np.pi, np.pi],
        'Rzphi_n_tot': nR * nz * nphi, 'Rzphi_n_grid': np.array([nR, nz, nphi]),
        'sample_for_integration': sob,
        'v0': np.zeros(n_ap), 's': np.full(n_ap, 150.0),
    }
    bootstrap_normals(dd, K, n_ap)
    sched = dict(N_max=cfg['N_max'], N_dynamical_time=cfg['N_dyn'], N_step_per_orb=100,
                 rtol=cfg['rtol'], atol=1e-7, dt_min=1e-5, dt_max=0.3, chunk=100,
                 n_realizations=r)
    if N_max is not None:
        sched['N_max'] = int(N_max)
        sched['N_dynamical_time'] = int(N_max) / sched['N_step_per_orb']
    return theta, dd, sched


def sky_projection(w, R):
    """Sky (x', y') and LOS velocity (km/s) of bar-frame phase-space points."""
    xr = w[:, :3] @ R.T
    vr = w[:, 3:] @ R.T
    return xr[:, 0], xr[:, 1], vr[:, 2] / KMS_TO_KPCGYR


def fit_base_gaussians(dd, w_ics, rot, weights=None):
    """Per-aperture (v0, s): weighted mean and dispersion of the tracers' LOS
    velocities, a Gaussian fit standing in for the data's own."""
    nX, nY = dd['nX_nY']
    Xm, Ym = dd['X_minmax'], dd['Y_minmax']
    x, y, v = sky_projection(w_ics, rot)
    ix = np.floor((x - Xm[0]) / (Xm[1] - Xm[0]) * nX).astype(np.int64)
    iy = np.floor((y - Ym[0]) / (Ym[1] - Ym[0]) * nY).astype(np.int64)
    ok = (ix >= 0) & (ix < nX) & (iy >= 0) & (iy < nY)
    B = int(dd['total_bins'])
    ap = dd['bin_mapping'][ix[ok] + iy[ok] * nX]
    wt = np.ones(ok.sum()) if weights is None else weights[ok]
    sw = np.bincount(ap, wt, minlength=B)
    m1 = np.bincount(ap, wt * v[ok], minlength=B)
    m2 = np.bincount(ap, wt * v[ok] ** 2, minlength=B)
    vall = np.average(v, weights=weights)
    sall = np.sqrt(np.average((v - vall) ** 2, weights=weights))
    with np.errstate(invalid='ignore', divide='ignore'):
        mean = np.where(sw > 0, m1 / sw, vall)
        var = np.where(sw > 0, m2 / sw - mean ** 2, sall ** 2)
    s = np.sqrt(np.clip(var, (0.3 * sall) ** 2, None))
    s = np.where(sw >= 3, s, sall)
    dd['v0'] = np.where(sw >= 3, mean, vall)
    dd['s'] = s


def observe(dd, A_xy, A_h, mmpo, M_to_L, seed=1, noise=0.02, h_err=0.03, floor=None):
    """Observed surface density and h1..h4 from library columns and Dirichlet
    weights (n,), following the model's own conventions (model.py:217-272).

    ``floor``: raise the surface density to at least floor x its median
    positive value, so that no aperture has zero data (the likelihood divides
    by 0.1 y_xy, likelihood.py:215, so zero data make its gradient NaN)."""
    rng = np.random.default_rng(seed)
    n = A_xy.shape[1]
    w = rng.dirichlet(np.ones(n)) * n
    sd_model = A_xy @ w                                   # rescaled mass units
    Xm, Ym, nXY = dd['X_minmax'], dd['Y_minmax'], dd['nX_nY']
    area = ((Xm[1] - Xm[0]) / nXY[0]) * ((Ym[1] - Ym[0]) / nXY[1]) * 1e6
    M_b = sd_model * dd['num_per_bin'] * area
    h4 = np.zeros_like(sd_model)
    for _ in range(3):
        gk = M_b / ((1.0 + h4 * np.sqrt(6.0) / 4.0) * 2.0 * np.sqrt(np.pi) * dd['s'] + 1e-12)
        gk = np.where(np.abs(gk) > 1e-8, gk, 1e-8)
        h = [np.clip((Ak @ w) / gk, -10, 10) for Ak in A_h]
        h4 = h[3]
    B = sd_model.size
    sd = sd_model * mmpo / M_to_L
    dd['XY_density_data'] = sd * (1.0 + noise * rng.standard_normal(B))
    if floor is not None:
        pos = dd['XY_density_data'][dd['XY_density_data'] > 0]
        dd['XY_density_data'] = np.maximum(dd['XY_density_data'], floor * np.median(pos))
    dd['XY_density_data_err'] = 0.01 * dd['XY_density_data'] + 1e-12
    for k in range(4):
        dd['h%d_data' % (k + 1)] = h[k] + noise * rng.standard_normal(B)
        dd['h%d_data_err' % (k + 1)] = np.full(B, h_err)
    sq2, sq3, sq6 = np.sqrt(2.0), np.sqrt(3.0), np.sqrt(6.0)
    dv = sq2 * dd['h1_data'] + sq3 * dd['h3_data']
    dd['V_data'] = dd['v0'] + dd['s'] * dv
    dd['sigma_data'] = dd['s'] * np.sqrt(np.clip(1 + 2 * sq2 * dd['h2_data'] + 2 * sq6 * dd['h4_data']
                                                 - dv ** 2, 1e-10, None))
    dd['V_data_err'] = np.full(B, 5.0)
    dd['sigma_data_err'] = np.full(B, 5.0)
    return w


def observed_workload(config, seed, n=None, return_library=False):
    """A galaxy of a benchmark configuration with observations, on the GPU.

    The initial conditions and the orbit library come from the device path
    (smx_initial_conditions, smx_build_library); (v0, s) are fitted to the
    tracers' line-of-sight velocities and the observations are made from
    that library by ``observe`` (surface density floored at 1e-3 of its median).
    Returns (theta, dict_data, schedule, info) with info = {'P', 'B',
    'valid'} (the valid fraction of the integrations) and, when
    ``return_library``, 'lib' (the device library, (n, P + 6B) FP64)."""
    import torch
    from . import model as M, likelihood as L, potentials as POT
    theta, dd, sched = make_galaxy(config, seed=seed, n=n)
    ph, pb = L.theta_to_params(theta)
    s = M.schedule_of(sched)
    pot = POT.pack_potential(M.baryon_params(pb), ph)
    wb, Tt, dt0 = M.initial_conditions(pot, dd, s['n_realizations'], s)
    rot = M.makeRotationMatrix(pb['alpha'], pb['beta'], pb['gamma'])
    fit_base_gaussians(dd, wb[:, 0, :].cpu().numpy(), rot)
    Rl, Rn, xyl, xyn = L.grid_args(dd)
    g = M.make_grids(int(dd['total_bins']), Rn, Rl, xyl, xyn, rot, True)
    libd = M.build_library(pot, g, M.make_schedule(s, -pb['Omega_bar']), wb, Tt, dt0, dd)
    lib = libd['lib'].cpu().numpy()
    P, B = libd['P'], libd['B']
    A_xy = lib[:, P:P + B].T
    A_h = [lib[:, P + B * k:P + B * (k + 1)].T for k in (1, 2, 3, 4)]
    y_R = M.density_target(pb, dd).cpu().numpy()
    observe(dd, A_xy, A_h, y_R.sum() / lib.shape[0], pb['mass_to_light_ratio'], floor=1e-3)
    info = {'P': P, 'B': B, 'valid': float(libd['valid'].mean().item())}
    if return_library:
        info['lib'] = libd['lib']
    del libd, wb, Tt, dt0
    torch.cuda.empty_cache()
    return theta, dd, sched, info
