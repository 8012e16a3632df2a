"""Forward model on the GPU: build_model / get_model_with_orbit drop-ins.

Same signatures, argument meaning and output keys as model.py:70-286
(build_model) and model.py:606-850 (get_model_with_orbit).  Every per-orbit,
per-cell and per-iteration computation runs in the HIP kernels of libsmx.so;
torch holds device memory and does the few O(K*B) elementwise steps between
kernels (data normalisation, clipping, V/sigma).

The reference draws the Jeans normals and jitter uniforms inside build_model
from fixed JAX PRNG keys (model.py:112, 128-138).  Here they are the same
draws, from the NumPy restatement of JAX's generator in ``prng``, unless
dict_data carries its own as ``dict_data['jeans_normals']`` (n, 3) and
``dict_data['jitter_uniforms']`` (n, r, 6).  The integration schedule that
build_model hard-codes (model.py:163-169) is the default of ``schedule``.
"""
import numpy as np
import torch

from . import _lib
from . import potentials as POT
from . import prng
from .utils import makeRotationMatrix

EPSILON = 1e-12
DEFAULT_SCHEDULE = dict(N_step_per_orb=100, N_dynamical_time=50, N_max=None, atol=1e-7,
                        rtol=1e-4, dt_min=1e-5, dt_max=0.3, chunk=100, n_realizations=4)


def schedule_of(sched):
    s = dict(DEFAULT_SCHEDULE)
    if sched:
        s.update(sched)
    if s['N_max'] is None:
        s['N_max'] = int(s['N_step_per_orb'] * s['N_dynamical_time'])
    return s


def _canonical(density_func, potential_func):
    from . import likelihood
    if density_func is not likelihood.density_func or potential_func is not likelihood.potential_func:
        raise NotImplementedError(
            "the GPU path implements the canonical density_func / potential_func of "
            "likelihood.py (NFW + Miyamoto-Nagai + T3 bar + V4 bulge); pass "
            "schwarmax_amd.likelihood.density_func / potential_func")


def baryon_params(params_disk_rho):
    """params_baryon of model.py:87-102."""
    keys = ('logM_disc', 'Rs_disc', 'Hs_disc', 'logM_bar', 'L_bar', 'a_bar', 'b_bar',
            'x_origin', 'y_origin', 'z_origin', 'dirx', 'diry', 'dirz')
    return {k: params_disk_rho[k] for k in keys}


def make_grids(num_Vbin, Rzphi_n_grid, Rzphi_lim_grid, xy_lim_grid, xy_n_grid, rot, symmetry):
    g = _lib.SmxGrids()
    nR, nz, nphi = (int(v) for v in np.asarray(Rzphi_n_grid).ravel())
    nX, nY = (int(v) for v in np.asarray(xy_n_grid).ravel())
    lim = np.asarray(Rzphi_lim_grid, dtype=np.float64)
    xyl = np.asarray(xy_lim_grid, dtype=np.float64)
    g.nR, g.nz, g.nphi, g.nX, g.nY = nR, nz, nphi, nX, nY
    g.num_vbin = int(num_Vbin)
    g.symmetry = 1 if symmetry else 0
    g.rz_min[:] = list(lim[:, 0])
    g.rz_max[:] = list(lim[:, 1])
    g.xy_min[:] = list(xyl[:, 0])
    g.xy_max[:] = list(xyl[:, 1])
    g.rot[:] = list(np.asarray(rot, dtype=np.float64).ravel())
    # integrates.py:70-71
    g.area_pixel = ((xyl[0, 1] - xyl[0, 0]) / nX) * ((xyl[1, 1] - xyl[1, 0]) / nY) * 1e6
    return g


def make_schedule(s, Omega, chunked=True):
    """The kernels' schedule.  The chunked integrator runs N_max // chunk
    chunks of chunk attempts (integrates.py:63), so a N_max that is not a
    multiple of the chunk size gives (N_max // chunk) * chunk attempts; the
    trajectory integrator runs all N_max (integrates.py:382-384)."""
    sc = _lib.SmxSchedule()
    N = int(s['N_max'])
    if chunked:
        N = (N // int(s['chunk'])) * int(s['chunk'])
    sc.N_max = N
    sc.Omega = float(Omega)
    sc.atol, sc.rtol = float(s['atol']), float(s['rtol'])
    sc.dt_min, sc.dt_max = float(s['dt_min']), float(s['dt_max'])
    return sc


def ic_draws(dict_data, r):
    """Jeans normals (n, 3) and jitter uniforms (n, r_total >= r, 6).

    dict_data's 'jeans_normals' and 'jitter_uniforms' when it has them;
    otherwise the reference's own draws (model.py:112-138, jeans.py:89: keys
    42, 109, 2026 and split(911, 6), in JAX's partitionable threefry layout)
    for max(r, 4) realizations, the reference's 4 first.
    """
    has = ('jeans_normals' in dict_data, 'jitter_uniforms' in dict_data)
    if all(has):
        normals, uni = dict_data['jeans_normals'], dict_data['jitter_uniforms']
    elif any(has):
        raise ValueError("dict_data must hold both 'jeans_normals' and 'jitter_uniforms' or neither")
    else:
        n = np.asarray(dict_data['w0']).shape[0]
        normals, uni = prng.reference_draws(n, max(int(r), 4))
    n = np.asarray(dict_data['w0']).shape[0]
    if tuple(normals.shape) != (n, 3) or len(uni.shape) != 3 or uni.shape[0] != n \
            or uni.shape[2] != 6 or uni.shape[1] < r:
        raise ValueError("draws must be (n, 3) normals and (n, >= r, 6) uniforms for n = %d" % n)
    return normals, uni


def initial_conditions(pot, dict_data, r, sched):
    """Jeans ICs, jitter and time scales on the GPU (model.py:112-169)."""
    L = _lib.lib()
    w0 = _lib.dev(np.asarray(dict_data['w0'])[:, :3])
    n = w0.shape[0]
    normals, uni = (_lib.dev(a) for a in ic_draws(dict_data, r))
    r_total = uni.shape[1]
    wb = torch.empty((n, r, 6), dtype=torch.float64, device='cuda')
    Tt = torch.empty((n, r), dtype=torch.float64, device='cuda')
    dt0 = torch.empty((n, r), dtype=torch.float64, device='cuda')
    _lib.check(L.smx_initial_conditions(pot, _lib.ptr(w0), _lib.ptr(normals), _lib.ptr(uni), n, r,
                                        r_total, float(sched['N_step_per_orb']),
                                        float(sched['N_dynamical_time']), _lib.ptr(wb), _lib.ptr(Tt),
                                        _lib.ptr(dt0), _lib.stream_ptr()), 'smx_initial_conditions')
    return wb, Tt, dt0


def build_library(pot, grids, sched_struct, wb, Tt, dt0, dict_data, traj=False):
    """Orbit library (n, P + 6B) columns [Rzphi | xy | h1 | h2 | h3 | h4 | h0]."""
    L = _lib.lib()
    n, r = wb.shape[0], wb.shape[1]
    n_int = n * r
    P = grids.nR * grids.nz * grids.nphi
    B = grids.num_vbin
    ws_bytes = (L.smx_library_workspace_bytes(grids, n_int) if traj
                else L.smx_library_build_bytes(grids, sched_struct, n_int))
    ws = torch.empty(int(ws_bytes), dtype=torch.uint8, device='cuda')
    lib = torch.empty((n, P + 6 * B), dtype=torch.float64, device='cuda')
    valid = torch.empty(n_int, dtype=torch.float64, device='cuda')
    T_int = torch.empty(n_int, dtype=torch.float64, device='cuda')
    bm = _lib.bin_map(dict_data['bin_mapping'], grids.nX, grids.nY)
    npb = _lib.dev(dict_data['num_per_bin'])
    v0 = _lib.velocities(dict_data['v0'])
    sd = _lib.dispersions(dict_data['s'])
    extra = {}
    if traj:
        N = sched_struct.N_max
        o_traj = torch.empty((n_int, N, 6), dtype=torch.float64, device='cuda')
        t_traj = torch.empty((n_int, N), dtype=torch.float64, device='cuda')
        _lib.check(L.smx_build_library_traj(pot, grids, sched_struct, _lib.ptr(wb), _lib.ptr(Tt),
                                            _lib.ptr(dt0), n, r, _lib.ptr(bm), _lib.ptr(npb),
                                            _lib.ptr(v0), _lib.ptr(sd), _lib.ptr(ws), _lib.ptr(lib),
                                            _lib.ptr(o_traj), _lib.ptr(t_traj), _lib.ptr(valid),
                                            _lib.ptr(T_int), _lib.stream_ptr()),
                   'smx_build_library_traj')
        extra = {'o_traj': o_traj, 't_traj': t_traj}
    else:
        t_final = torch.empty(n_int, dtype=torch.float64, device='cuda')
        _lib.check(L.smx_build_library(pot, grids, sched_struct, _lib.ptr(wb), _lib.ptr(Tt),
                                       _lib.ptr(dt0), n, r, _lib.ptr(bm), _lib.ptr(npb), _lib.ptr(v0),
                                       _lib.ptr(sd), _lib.ptr(ws), ws.numel(), _lib.ptr(lib),
                                       _lib.ptr(valid), _lib.ptr(T_int), _lib.ptr(t_final),
                                       _lib.stream_ptr()),
                   'smx_build_library')
        extra = {'t_final': t_final}
    del ws
    return dict(lib=lib, valid=valid, T=T_int, P=P, B=B, **extra)


def density_target(params_disk_rho, dict_data):
    """y_Rzphi by Sobol quadrature on the GPU (model.py:192-213)."""
    L = _lib.lib()
    Rg = _lib.dev(dict_data['R_grid'])
    zg = _lib.dev(dict_data['z_grid'])
    pg = _lib.dev(dict_data['phi_grid'])
    smp = _lib.dev(dict_data['sample_for_integration'])
    P = Rg.numel()
    out = torch.empty(P, dtype=torch.float64, device='cuda')
    _lib.check(L.smx_density_target(POT.pack_potential(params_disk_rho), _lib.ptr(Rg), _lib.ptr(zg),
                                    _lib.ptr(pg), P, float(dict_data['dR']), float(dict_data['dz']),
                                    float(dict_data['dphi']), _lib.ptr(smp), smp.shape[0],
                                    _lib.ptr(out), _lib.stream_ptr()), 'smx_density_target')
    return out


def compute_model_single(W, lib, P, B, gamma_kin, v0, s, mmpo=1.0, lm=1.0):
    """Model maps for weight rows W (K, n) (model.py:14-66): one FP64 GEMM
    (smx_model_maps), then smx_model_epilogue for the surface density
    (map * mmpo) * lm, the clamped h1..h4 and V, sigma (ghmoments.py:53-79).
    d3 is a view of the maps."""
    from .nnls import model_maps
    L = _lib.lib()
    m = P + 5 * B
    maps = model_maps(lib, m, W)
    K = maps.shape[0]
    f64 = torch.float64
    d2 = torch.empty((K, B), dtype=f64, device='cuda')
    h = torch.empty((4, K, B), dtype=f64, device='cuda')
    V = torch.empty((K, B), dtype=f64, device='cuda')
    sig = torch.empty((K, B), dtype=f64, device='cuda')
    g, v0, s = (_lib.dev(t) for t in (gamma_kin, v0, s))
    _lib.check(L.smx_model_epilogue(_lib.ptr(maps, f64), m, K, P, B, _lib.ptr(g, f64),
                                    _lib.ptr(v0, f64), _lib.ptr(s, f64), float(mmpo), float(lm),
                                    _lib.ptr(d2, f64), _lib.ptr(h, f64), _lib.ptr(V, f64),
                                    _lib.ptr(sig, f64), _lib.stream_ptr()), 'smx_model_epilogue')
    return maps[:, :P], d2, h[0], h[1], h[2], h[3], V, sig


def _forward(density_func, potential_func, params_halo_pot, params_disk_rho, dict_data, num_Vbin,
             Rzphi_n_tot, Rzphi_n_grid, Rzphi_lim_grid, xy_lim_grid, xy_n_grid, nnls_maxiter,
             regularization, schedule, traj, channel_w, nnls_tol):
    _canonical(density_func, potential_func)
    sched = schedule_of(schedule)
    pb = baryon_params(params_disk_rho)
    pot = POT.pack_potential(pb, params_halo_pot)
    Omega_bar = float(params_disk_rho['Omega_bar'])
    rot = makeRotationMatrix(float(params_disk_rho['alpha']), float(params_disk_rho['beta']),
                             float(params_disk_rho['gamma']))
    r = 1 if traj else int(sched['n_realizations'])
    grids = make_grids(num_Vbin, Rzphi_n_grid, Rzphi_lim_grid, xy_lim_grid, xy_n_grid, rot,
                       symmetry=not traj)
    if grids.nR * grids.nz * grids.nphi != int(Rzphi_n_tot):
        raise ValueError("Rzphi_n_tot must equal prod(Rzphi_n_grid)")
    wb, Tt, dt0 = initial_conditions(pot, dict_data, r, sched)
    libd = build_library(pot, grids, make_schedule(sched, -Omega_bar, chunked=not traj), wb, Tt,
                         dt0, dict_data, traj)
    out, aux = model_from_library(libd['lib'], libd['P'], libd['B'], params_disk_rho, dict_data,
                                  nnls_maxiter, regularization, channel_w, nnls_tol)
    aux.update(libd=libd, rot=rot, Omega_bar=Omega_bar, wb=wb, T_total=Tt, dt0=dt0)
    return out, aux


def prepare_targets(n, params_disk_rho, dict_data):
    """Density target and normalised data of model.py:192-258 (mass per
    orbit, M/L, gamma_kin), as device tensors."""
    lm = 1.0 / float(params_disk_rho['mass_to_light_ratio'])
    y_R = density_target(params_disk_rho, dict_data)
    sig_R = 0.02 * y_R + 1e-10
    D = lambda k: _lib.dev(dict_data[k])  # noqa: E731
    y_xy = D('XY_density_data') / lm
    sig_xy = (D('XY_density_data_err') + EPSILON) / lm
    y_h = [D('h%d_data' % k) for k in (1, 2, 3, 4)]
    sig_h = [D('h%d_data_err' % k) + EPSILON for k in (1, 2, 3, 4)]
    mmpo = torch.sum(y_R) / n                                   # model.py:231-235
    y_xy = y_xy / mmpo
    sig_xy = sig_xy / mmpo
    y_R = y_R / mmpo
    sig_R = sig_R / mmpo
    Xm = np.asarray(dict_data['X_minmax'], dtype=np.float64)
    Ym = np.asarray(dict_data['Y_minmax'], dtype=np.float64)
    nXY = np.asarray(dict_data['nX_nY'])
    area = ((Xm[1] - Xm[0]) / nXY[0]) * ((Ym[1] - Ym[0]) / nXY[1]) * 1e6
    npb = D('num_per_bin')
    s_ap, v0_ap = D('s'), D('v0')
    M_per_bin = y_xy * npb * area
    norm_GH = 1.0 + y_h[3] * np.sqrt(6.0) / 4.0
    gamma_kin = M_per_bin / (norm_GH * 2.0 * np.sqrt(np.pi) * s_ap + EPSILON)
    return dict(lm=lm, y_R=y_R, sig_R=sig_R, y_xy=y_xy, sig_xy=sig_xy, y_h=y_h, sig_h=sig_h,
                mmpo=mmpo, gamma_kin=gamma_kin, s_ap=s_ap, v0_ap=v0_ap)


def _bootstrap_data(n, params_disk_rho, dict_data):
    """Targets and the bootstrapped observations (model.py:192-258)."""
    t = prepare_targets(n, params_disk_rho, dict_data)
    D = lambda k: _lib.dev(dict_data[k])  # noqa: E731
    y_xy, sig_xy, y_h, sig_h = t['y_xy'], t['sig_xy'], t['y_h'], t['sig_h']
    t['y_xy_boot'] = y_xy[None, :] + D('XY_standard_normal') * sig_xy[None, :]
    t['y_h_boot'] = [y_h[k][None, :] + D('h%d_standard_normal' % (k + 1)) * sig_h[k][None, :]
                     for k in range(4)]
    return t


def _model_maps(W, lib, P, B, t):
    """The model dict of model.py:260-286 from the weights."""
    d3, d2, h1, h2, h3, h4, V, sig = compute_model_single(W, lib, P, B, t['gamma_kin'], t['v0_ap'],
                                                          t['s_ap'], t['mmpo'], t['lm'])
    return {'density_3d': t['y_R'], 'density_3d_model': d3, 'surface_density': d2, 'h1': h1,
            'h2': h2, 'h3': h3, 'h4': h4, 'V': V, 'sigma': sig, 'weights': W}


def model_from_library(lib, P, B, params_disk_rho, dict_data, nnls_maxiter, regularization,
                       channel_w, nnls_tol=0.0):
    """Everything build_model does after the library (model.py:192-286):
    density target, data normalisation, bootstrap NNLS and the model maps."""
    from .nnls import solve_nnls_admm_bootstrap_lib
    n = lib.shape[0]
    t = _bootstrap_data(n, params_disk_rho, dict_data)
    W, info = solve_nnls_admm_bootstrap_lib(lib, P, B, t['y_R'], t['gamma_kin'], t['y_xy_boot'],
                                            t['y_h_boot'], t['sig_R'], t['sig_xy'], t['sig_h'],
                                            lambda_reg=regularization, maxiter=nnls_maxiter,
                                            w_rzphi=channel_w[0], w_xy=channel_w[1],
                                            w_h=channel_w[2], tol=nnls_tol)
    return _model_maps(W, lib, P, B, t), dict(mmpo=t['mmpo'], n=n, nnls_iters=info['iters'])


def models_from_libraries(libs, P, B, params_disk_rhos, dict_data, nnls_maxiter, regularization,
                          channel_w, arena=None):
    """model_from_library for several models (one galaxy, the same P and B) with
    their NNLS solves grouped: one ADMM product launch and one update launch
    per iteration for all of them (``nnls.solve_group``; ``arena`` as there).
    Each model's result equals its model_from_library call (nnls_tol = 0)."""
    from .nnls import solve_group
    ts = [_bootstrap_data(lib.shape[0], pb, dict_data) for lib, pb in zip(libs, params_disk_rhos)]
    Ws = solve_group(libs, P, B, [(t['y_R'], t['gamma_kin'], t['y_xy_boot'], t['y_h_boot'],
                                   t['sig_R'], t['sig_xy'], t['sig_h']) for t in ts],
                     regularization, nnls_maxiter, *channel_w, arena=arena)
    return [(_model_maps(W, lib, P, B, t), dict(mmpo=t['mmpo'], n=lib.shape[0],
                                                nnls_iters=int(nnls_maxiter)))
            for W, lib, t in zip(Ws, libs, ts)]


def build_model(density_func, potential_func, params_halo_pot, params_disk_rho, dict_data, num_Vbin,
                Rzphi_n_tot=360, Rzphi_n_grid=(10, 6, 6),
                Rzphi_lim_grid=((0, 10.), (-3, 3), (-np.pi, np.pi)),
                xy_lim_grid=((-10., 10.), (-3., 3.)), xy_n_grid=(60, 40),
                nnls_maxiter=200, regularization=1.0, schedule=None, nnls_tol=0.0,
                return_aux=False):
    """model.py:70-286 on the GPU.  Returns the reference's dict (torch, on device)."""
    out, aux = _forward(density_func, potential_func, params_halo_pot, params_disk_rho, dict_data,
                        num_Vbin, Rzphi_n_tot, Rzphi_n_grid, Rzphi_lim_grid, xy_lim_grid, xy_n_grid,
                        nnls_maxiter, regularization, schedule, False,
                        (np.sqrt(5.0), np.sqrt(5.0), np.sqrt(1.0)), nnls_tol)
    return (out, aux) if return_aux else out


def get_model_with_orbit(density_func, potential_func, params_halo_pot, params_disk_rho, dict_data,
                         num_Vbin, Rzphi_n_tot=360, Rzphi_n_grid=(10, 6, 6),
                         Rzphi_lim_grid=((0, 10.), (-3, 3), (-np.pi, np.pi)),
                         xy_lim_grid=((-10., 10.), (-3., 3.)), xy_n_grid=(60, 40),
                         nnls_maxiter=200, regularization=1.0, schedule=None, nnls_tol=0.0):
    """model.py:606-850 on the GPU (one realization, trajectories kept)."""
    out, aux = _forward(density_func, potential_func, params_halo_pot, params_disk_rho, dict_data,
                        num_Vbin, Rzphi_n_tot, Rzphi_n_grid, Rzphi_lim_grid, xy_lim_grid, xy_n_grid,
                        nnls_maxiter, regularization, schedule, True, (1.0, 1.0, 1.0), nnls_tol)
    libd = aux['libd']
    n = aux['n']
    N = libd['o_traj'].shape[1]
    res = {
        'weights': out['weights'][0], 'weights_all_boot': out['weights'],
        'y_traj': libd['o_traj'].reshape(n, 1, N, 6), 't_traj': libd['t_traj'].reshape(n, 1, N),
        'rotation_matrix': aux['rot'], 'Omega_bar': aux['Omega_bar'], 'n_particles': n,
        'n_realizations': 1, 'mass_per_orbit': aux['mmpo'],
    }
    for k in ('density_3d', 'density_3d_model', 'surface_density', 'h1', 'h2', 'h3', 'h4', 'V', 'sigma'):
        res[k] = out[k]
    return res
