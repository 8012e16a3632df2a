"""Adaptive BS2(3) integration in the rotating bar frame, with binning.

Restates integrates.py:8-313 (chunked path with 4-fold symmetry) and
integrates.py:319-581 (trajectory path, no symmetry), utils.py:10-42 (grid
assignment) and ghmoments.py:6-36 (Gauss-Hermite basis), vectorised over
integrations: state arrays have shape (6, M) for M independent integrations.
"""
import numpy as np

from .constants import EPSILON, KPCGYR_TO_KMS
from . import potential as P

SQ2, SQ3 = 2 ** 0.5, 3 ** 0.5
INV_SQRT_2PI = (2 * np.pi) ** -0.5

# 4-fold bar symmetry (integrates.py:74-79)
SIGN_SYM = np.array([
    [1, 1, 1, 1, 1, 1],
    [1, 1, -1, 1, 1, -1],
    [-1, -1, 1, -1, -1, 1],
    [-1, -1, -1, -1, -1, -1],
], dtype=np.float64)


def make_acc_pot(params_baryon, params_halo):
    """acc_fn / pot_fn closures of build_model (model.py:152-161)."""
    def acc_fn(x, y, z):
        ax, ay, az, _ = P.accel(x, y, z, params_baryon, params_halo)
        return ax, ay, az

    def pot_fn(x, y, z):
        return P.potential_func(x, y, z, params_baryon, params_halo)
    return acc_fn, pot_fn


def deriv_barred(y, acc_fn, Omega):
    """Co-rotating-frame RHS (integrates.py:8-27); y has shape (6, M)."""
    ax, ay, az = acc_fn(y[0], y[1], y[2])
    return np.stack([y[3] + Omega * y[1], y[4] - Omega * y[0], y[5],
                     ax + Omega * y[4], ay - Omega * y[3], az])


def compute_EJ(y, pot_fn, Omega):
    """Jacobi energy (integrates.py:29-35)."""
    E_kin = 0.5 * (y[3] ** 2 + y[4] ** 2 + y[5] ** 2)
    return E_kin + pot_fn(y[0], y[1], y[2]) - Omega * (y[0] * y[4] - y[1] * y[3])


def _nan_max(a, b):
    return np.maximum(a, b)          # NaN-propagating, like jnp.maximum


def _clip(a, lo, hi):
    return np.minimum(np.maximum(a, lo), hi)


def rk_step(t, y, dt, k1, acc_fn, Omega, atol, rtol, dt_min, dt_max):
    """One BS3 attempt (integrates.py:90-112).  Returns (t, y, dt, k1, dt_out)."""
    return rk_step_dec(t, y, dt, k1, acc_fn, Omega, atol, rtol, dt_min, dt_max)[:5]


def rk_step_dec(t, y, dt, k1, acc_fn, Omega, atol, rtol, dt_min, dt_max, dec_in=None):
    """rk_step that also returns its branch decisions; ``dec_in`` replays them.

    The decisions are where the attempt's map is not smooth: accept / reject,
    which of |y| and |y_new| the error scale takes, whether the error floor
    1e-10 and the two clips of the controller are active.  jax.grad
    differentiates the branch taken; replaying the branches of a base run
    keeps nearby evaluations on that branch, for finite differences.
    Returns (t, y, dt, k1, dt_out, decisions).
    """
    e1, e2, e3, e4 = -5.0 / 72.0, 1.0 / 12.0, 1.0 / 9.0, -1.0 / 8.0
    k2 = deriv_barred(y + 0.5 * dt * k1, acc_fn, Omega)
    k3 = deriv_barred(y + 0.75 * dt * k2, acc_fn, Omega)
    y_new = y + dt * (2.0 / 9.0 * k1 + 1.0 / 3.0 * k2 + 4.0 / 9.0 * k3)
    k4 = deriv_barred(y_new, acc_fn, Omega)
    err_vec = dt * (e1 * k1 + e2 * k2 + e3 * k3 + e4 * k4)
    ay, an = np.fabs(y), np.fabs(y_new)
    sel = ay >= an if dec_in is None else dec_in['sel']
    scale = atol + rtol * np.where(sel, ay, an)
    es = (err_vec / scale) ** 2
    err_norm = np.sqrt((((((es[0] + es[1]) + es[2]) + es[3]) + es[4]) + es[5]) / 6.0)
    accept = err_norm <= 1.0 if dec_in is None else dec_in['accept']
    y_out = np.where(accept, y_new, y)
    t_new = np.where(accept, t + dt, t)
    k1_new = np.where(accept, k4, k1)
    floor = (err_norm < 1e-10) if dec_in is None else dec_in['floor']
    sf_raw = 0.9 * np.power(np.where(floor, 1e-10, err_norm), -1.0 / 3.0)
    if dec_in is None:
        sf_st = np.where(sf_raw < 0.2, -1, np.where(sf_raw > 5.0, 1, 0))
    else:
        sf_st = dec_in['sf']
    sf = np.where(sf_st < 0, 0.2, np.where(sf_st > 0, 5.0, sf_raw))
    dtp = dt * sf
    if dec_in is None:
        dt_st = np.where(dtp < dt_min, -1, np.where(dtp > dt_max, 1, 0))
    else:
        dt_st = dec_in['dt']
    dt_new = np.where(dt_st < 0, dt_min, np.where(dt_st > 0, dt_max, dtp))
    dt_out = np.where(accept, dt, 0.0)
    dec = {'accept': accept, 'sel': sel, 'floor': floor, 'sf': sf_st, 'dt': dt_st, 'err': err_norm}
    return t_new, y_out, dt_new, k1_new, dt_out, dec


# --------------------------------------------------------------------------
# grid assignment and Gauss-Hermite basis
# --------------------------------------------------------------------------
def assign_regular_grid(positions, grid_min, grid_max, n_bins):
    """utils.py:10-42.  positions: (N, D).  OOB rows -> prod(n_bins)."""
    grid_min = np.asarray(grid_min, dtype=np.float64)
    grid_max = np.asarray(grid_max, dtype=np.float64)
    n_bins = np.asarray(n_bins, dtype=np.int64)
    strides = np.concatenate([[1], np.cumprod(n_bins[:-1])])
    normalized = (positions - grid_min) / (grid_max - grid_min)
    scaled = normalized * n_bins.astype(np.float64)
    with np.errstate(invalid='ignore'):
        coords = np.where(np.isnan(scaled), 0.0, np.floor(scaled))
    coords = np.clip(coords, 0, n_bins - 1).astype(np.int64)
    idx = np.sum(coords * strides, axis=-1)
    oob = np.any((normalized >= 1.0) | (normalized < 0.0), axis=-1)
    return np.where(oob, int(np.prod(n_bins)), idx)


def gauss(x, v, s):
    """ghmoments.py:6-9."""
    y = (x - v) / s
    return y, INV_SQRT_2PI / s * np.exp(-0.5 * y ** 2)


def hermite(y):
    """H0..H4 by the orthonormal recurrence (ghmoments.py:11-36)."""
    H0 = 1.0 + y * 0
    H1 = SQ2 * y
    H2 = (SQ2 * y * H1 - 1.0 * H0) / SQ2
    H3 = (SQ2 * y * H2 - SQ2 * H1) / SQ3
    H4 = (SQ2 * y * H3 - SQ3 * H2) / 2.0
    return H0, H1, H2, H3, H4


def h_to_V_sigma(h1, h2, v0, s, h3=0.0, h4=0.0):
    """ghmoments.py:53-79."""
    sqrt2, sqrt3, sqrt6 = np.sqrt(2.0), np.sqrt(3.0), np.sqrt(6.0)
    delta_v = sqrt2 * h1 + sqrt3 * h3
    v_mean = v0 + s * delta_v
    sig2 = 1.0 + 2.0 * sqrt2 * h2 + 2.0 * sqrt6 * h4 - delta_v ** 2
    return v_mean, s * np.sqrt(np.maximum(sig2, 1e-10))


class Grids:
    """Binning geometry shared by both integrator paths."""

    def __init__(self, num_Vbin, bin_mapping, num_per_bin, Rzphi_minmax, XY_minmax,
                 nRzphi, nXY, v0, s, rotation_matrix):
        self.num_Vbin = int(num_Vbin)
        self.bin_mapping = np.asarray(bin_mapping, dtype=np.int64)
        self.num_per_bin = np.asarray(num_per_bin, dtype=np.float64)
        self.Rzphi_minmax = np.asarray(Rzphi_minmax, dtype=np.float64)
        self.XY_minmax = np.asarray(XY_minmax, dtype=np.float64)
        self.nRzphi = np.asarray(nRzphi, dtype=np.int64)
        self.nXY = np.asarray(nXY, dtype=np.int64)
        self.P = int(np.prod(self.nRzphi))
        self.v0 = np.asarray(v0, dtype=np.float64)
        self.s = np.asarray(s, dtype=np.float64)
        self.R = np.asarray(rotation_matrix, dtype=np.float64)
        # integrates.py:70-71
        self.area_pixel = ((self.XY_minmax[0, 1] - self.XY_minmax[0, 0]) / self.nXY[0]) * \
                          ((self.XY_minmax[1, 1] - self.XY_minmax[1, 0]) / self.nXY[1]) * 1e6

    def indices(self, ys):
        """Rzphi cell and aperture of samples ys (6, S) -> (idx_rzphi, vbin, vlos)."""
        R_vals = np.sqrt(ys[0] ** 2 + ys[1] ** 2)
        phi = np.arctan2(ys[1], ys[0])
        rzphi = assign_regular_grid(np.stack([R_vals, ys[2], phi], axis=-1),
                                    self.Rzphi_minmax[:, 0], self.Rzphi_minmax[:, 1], self.nRzphi)
        Rm = self.R
        xr = (Rm[0, 0] * ys[0] + Rm[0, 1] * ys[1]) + Rm[0, 2] * ys[2]
        yr = (Rm[1, 0] * ys[0] + Rm[1, 1] * ys[1]) + Rm[1, 2] * ys[2]
        vzr = (Rm[2, 0] * ys[3] + Rm[2, 1] * ys[4]) + Rm[2, 2] * ys[5]
        xy = assign_regular_grid(np.stack([xr, yr], axis=-1),
                                 self.XY_minmax[:, 0], self.XY_minmax[:, 1], self.nXY)
        # bin_mapping[XY_indices] (integrates.py:140): off-field samples carry
        # id nX*nY; a map without the sentinel entry (utils.py:429 allows both
        # lengths) is gathered with JAX's clamp of out-of-range indices.
        vbin = self.bin_mapping[np.minimum(xy, self.bin_mapping.size - 1)]
        return rzphi, vbin, vzr * KPCGYR_TO_KMS

    def gh_terms(self, vbin, vlos):
        """gauss weight and H0..H4 at the (wrapped) aperture of each sample."""
        vb = np.where(vbin < 0, vbin + len(self.v0), vbin)    # jnp negative-index wrap
        vb = np.clip(vb, 0, len(self.v0) - 1)
        y, g = gauss(vlos, self.v0[vb], self.s[vb])
        return g, hermite(y)


def _segsum(M, nseg, owner, idx, w):
    """Per-integration segment_sum that drops idx outside [0, nseg)."""
    ok = (idx >= 0) & (idx < nseg)
    flat = owner[ok] * nseg + idx[ok]
    return np.bincount(flat, weights=w[ok], minlength=M * nseg).reshape(M, nseg)


def integrate_chunked(w0, acc_fn, pot_fn, N_max, T_total, dt_init, Omega, atol, rtol,
                      dt_min, dt_max, grids, chunk_size, symmetry=True, want_h0=True, tape=None):
    """integrate_adaptive_barred_chunked (integrates.py:37-238) for M integrations.

    w0: (M, 6); T_total, dt_init: (M,).  Returns dict of per-integration
    Rzphi (M,P), surface_density (M,B), h0..h4 (M,B), valid (M,), T (M,),
    t_final (M,).

    ``tape``: None, or a dict with 'mode' 'record' (the discrete decisions of
    this run -- accept flags, bin ids, validity -- are stored in it) or
    'replay' (the stored decisions are used instead of recomputed), so that
    nearby parameters are evaluated on the same piece of the piecewise-smooth
    map jax.grad differentiates.
    """
    rec = tape is not None and tape['mode'] == 'record'
    rep = tape is not None and tape['mode'] == 'replay'
    if rec:
        tape.update(accept=[], irz=[], vbin=[])
    it_acc = iter(tape['accept']) if rep else None
    it_bin = iter(zip(tape['irz'], tape['vbin'])) if rep else None
    w0 = np.asarray(w0, dtype=np.float64)
    M = w0.shape[0]
    B = grids.num_Vbin
    y = w0.T.copy()
    k1 = deriv_barred(y, acc_fn, Omega)
    EJ0 = compute_EJ(y, pot_fn, Omega)
    t = np.zeros(M)
    dt = np.asarray(dt_init, dtype=np.float64) * np.ones(M)
    rz = np.zeros((M, grids.P))
    counts = np.zeros((M, B))
    sw = np.zeros((5, M, B))
    T_acc = np.zeros(M)
    signs = SIGN_SYM if symmetry else SIGN_SYM[:1]
    nsym = signs.shape[0]
    owner = np.broadcast_to(np.arange(M), (nsym, chunk_size, M)).reshape(-1)
    for _ in range(N_max // chunk_size):          # n_chunks = N_max // chunk_size (integrates.py:63)
        ys, dts = [], []
        for _ in range(chunk_size):
            t, y, dt, k1, dt_out, dec = rk_step_dec(t, y, dt, k1, acc_fn, Omega, atol, rtol,
                                                    dt_min, dt_max,
                                                    next(it_acc) if rep else None)
            if rec:
                tape['accept'].append(dec)
            ys.append(y)
            dts.append(dt_out)
        yc = np.stack(ys, axis=1)                         # (6, C, M)
        dc = np.stack(dts, axis=0)                        # (C, M)
        ysym = (signs[:, :, None, None] * yc[None]).transpose(1, 0, 2, 3).reshape(6, -1)
        dsym = np.broadcast_to(dc, (nsym, chunk_size, M)).reshape(-1)
        irz, vbin, vlos = grids.indices(ysym)
        if rec:
            tape['irz'].append(irz)
            tape['vbin'].append(vbin)
        if rep:
            irz, vbin = next(it_bin)
        rz += _segsum(M, grids.P, owner, irz, dsym)
        counts += _segsum(M, B, owner, vbin, dsym)
        g, H = grids.gh_terms(vbin, vlos)
        for k in range(5):
            if k == 0 and not want_h0:
                continue
            sw[k] += _segsum(M, B, owner, vbin, g * dsym * H[k])
        T_acc += dsym.reshape(nsym * chunk_size, M).sum(axis=0)
    EJ = compute_EJ(y, pot_fn, Omega)
    valid = np.where((np.fabs(EJ / EJ0 - 1.0) < 0.1) & (t > np.asarray(T_total) / 10), 1.0, 0.0)
    if rec:
        tape['valid'] = valid
    if rep:
        valid = tape['valid']
    T_int = T_acc + EPSILON
    v = valid[:, None] > 0.5
    out = {
        'Rzphi': np.where(v, rz / T_int[:, None], 0.0),
        'surface_density': np.where(v, (counts / T_int[:, None]) /
                                    (grids.num_per_bin * grids.area_pixel + EPSILON), 0.0),
        'valid': valid, 'T': T_int, 't_final': t, 'y_final': y.T.copy(),
    }
    for k in range(5):
        out['h%d' % k] = np.where(v, sw[k] / T_int[:, None], 0.0)
    return out


def realization_average(per, n, r):
    """integrate_adaptive_batch_chunked averaging (integrates.py:279-300).

    per: output of integrate_chunked over n*r integrations ordered (column,
    realization).  Returns the library columns (n, .) for each quantity and
    the number of valid realizations per column.
    """
    valid = per['valid'].reshape(n, r)
    wt = 1.0 / (valid.sum(axis=1) + 0.1)
    out = {}
    for key in ('Rzphi', 'surface_density', 'h0', 'h1', 'h2', 'h3', 'h4'):
        a = per[key].reshape(n, r, -1)
        out[key] = np.einsum('nrk,n->nk', a, wt)
    out['n_valid'] = valid.sum(axis=1)
    return out


def integrate_withtraj(w0, acc_fn, pot_fn, N_max, T_total, dt_init, Omega, atol, rtol,
                       dt_min, dt_max, grids, record=None):
    """integrate_adaptive_barred_withtraj (integrates.py:319-512), M integrations.

    Returns per-integration binned quantities (no symmetry, dt pre-normalised)
    plus o_traj (M, N, 6) with velocities in km/s and t_traj (M, N).
    ``record``: a dict that receives each attempt's error norm 'err' (M, N)
    and accept flag 'accept' (M, N), for first-divergence reports.
    """
    w0 = np.asarray(w0, dtype=np.float64)
    M = w0.shape[0]
    B = grids.num_Vbin
    y = w0.T.copy()
    k1 = deriv_barred(y, acc_fn, Omega)
    EJ0 = compute_EJ(y, pot_fn, Omega)
    t = np.zeros(M)
    dt = np.asarray(dt_init, dtype=np.float64) * np.ones(M)
    traj = np.empty((N_max, 6, M))
    dts = np.empty((N_max, M))
    errs = np.empty((N_max, M))
    accs = np.empty((N_max, M), dtype=bool)
    for i in range(N_max):
        t, y, dt, k1, dt_out, dec = rk_step_dec(t, y, dt, k1, acc_fn, Omega, atol, rtol, dt_min,
                                                dt_max)
        traj[i] = y
        dts[i] = dt_out
        errs[i] = dec['err']
        accs[i] = dec['accept']
    if record is not None:
        record['err'] = errs.T.copy()
        record['accept'] = accs.T.copy()
    t_traj = np.cumsum(dts, axis=0).T                     # (M, N)
    EJ = compute_EJ(y, pot_fn, Omega)
    valid = np.where((np.fabs(EJ / EJ0 - 1.0) < 0.1) & (t > np.asarray(T_total) / 10), 1.0, 0.0)
    T_int = dts.sum(axis=0) + EPSILON
    dt_norm = (dts / T_int[None, :]).reshape(-1)         # (N*M,) step-major
    ys = traj.transpose(1, 0, 2).reshape(6, -1)
    owner = np.broadcast_to(np.arange(M), (N_max, M)).reshape(-1)
    irz, vbin, vlos = grids.indices(ys)
    rz = _segsum(M, grids.P, owner, irz, dt_norm)
    g, H = grids.gh_terms(vbin, vlos)
    counts = _segsum(M, B, owner, vbin, dt_norm)
    v = valid[:, None] > 0.5
    out = {
        'Rzphi': np.where(v, rz, 0.0),
        'surface_density': np.where(v, counts / (grids.num_per_bin * grids.area_pixel + EPSILON), 0.0),
        'valid': valid, 'T': T_int, 't_final': t,
    }
    for k in range(5):
        out['h%d' % k] = np.where(v, _segsum(M, B, owner, vbin, dt_norm * g * H[k]), 0.0)
    o = traj.transpose(2, 0, 1).copy()                    # (M, N, 6)
    o[:, :, 3:] *= KPCGYR_TO_KMS
    out['o_traj'] = o
    out['t_traj'] = t_traj
    return out
