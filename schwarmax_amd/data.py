"""Observed-data boundary: the reference's loader without JAX, and a
documented on-disk format for its binned data and for dict_data.

``load_data_bootstrap`` (utils.py:572-743) turns a binned-kinematics
dictionary (the "bin dict") into the dict_data the model reads.  The
reference unpickles the bin dict; pickles can run code on load, and the
reference's own carry JAX array reconstructors, so here the bin dict is an
``.npz`` archive read with ``allow_pickle=False``:

============================  =====================================  =========
key                           meaning                                shape
============================  =====================================  =========
X_minmax, Y_minmax            field of view (kpc)                    (2,)
nX_nY                         regular pixel grid                     (2,) int
X_regular_grid,               pixel centres                          (nY, nX)
Y_regular_grid                                                       or (nX*nY,)
bin_mapping                   pixel -> aperture id (+ sentinel)      (nX*nY+1,) int
num_per_bin                   pixels per aperture                    (B,)
total_bins                    B                                      () int
surface_density               aperture surface density               (B,)
V_mean, V_sigma, h1..h4       kinematic moments                      (B,)
V_mean_err, V_sigma_err,      their errors                           (B,)
h1_err..h4_err
v0, s                         Gauss-Hermite base Gaussian            (B,)
orientation                   (alpha, beta, gamma)                   (3,)
============================  =====================================  =========

``write_bin_dict`` writes one from a dict of arrays (for instance a
reference pickle loaded once in the user's own JAX environment).

``save_dict_data`` / ``load_dict_data`` store a complete dict_data (the
loader's output, or a synthetic galaxy) in the same way, with every key the
model reads, so an analysis can skip the loader.

Reproduced exactly: the bootstrap normals (NumPy ``default_rng(42)``), the
(R, z, phi) cell centres and their order, the Sobol points, the azimuths of
the initial positions (a fresh ``default_rng(42)``) and the uniforms behind
the initial radii and heights (JAX's threefry stream, ``schwarmax_amd.prng``).
The inverse-CDF radii and heights are computed in float32 like the
reference's (JAX float32 defaults); their CDF sums follow NumPy's order,
so they agree with JAX's to float32 rounding.  The cell centres, spacings
and errors are float64 here where the reference holds float32 copies.
"""
import os

import numpy as np
from scipy.stats import qmc

from . import prng

EPSILON = 1e-12
FORMAT_KEY = '__smx_dict_data__'
FORMAT_VERSION = 1

BIN_DICT_KEYS = ('X_minmax', 'Y_minmax', 'nX_nY', 'num_per_bin', 'total_bins', 'bin_mapping',
                 'surface_density', 'V_mean', 'V_sigma', 'h1', 'h2', 'h3', 'h4', 'v0', 's',
                 'orientation', 'V_mean_err', 'V_sigma_err', 'h1_err', 'h2_err', 'h3_err',
                 'h4_err', 'X_regular_grid', 'Y_regular_grid')

_LIST_KEYS = ('R_minmax', 'z_minmax', 'phi_minmax')
_INT_KEYS = ('total_bins', 'Rzphi_n_tot')


def _npz_path(filename):
    if str(filename).endswith('.pkl') or str(filename).endswith('.pickle'):
        raise ValueError(
            "%s is a pickle; pickles are not read here.  Convert it once where it can be "
            "trusted and unpickled: np.savez(out, **{k: np.asarray(v) for k, v in d.items()}) "
            "and pass the .npz" % filename)
    return filename


def read_bin_dict(filename):
    """The bin dict in an .npz archive (no pickled objects)."""
    with np.load(_npz_path(filename), allow_pickle=False) as z:
        d = {k: z[k] for k in z.files}
    missing = [k for k in BIN_DICT_KEYS if k not in d]
    if missing:
        raise KeyError("bin dict %s lacks %s" % (filename, ', '.join(missing)))
    return d


def _plain(v, key):
    a = np.asarray(v)
    if a.dtype == object:
        raise TypeError("%s holds Python objects; store numbers only" % key)
    return a


def write_bin_dict(filename, bin_dict):
    """Write the bin dict keys of bin_dict to an .npz archive."""
    missing = [k for k in BIN_DICT_KEYS if k not in bin_dict]
    if missing:
        raise KeyError("bin dict lacks %s" % ', '.join(missing))
    np.savez(filename, **{k: _plain(bin_dict[k], k) for k in BIN_DICT_KEYS})


def assign_regular_grid(points, lo, hi, n):
    """Flat cell id of each point (utils.py:10-42): q = (x - lo)/(hi - lo),
    out of range if q < 0 or q >= 1, else floor(q n) clipped to [0, n-1];
    id = i0 + i1 n0 + i2 n0 n1, and prod(n) (one past the last cell) when out
    of range."""
    points = np.asarray(points, dtype=np.float64)
    lo, hi, n = (np.asarray(a) for a in (lo, hi, n))
    q = (points - lo) / (hi - lo)
    oob = np.any((q < 0) | (q >= 1), axis=-1)
    c = np.clip(np.floor(q * n).astype(np.int64), 0, n - 1)
    strides = np.concatenate([[1], np.cumprod(n[:-1])])
    return np.where(oob, int(np.prod(n)), c @ strides)


def rzphi_cells(nR, nz, nphi, Rmax=10.0, zmax=3.0):
    """(R, z, phi) cell centres in cell-id order and the spacings
    (utils.py:627-657)."""
    lo = np.array([0.0, -zmax, -np.pi])
    hi = np.array([Rmax, zmax, np.pi])
    n = np.array([nR, nz, nphi])
    edges = [np.linspace(lo[k], hi[k], n[k] + 1) for k in range(3)]
    mids = [0.5 * (e[:-1] + e[1:]) for e in edges]
    mesh = np.meshgrid(*mids, indexing='ij')
    g = np.stack([m.ravel() for m in mesh], axis=-1)
    order = np.argsort(assign_regular_grid(g, lo, hi, n), kind='stable')
    g = g[order]
    d = [np.unique(g[:, k])[1] - np.unique(g[:, k])[0] for k in range(3)]
    return g[:, 0], g[:, 1], g[:, 2], d[0], d[1], d[2]


def bootstrap_normals(dd, K, B):
    """Bootstrap standard normals as utils.py:604-618 draws them: one
    default_rng(42) stream, (K, B) each for XY, h1..h4, V, sigma in that
    order, row 0 zeroed (the unperturbed data)."""
    rng = np.random.default_rng(42)
    for nm in ('XY', 'h1', 'h2', 'h3', 'h4', 'V', 'sigma'):
        a = rng.normal(size=(K, B))
        a[0, :] = 0.0
        dd[nm + '_standard_normal'] = a


def XexpX_pdf_log(x, a):
    """log of x exp(-x/a) / a^2 on x >= 0 (utils.py:369-389), float32."""
    x = np.asarray(x, dtype=np.float32)
    with np.errstate(divide='ignore'):
        pdf = np.log(x) - np.log(np.float32(a * a)) - x / np.float32(a)
    return np.where(x >= 0, pdf, -np.inf).astype(np.float32)


def expX_pdf_log(x, a):
    """log(a) - x/a on x >= 0 (utils.py:391-411, as written there), float32."""
    x = np.asarray(x, dtype=np.float32)
    pdf = np.log(np.float32(a)) - x / np.float32(a)
    return np.where(x >= 0, pdf, -np.inf).astype(np.float32)


def initial_positions(n_samples, partitionable=True):
    """The loader's tracer positions (utils.py:668-686): R from x exp(-x/3) on
    [0, 12] (key 10086), z from exp(-x/1.2) on [0, 5] (key 10010), both by
    inverse CDF on 1,000 grid points; azimuth uniform on [0, 2 pi) from a
    fresh default_rng(42).  float32 (n_samples, 3)."""
    xg = np.linspace(0.0, 12.0, 1000)
    R = prng.sample_from_logP(xg, XexpX_pdf_log(xg, 3), n_samples, prng.PRNGKey(10086),
                              partitionable)
    phi = np.random.default_rng(42).uniform(0, 2 * np.pi, size=n_samples)
    x = R * np.cos(phi).astype(np.float32)
    y = R * np.sin(phi).astype(np.float32)
    zg = np.linspace(0, 5, 1000)
    z = prng.sample_from_logP(zg, expX_pdf_log(zg, 1.2), n_samples, prng.PRNGKey(10010),
                              partitionable)
    return np.stack([x, y, z], axis=-1).astype(np.float32)


def _spacing(a):
    u = np.unique(np.asarray(a, dtype=np.float64))
    return float(u[1] - u[0])


def load_data_bootstrap(path, filename, N_BOOTSTRAP=100, n_samples=5_000,
                        jax_threefry_partitionable=True):
    """utils.py:572-743: dict_data from the bin dict at path + filename.

    Same keys and meaning as the reference's output; arrays are NumPy.
    ``jax_threefry_partitionable=False`` reproduces the initial positions of
    JAX releases before 0.5.0.
    """
    bd = read_bin_dict(os.fspath(path) + os.fspath(filename))
    f64 = lambda k: np.asarray(bd[k], dtype=np.float64)  # noqa: E731
    sd = f64('surface_density')
    B = sd.shape[0]
    dd = {
        'v0': f64('v0'), 's': f64('s'),
        'XY_density_data': sd, 'XY_density_data_err': 0.01 * sd + EPSILON,
        'V_data': f64('V_mean'), 'V_data_err': f64('V_mean_err'),
        'sigma_data': f64('V_sigma'), 'sigma_data_err': f64('V_sigma_err'),
    }
    for k in (1, 2, 3, 4):
        dd['h%d_data' % k] = f64('h%d' % k)
        dd['h%d_data_err' % k] = f64('h%d_err' % k)
    dd['num_per_bin'] = np.asarray(bd['num_per_bin'])
    dd['bin_mapping'] = np.asarray(bd['bin_mapping'])
    dd['total_bins'] = int(np.asarray(bd['total_bins']))
    for k in ('V', 'sigma', 'h1', 'h2', 'h3', 'h4'):
        if dd[k + '_data'].shape[0] != B:
            raise ValueError("%s_data has %d apertures, surface_density %d"
                             % (k, dd[k + '_data'].shape[0], B))
    bootstrap_normals(dd, int(N_BOOTSTRAP), B)

    nR, nz, nphi = 10, 6, 10
    Rg, zg, pg, dR, dz, dphi = rzphi_cells(nR, nz, nphi)
    sob = qmc.Sobol(d=3, scramble=False).random_base2(m=10)
    Xr, Yr = np.asarray(bd['X_regular_grid']), np.asarray(bd['Y_regular_grid'])
    dd.update({
        'R_grid': Rg, 'z_grid': zg, 'phi_grid': pg,
        'R_minmax': [0.0, 10.0], 'z_minmax': [-3.0, 3.0], 'phi_minmax': [-np.pi, np.pi],
        'Rzphi_n_tot': nR * nz * nphi, 'Rzphi_n_grid': np.array([nR, nz, nphi]),
        'dR': float(dR), 'dz': float(dz), 'dphi': float(dphi),
        'sample_for_integration': sob,
        'X_regular_grid': Xr, 'Y_regular_grid': Yr, 'dX': _spacing(Xr), 'dY': _spacing(Yr),
        'sample_for_integration_XY': qmc.Sobol(d=3, scramble=False).random_base2(m=10),
        'X_minmax': f64('X_minmax'), 'Y_minmax': f64('Y_minmax'),
        'nX_nY': np.asarray(bd['nX_nY']),
        'w0': initial_positions(int(n_samples), jax_threefry_partitionable),
    })
    return dd


def save_dict_data(filename, dict_data):
    """Store dict_data as an .npz archive of plain arrays."""
    out = {FORMAT_KEY: np.array(FORMAT_VERSION)}
    for k, v in dict_data.items():
        if k == FORMAT_KEY:
            continue
        a = np.asarray(v.detach().cpu().numpy() if hasattr(v, 'detach') else v)
        out[k] = _plain(a, k)
    np.savez(filename, **out)


def load_dict_data(filename):
    """dict_data written by save_dict_data (no pickled objects)."""
    with np.load(_npz_path(filename), allow_pickle=False) as z:
        if FORMAT_KEY not in z.files:
            raise ValueError("%s was not written by save_dict_data" % filename)
        if int(z[FORMAT_KEY]) != FORMAT_VERSION:
            raise ValueError("%s has format version %d, this reader %d"
                             % (filename, int(z[FORMAT_KEY]), FORMAT_VERSION))
        dd = {k: z[k] for k in z.files if k != FORMAT_KEY}
    for k, v in list(dd.items()):
        if k in _LIST_KEYS:
            dd[k] = [float(x) for x in v]
        elif k in _INT_KEYS:
            dd[k] = int(v)
        elif v.ndim == 0:
            dd[k] = v.item()
    return dd
