"""The data boundary (CPU): load_data_bootstrap (utils.py:572-743) without JAX,
the .npz bin-dict and dict_data formats, and a loader-made dict_data run
through the oracle end to end."""
import math

import numpy as np
import pytest

from schwarmax_amd import data as D, prng as P, synth

# The keys the reference's loader returns (utils.py:689-741).
REFERENCE_KEYS = {
    'v0', 's', 'XY_density_data', 'XY_density_data_err', 'V_data', 'V_data_err', 'sigma_data',
    'sigma_data_err', 'h1_data', 'h1_data_err', 'h2_data', 'h2_data_err', 'h3_data',
    'h3_data_err', 'h4_data', 'h4_data_err', 'num_per_bin', 'bin_mapping', 'total_bins',
    'XY_standard_normal', 'h1_standard_normal', 'h2_standard_normal', 'h3_standard_normal',
    'h4_standard_normal', 'V_standard_normal', 'sigma_standard_normal', 'R_grid', 'z_grid',
    'phi_grid', 'R_minmax', 'z_minmax', 'phi_minmax', 'Rzphi_n_tot', 'Rzphi_n_grid', 'dR', 'dz',
    'dphi', 'sample_for_integration', 'X_regular_grid', 'Y_regular_grid', 'dX', 'dY',
    'sample_for_integration_XY', 'X_minmax', 'Y_minmax', 'nX_nY', 'w0'}


def make_bin_dict(seed=0, nX=12, nY=10, B=9):
    rng = np.random.default_rng(seed)
    bm, num = synth.voronoi_apertures(nX, nY, B, rng)
    half_x, half_y = 3.0, 2.5
    xc = np.linspace(-half_x, half_x, nX + 1)
    yc = np.linspace(-half_y, half_y, nY + 1)
    Xr, Yr = np.meshgrid(0.5 * (xc[1:] + xc[:-1]), 0.5 * (yc[1:] + yc[:-1]))
    bd = {'X_minmax': np.array([-half_x, half_x]), 'Y_minmax': np.array([-half_y, half_y]),
          'nX_nY': np.array([nX, nY]), 'num_per_bin': num, 'total_bins': np.array(B),
          'bin_mapping': bm, 'surface_density': rng.uniform(1, 2, B),
          'V_mean': rng.normal(0, 50, B), 'V_sigma': rng.uniform(80, 120, B),
          'v0': rng.normal(0, 30, B), 's': rng.uniform(90, 110, B),
          'orientation': np.array([0.1, 1.1, 0.8]),
          'X_regular_grid': Xr, 'Y_regular_grid': Yr}
    for k in ('V_mean_err', 'V_sigma_err'):
        bd[k] = rng.uniform(2, 5, B)
    for k in (1, 2, 3, 4):
        bd['h%d' % k] = rng.normal(0, 0.03, B)
        bd['h%d_err' % k] = rng.uniform(0.01, 0.02, B)
    return bd


@pytest.fixture(scope="module")
def loaded(tmp_path_factory):
    d = tmp_path_factory.mktemp("bins")
    bd = make_bin_dict()
    D.write_bin_dict(str(d / "galaxy.npz"), bd)
    return bd, D.load_data_bootstrap(str(d) + "/", "galaxy.npz", N_BOOTSTRAP=7, n_samples=3000)


def test_keys_and_data(loaded):
    bd, dd = loaded
    assert set(dd) == REFERENCE_KEYS
    B = int(bd['total_bins'])
    assert dd['total_bins'] == B and isinstance(dd['total_bins'], int)
    assert np.array_equal(dd['XY_density_data_err'], 0.01 * bd['surface_density'] + 1e-12)
    for ours, theirs in (('V_data', 'V_mean'), ('sigma_data', 'V_sigma'),
                         ('V_data_err', 'V_mean_err'), ('sigma_data_err', 'V_sigma_err'),
                         ('h3_data', 'h3'), ('h4_data_err', 'h4_err'), ('v0', 'v0')):
        assert np.array_equal(dd[ours], bd[theirs]), ours
    assert np.array_equal(dd['bin_mapping'], bd['bin_mapping'])
    assert dd['bin_mapping'].shape == (12 * 10 + 1,) and dd['bin_mapping'][-1] == -1


def test_bootstrap_normals(loaded):
    """One default_rng(42) stream in the order XY, h1..h4, V, sigma, row 0
    zeroed (utils.py:604-618)."""
    bd, dd = loaded
    B = int(bd['total_bins'])
    rng = np.random.default_rng(42)
    for nm in ('XY', 'h1', 'h2', 'h3', 'h4', 'V', 'sigma'):
        a = rng.normal(size=(7, B))
        got = dd[nm + '_standard_normal']
        assert got.shape == (7, B) and np.all(got[0] == 0.0)
        assert np.array_equal(got[1:], a[1:]), nm


def test_cells(loaded):
    """600 cell centres in cell-id order: each centre falls in its own cell,
    the centres are the mid-points of the 10 x 6 x 10 grid over R in [0, 10],
    z in [-3, 3], phi in [-pi, pi], and R varies fastest."""
    _, dd = loaded
    g = np.stack([dd['R_grid'], dd['z_grid'], dd['phi_grid']], axis=-1)
    assert g.shape == (600, 3) and dd['Rzphi_n_tot'] == 600
    assert dd['Rzphi_n_grid'].tolist() == [10, 6, 10]
    ids = D.assign_regular_grid(g, [0, -3, -np.pi], [10, 3, np.pi], [10, 6, 10])
    assert np.array_equal(ids, np.arange(600))
    assert np.array_equal(np.unique(dd['R_grid']), np.arange(10) + 0.5)
    assert np.array_equal(np.unique(dd['z_grid']), np.arange(-3, 3) + 0.5)
    assert np.allclose(np.unique(dd['phi_grid']), -np.pi + (np.arange(10) + 0.5) * np.pi / 5,
                       rtol=0, atol=1e-15)
    assert np.array_equal(dd['R_grid'][:10], np.arange(10) + 0.5)
    assert (dd['dR'], dd['dz']) == (1.0, 1.0) and abs(dd['dphi'] - np.pi / 5) < 1e-15
    assert dd['dX'] == pytest.approx(0.5) and dd['dY'] == pytest.approx(0.5)


def test_assign_regular_grid_brute_force():
    """Vectorised cell ids against a per-point loop of the rule: out of range
    (id = prod(n)) when any normalised coordinate is < 0 or >= 1, else floor
    then clip, id = i + j nx + k nx ny; edges included."""
    rng = np.random.default_rng(3)
    lo, hi, n = np.array([0.0, -3.0, -np.pi]), np.array([10.0, 3.0, np.pi]), np.array([4, 3, 5])
    pts = rng.uniform(lo - 1, hi + 1, size=(400, 3))
    pts[:6] = [lo, hi, [10.0, 0.0, 0.0], [0.0, -3.0, np.pi], [9.999999, 2.9999, 3.14159],
               [5.0, 3.0, 0.0]]
    got = D.assign_regular_grid(pts, lo, hi, n)
    for p, g in zip(pts, got):
        q = [(p[k] - lo[k]) / (hi[k] - lo[k]) for k in range(3)]
        if any(v < 0 or v >= 1 for v in q):
            assert g == 60
            continue
        c = [min(max(int(math.floor(q[k] * n[k])), 0), n[k] - 1) for k in range(3)]
        assert g == c[0] + c[1] * 4 + c[2] * 12
    assert got[0] == 0 and got[1] == 60 and got[2] == 60 and got[3] == 60


def test_sobol(loaded):
    """Unscrambled 3-d Sobol, 2^10 points: the first eight are the textbook
    ones, and every 1-d projection is {k / 1024} (each coordinate stratified)."""
    _, dd = loaded
    for key in ('sample_for_integration', 'sample_for_integration_XY'):
        s = dd[key]
        assert s.shape == (1024, 3)
        assert s[:8].tolist() == [[0, 0, 0], [.5, .5, .5], [.75, .25, .25], [.25, .75, .75],
                                  [.375, .375, .625], [.875, .875, .125], [.625, .125, .875],
                                  [.125, .625, .375]]
        for k in range(3):
            assert np.array_equal(np.sort(s[:, k]), np.arange(1024) / 1024)


def test_initial_positions(loaded):
    """Azimuths from a fresh default_rng(42) on [0, 2 pi); radii and heights by
    inverse CDF at JAX-keyed uniforms (keys 10086, 10010) within one grid step
    of the closed-form truncated x e^{-x/3} and e^{-z/1.2} quantiles."""
    _, dd = loaded
    w0 = dd['w0']
    assert w0.shape == (3000, 3) and w0.dtype == np.float32
    x, y, z = (w0[:, k].astype(np.float64) for k in range(3))
    phi = np.random.default_rng(42).uniform(0, 2 * np.pi, size=3000)
    dphi = np.angle(np.exp(1j * (np.arctan2(y, x) - phi)))
    R = np.hypot(x, y)
    assert np.abs(dphi[R > 0.05]).max() < 1e-5
    u = P.uniform(P.PRNGKey(10086), (3000,)).astype(np.float64)
    xx = np.linspace(0, 12, 400001)
    F = 1 - (1 + xx / 3) * np.exp(-xx / 3)
    assert np.abs(R - np.interp(u, F / F[-1], xx)).max() <= 1.01 * 12 / 999 + 1e-5
    uz = P.uniform(P.PRNGKey(10010), (3000,)).astype(np.float64)
    zz = np.linspace(0, 5, 400001)
    G = 1 - np.exp(-zz / 1.2)
    assert np.abs(z - np.interp(uz, G / G[-1], zz)).max() <= 1.01 * 5 / 999
    assert z.min() >= 0 and z.max() <= 5


def test_dict_data_round_trip(loaded, tmp_path):
    _, dd = loaded
    f = str(tmp_path / "dd.npz")
    D.save_dict_data(f, dd)
    back = D.load_dict_data(f)
    assert set(back) == set(dd)
    for k, v in dd.items():
        if isinstance(v, list):
            assert back[k] == v and isinstance(back[k], list), k
        elif isinstance(v, (int, float)):
            assert back[k] == v and type(back[k]) is type(v), k
        else:
            assert np.array_equal(back[k], v) and back[k].dtype == np.asarray(v).dtype, k


def test_pickles_refused(tmp_path):
    """A pickle path is refused before anything opens it; a missing key and
    an archive with object arrays are refused too."""
    with pytest.raises(ValueError, match="pickle"):
        D.load_data_bootstrap(str(tmp_path) + "/", "does_not_exist.pkl")
    bd = make_bin_dict()
    del bd['orientation']
    with pytest.raises(KeyError):
        D.write_bin_dict(str(tmp_path / "x.npz"), bd)
    with pytest.raises(TypeError):
        D.save_dict_data(str(tmp_path / "y.npz"), {'a': np.array([{'b': 1}], dtype=object)})
    np.savez(str(tmp_path / "z.npz"), a=np.zeros(3))
    with pytest.raises(ValueError):
        D.load_dict_data(str(tmp_path / "z.npz"))


def test_loader_output_runs_through_oracle(loaded):
    """A loader-made dict_data (with the reference's own PRNG draws) drives the
    oracle's model_likelihood to a finite value; its cell grid and field of
    view are what likelihood.grid_args hands build_model."""
    from schwarmax_amd import likelihood as L
    from oracle import model as OM
    bd, dd = loaded
    sub = dict(dd)
    sub['w0'] = dd['w0'][:24]
    nrm, uni = P.reference_draws(24, 4)
    sub['jeans_normals'], sub['jitter_uniforms'] = nrm, uni
    Rl, Rn, xyl, xyn = L.grid_args(sub)
    assert Rn.tolist() == [10, 6, 10] and xyn.tolist() == [12, 10]
    theta = synth.base_theta()
    sched = dict(N_max=100, N_dynamical_time=2)
    v = OM.model_likelihood(theta, sub, int(bd['total_bins']), 600, schedule=sched,
                            nnls_maxiter=20)
    assert np.isfinite(v)
