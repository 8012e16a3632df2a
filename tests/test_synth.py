"""Workload generator sanity (CPU)."""
import numpy as np

from schwarmax_amd import synth


def test_geometry_layout():
    theta, dd, sched = synth.make_galaxy('plain-2k', seed=0)
    nX, nY = dd['nX_nY']
    bm = dd['bin_mapping']
    assert bm.shape == (nX * nY + 1,) and bm[-1] == -1
    assert dd['num_per_bin'].sum() == nX * nY and dd['num_per_bin'].min() >= 1
    assert len(dd['R_grid']) == dd['Rzphi_n_tot'] == 20 * 10 * 16
    assert np.allclose(dd['X_minmax'], [-4.848136811095359, 4.848136811095359])
    assert dd['XY_standard_normal'].shape == (1, 300) and np.all(dd['XY_standard_normal'][0] == 0)
    assert dd['jitter_uniforms'].shape == (2000, 4, 6)
    assert sched['N_max'] == 1000 and sched['rtol'] == 1e-6


def test_bootstrap_normals_match_loader_draws():
    """utils.py:604-618: default_rng(42), XY then h1..h4, V, sigma; row 0 zeroed."""
    dd = {}
    synth.bootstrap_normals(dd, 5, 7)
    rng = np.random.default_rng(42)
    a = rng.normal(size=(5, 7))
    b = rng.normal(size=(5, 7))
    a[0] = 0
    b[0] = 0
    assert np.array_equal(dd['XY_standard_normal'], a)
    assert np.array_equal(dd['h1_standard_normal'], b)


def test_base_theta_reproduces_configuration():
    th = synth.base_theta()
    assert abs(th[11] - np.log10(40 * 1.022712165045695)) < 1e-15
    assert abs(np.degrees(th[8]) - 65.0) < 1e-12
