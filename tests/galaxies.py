"""Small synthetic galaxies whose orbit libraries make the parity tests bite.

Every fixture is built on the CPU-runnable configuration's geometry with a
short schedule whose horizon follows the reference's rule N_dynamical_time =
N_max / N_step_per_orb (model.py:163-165), so that the integrations pass the
validity gate (integrates.py:210).  Before any observation is made, the
oracle's library is checked with ``oracle.model.library_guard``: at least 80%
valid integrations and non-zero density, surface density and h1..h4 rows.
Observations then come from that library and a Dirichlet weight vector
(``synth.observe``).
"""
import numpy as np


def blank_obs(B):
    """Placeholder data for the library-only oracle call (the library does not
    depend on the data)."""
    d = {'XY_density_data': np.ones(B), 'XY_density_data_err': np.full(B, 0.01)}
    for k in (1, 2, 3, 4):
        d['h%d_data' % k] = np.zeros(B)
        d['h%d_data_err' % k] = np.full(B, 0.03)
    return d


def observed_galaxy(seed=11, n=40, n_ap=24, grid=(8, 4, 6), K=5, nxy=16, N_max=500, floor=1e-3,
                    min_valid=0.8, config='plain-2k', obs_seed=1):
    """(theta, dict_data, schedule, oracle library, valid fraction).

    The schedule is the configuration's (rtol 1e-6 for plain-2k) with N_max
    attempts and the horizon N_max / 100 orbital times."""
    from schwarmax_amd import synth
    from oracle import model as OM, potential as OP
    theta, dd, sched = synth.make_galaxy(config, seed=seed, n=n, n_ap=n_ap, grid=grid, K=K, nxy=nxy,
                                         N_max=N_max)
    ph, pb = OM.theta_to_params(theta)
    Rl, Rn, xyl, xyn = OM.grid_args(dd)
    _, lib = OM.build_model(OP.density_func, OP.potential_func, ph, pb, {**dd, **blank_obs(n_ap)},
                            n_ap, dd['Rzphi_n_tot'], Rn, Rl, xyl, xyn, nnls_maxiter=1,
                            schedule=sched, return_library=True)
    frac = OM.library_guard(lib, min_valid)
    A_xy = lib['surface_density'].T
    A_h = [lib['h%d' % k].T for k in (1, 2, 3, 4)]
    y_R = OM.density_target(dd, pb)
    synth.observe(dd, A_xy, A_h, y_R.sum() / A_xy.shape[1], pb['mass_to_light_ratio'],
                  seed=obs_seed, floor=floor)
    print(f"galaxy seed={seed} n={n} N_max={N_max}: {frac:.3f} of the integrations valid",
          flush=True)
    return theta, dd, sched, lib, frac


def rowwise_rel(got, ref):
    """max over rows of ||got - ref|| / ||ref|| (rows: bootstrap right-hand sides)."""
    got = np.atleast_2d(np.asarray(got, dtype=np.float64))
    ref = np.atleast_2d(np.asarray(ref, dtype=np.float64))
    num = np.linalg.norm(got - ref, axis=-1)
    den = np.maximum(np.linalg.norm(ref, axis=-1), 1e-300)
    return float(np.max(num / den))


def oracle_likelihood(theta, dd, n_ap, sched, nnls_maxiter, min_valid=0.8):
    """The oracle's model_likelihood (likelihood.py:96-244) with its library
    checked by ``library_guard``.  Returns (logL, model dict, library, valid
    fraction)."""
    from oracle import model as OM, potential as OP
    ph, pb = OM.theta_to_params(theta)
    Rl, Rn, xyl, xyn = OM.grid_args(dd)
    md, lib = OM.build_model(OP.density_func, OP.potential_func, ph, pb, dd, n_ap, dd['Rzphi_n_tot'],
                             Rn, Rl, xyl, xyn, nnls_maxiter=nnls_maxiter, schedule=sched,
                             return_library=True)
    frac = OM.library_guard(lib, min_valid)
    return OM.bootstrap_loglike(md, dd)[0], md, lib, frac


def library_rows(lib):
    """Oracle library as the device lays it out: (n, P + 6B) rows
    [Rzphi | surface density | h1 | h2 | h3 | h4 | h0]."""
    return np.concatenate([lib['Rzphi'], lib['surface_density'], lib['h1'], lib['h2'], lib['h3'],
                           lib['h4'], lib['h0']], axis=1)


def column_errors(got, ref):
    """Per-orbit (row) max relative error of two libraries."""
    return np.abs(got - ref).max(axis=1) / np.maximum(np.abs(ref).max(axis=1), 1e-300)
