"""Golden vectors for the CPU-runnable configuration at full size (plain-2k):
n = 2,000 orbit columns x 4 realizations, 1,000 BS3 attempts at rtol 1e-6
(horizon 10 orbital times), 300 Voronoi apertures on a 40" x 40" field, a
20 x 10 x 16 (R, z, phi) grid, K = 1, T = 500 ADMM iterations.

Runs the FP64 oracle (oracle/) on the synthetic galaxy of seed 0 and writes
config1_golden.npz beside this script: the fitted (v0, s), the observations
made from the oracle's library (synth.observe, seed 1), the validity flags,
per-orbit block sums of the library, the NNLS weights, the model maps and
logL.  tests/test_gpu_config1.py rebuilds the same galaxy (synth is seeded
NumPy) and checks the device path against these values.

    python tests/golden/make_config1.py        # about 10 minutes on 8 cores
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

OBS_KEYS = ('XY_density_data', 'XY_density_data_err', 'h1_data', 'h1_data_err', 'h2_data',
            'h2_data_err', 'h3_data', 'h3_data_err', 'h4_data', 'h4_data_err', 'V_data',
            'V_data_err', 'sigma_data', 'sigma_data_err', 'v0', 's')
MAP_KEYS = ('density_3d', 'density_3d_model', 'surface_density', 'h1', 'h2', 'h3', 'h4', 'V',
            'sigma', 'weights')
SEED, T = 0, 500


def galaxy():
    from schwarmax_amd import synth
    return synth.make_galaxy('plain-2k', seed=SEED)


def main(out=os.path.join(HERE, 'config1_golden.npz')):
    from schwarmax_amd import synth
    from oracle import model as OM, potential as OP
    t0 = time.time()
    theta, dd, sched = galaxy()
    ph, pb = OM.theta_to_params(theta)
    s = OM.schedule_of(sched)
    wb, _, _ = OM.initial_conditions(dd, OM.baryon_params(pb), ph, 1, s['N_step_per_orb'],
                                     s['N_dynamical_time'])
    rot = OM.make_rotation_matrix(pb['alpha'], pb['beta'], pb['gamma'])
    synth.fit_base_gaussians(dd, wb[:, 0, :], rot)
    B = int(dd['total_bins'])
    Rl, Rn, xyl, xyn = OM.grid_args(dd)
    blank = {'XY_density_data': np.ones(B), 'XY_density_data_err': np.full(B, 0.01)}
    for k in (1, 2, 3, 4):
        blank['h%d_data' % k] = np.zeros(B)
        blank['h%d_data_err' % k] = np.full(B, 0.03)
    _, lib = OM.build_model(OP.density_func, OP.potential_func, ph, pb, {**dd, **blank}, B,
                            dd['Rzphi_n_tot'], Rn, Rl, xyl, xyn, nnls_maxiter=1, schedule=sched,
                            return_library=True)
    frac = OM.library_guard(lib)
    print(f"library: {time.time() - t0:.0f} s, {frac:.4f} of the integrations valid", flush=True)
    y_R = OM.density_target(dd, pb)
    synth.observe(dd, lib['surface_density'].T, [lib['h%d' % k].T for k in (1, 2, 3, 4)],
                  y_R.sum() / lib['surface_density'].shape[0], pb['mass_to_light_ratio'], floor=1e-3)
    md, _ = OM.model_from_library(lib, dd, pb, T, 1.0)
    logl = float(OM.bootstrap_loglike(md, dd)[0])
    print(f"logL {logl:.12g}, {time.time() - t0:.0f} s", flush=True)
    blocks = [lib['Rzphi'], lib['surface_density'], lib['h1'], lib['h2'], lib['h3'], lib['h4'],
              lib['h0']]
    np.savez_compressed(
        out, seed=SEED, T=T, logL=logl, valid=lib['per']['valid'].astype(np.uint8),
        t_final=lib['per']['t_final'], block_sums=np.stack([b.sum(axis=1) for b in blocks], axis=1),
        **{'obs_' + k: np.asarray(dd[k]) for k in OBS_KEYS},
        **{'model_' + k: np.asarray(md[k]) for k in MAP_KEYS})
    print("wrote", out, flush=True)


if __name__ == '__main__':
    main()
