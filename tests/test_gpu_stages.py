"""Stage-by-stage parity of the HIP path against the FP64 oracle (GPU)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _pts(seed, n):
    rng = np.random.default_rng(seed)
    p = rng.normal(scale=[4.0, 4.0, 1.0], size=(n, 3))
    p[:5] = [[0.3, 0.0, 0.0], [8.0, 0.1, 0.0], [-3.0, 2.0, 0.5], [20.0, -15.0, 4.0], [0.01, 0.02, 0.0]]
    return p


def test_potential_and_force(torch_cuda, canon_params):
    from schwarmax_amd import potentials as POT
    from oracle import potential as OP
    from oracle.model import baryon_params
    _, ph, pb = canon_params
    p = _pts(1, 4000)
    phi, acc = POT.potential_eval(p[:, 0], p[:, 1], p[:, 2], baryon_params(pb), ph, with_accel=True)
    ax, ay, az, ophi = OP.accel(p[:, 0], p[:, 1], p[:, 2], baryon_params(pb), ph)
    phi = phi.cpu().numpy()
    acc = acc.cpu().numpy()
    assert np.max(np.abs(phi - ophi) / np.abs(ophi)) < 1e-13
    oacc = np.stack([ax, ay, az], axis=-1)
    scale = np.linalg.norm(oacc, axis=1)
    assert np.max(np.abs(acc - oacc).max(axis=1) / scale) < 1e-11


def test_density(torch_cuda, canon_params):
    from schwarmax_amd import potentials as POT
    from oracle import potential as OP
    _, ph, pb = canon_params
    p = _pts(2, 4000)
    rho = POT.density_eval(p[:, 0], p[:, 1], p[:, 2], pb).cpu().numpy()
    orho = OP.density_func(p[:, 0], p[:, 1], p[:, 2], pb)
    assert np.max(np.abs(rho - orho) / np.abs(orho)) < 1e-12


def _geometry(seed=0, n=32, n_ap=24, grid=(8, 4, 6), K=4, nxy=16, r=4, N_max=None):
    from schwarmax_amd import synth
    theta, dd, sched = synth.make_galaxy('plain-2k', seed=seed, n=n, n_ap=n_ap, grid=grid, K=K,
                                         nxy=nxy, r=r, N_max=N_max)
    return theta, dd, sched


def test_initial_conditions(torch_cuda):
    from schwarmax_amd import model as M, potentials as POT, likelihood as L
    from oracle import model as OM
    theta, dd, sched = _geometry(n=40)
    ph, pb = L.theta_to_params(theta)
    s = M.schedule_of(sched)
    pot = POT.pack_potential(M.baryon_params(pb), ph)
    wb, Tt, dt0 = M.initial_conditions(pot, dd, 4, s)
    owb, oT, odt = OM.initial_conditions(dd, OM.baryon_params(pb), ph, 4, s['N_step_per_orb'],
                                         s['N_dynamical_time'])
    wb = wb.cpu().numpy()
    # velocities come from FD of trapezoid integrals: agreement at the 1e-9 level
    assert np.max(np.abs(wb - owb)) / np.max(np.abs(owb)) < 1e-9
    assert np.max(np.abs(Tt.cpu().numpy() - oT) / oT) < 1e-10
    assert np.max(np.abs(dt0.cpu().numpy() - odt) / odt) < 1e-10


def _grids_both(dd, pb, symmetry):
    from schwarmax_amd import model as M
    from oracle import integrate as OI, model as OM
    from schwarmax_amd.likelihood import grid_args
    Rl, Rn, xyl, xyn = grid_args(dd)
    rot = OM.make_rotation_matrix(pb['alpha'], pb['beta'], pb['gamma'])
    g = M.make_grids(int(dd['total_bins']), Rn, Rl, xyl, xyn, rot, symmetry)
    og = OI.Grids(int(dd['total_bins']), dd['bin_mapping'], dd['num_per_bin'], Rl, xyl, Rn, xyn,
                  dd['v0'], dd['s'], rot)
    return g, og


@pytest.mark.parametrize("symmetry", [1, 0])
def test_binning_bit_exact(torch_cuda, symmetry):
    """Cell / aperture ids bit-identical to the oracle; unit weights give
    integer counts that must match exactly."""
    import ctypes
    import torch
    from schwarmax_amd import _lib, likelihood as L
    from oracle import integrate as OI
    theta, dd, _ = _geometry(n=8, n_ap=40, nxy=20)
    ph, pb = L.theta_to_params(theta)
    rng = np.random.default_rng(5)
    S = 20000
    smp = np.concatenate([rng.normal(scale=[3.5, 3.5, 1.2], size=(S, 3)),
                          rng.normal(scale=150.0, size=(S, 3))], axis=1)
    # exact cell edges and the field boundary as hard cases
    smp[:4, 0] = [0.0, 10.0, -10.0, 5.0]
    smp[:4, 1] = [0.0, 0.0, 0.0, 0.0]
    smp[4:8, 2] = [3.0, -3.0, 0.0, 1.5]
    g, og = _grids_both(dd, pb, symmetry)
    nsym = 4 if symmetry else 1
    ones = np.ones(S)
    d = lambda a, t=torch.float64: _lib.dev(a, t)  # noqa: E731
    P = g.nR * g.nz * g.nphi
    B = g.num_vbin
    idx_rz = torch.empty((S, nsym), dtype=torch.int32, device='cuda')
    idx_ap = torch.empty((S, nsym), dtype=torch.int32, device='cuda')
    acc_rz = torch.empty(P, dtype=torch.float64, device='cuda')
    acc_ap = torch.empty((B, 6), dtype=torch.float64, device='cuda')
    smp_d, ones_d = d(smp), d(ones)
    bm, v0, sd = d(dd['bin_mapping'], torch.int32), d(dd['v0']), d(dd['s'])
    _lib.check(_lib.lib().smx_bin_samples(g, _lib.ptr(smp_d), _lib.ptr(ones_d), S, _lib.ptr(bm),
                                          _lib.ptr(v0), _lib.ptr(sd), _lib.ptr(idx_rz), _lib.ptr(idx_ap),
                                          _lib.ptr(acc_rz), _lib.ptr(acc_ap), _lib.stream_ptr()),
               'smx_bin_samples')
    signs = OI.SIGN_SYM[:nsym]
    ys = (signs[:, :, None] * smp.T[None]).transpose(1, 2, 0).reshape(6, -1)   # sample-major
    irz, vb, vlos = og.indices(ys)
    irz = np.where(irz >= P, -1, irz)
    vb = np.where((vb < 0) | (vb >= B), -1, vb)
    assert np.array_equal(idx_rz.cpu().numpy().reshape(-1), irz)
    assert np.array_equal(idx_ap.cpu().numpy().reshape(-1), vb)
    cnt_rz = np.bincount(irz[irz >= 0], minlength=P).astype(np.float64)
    cnt_ap = np.bincount(vb[vb >= 0], minlength=B).astype(np.float64)
    assert np.array_equal(acc_rz.cpu().numpy(), cnt_rz)
    assert np.array_equal(acc_ap.cpu().numpy()[:, 0], cnt_ap)
    gw, H = og.gh_terms(vb, vlos)
    for k in range(5):
        ref = np.bincount(vb[vb >= 0], weights=(gw * H[k])[vb >= 0], minlength=B)
        got = acc_ap.cpu().numpy()[:, 1 + k]
        assert np.allclose(got, ref, rtol=1e-11, atol=1e-14 * np.abs(ref).max())


def test_library_matches_oracle(torch_cuda):
    """Library columns from identical initial conditions, on a schedule whose
    integrations pass the validity gate (the oracle's library is checked
    non-trivial first)."""
    from schwarmax_amd import model as M, potentials as POT, likelihood as L
    from oracle import integrate as OI, model as OM
    theta, dd, sched = _geometry(n=24, n_ap=24, nxy=16, N_max=600)
    ph, pb = L.theta_to_params(theta)
    s = M.schedule_of(sched)
    pot = POT.pack_potential(M.baryon_params(pb), ph)
    wb, Tt, dt0 = M.initial_conditions(pot, dd, 4, s)
    g, og = _grids_both(dd, pb, 1)
    libd = M.build_library(pot, g, M.make_schedule(s, -pb['Omega_bar']), wb, Tt, dt0, dd)
    acc_fn, pot_fn = OI.make_acc_pot(OM.baryon_params(pb), ph)
    per = OI.integrate_chunked(wb.cpu().numpy().reshape(-1, 6), acc_fn, pot_fn, s['N_max'],
                               Tt.cpu().numpy().reshape(-1), dt0.cpu().numpy().reshape(-1),
                               -pb['Omega_bar'], s['atol'], s['rtol'], s['dt_min'], s['dt_max'], og,
                               chunk_size=s['chunk'])
    lib_o = OI.realization_average(per, wb.shape[0], 4)
    frac = OM.library_guard(dict(lib_o, per=per))
    print(f"library: {frac:.3f} of the integrations valid", flush=True)
    P, B = libd['P'], libd['B']
    got = libd['lib'].cpu().numpy()
    ref = np.concatenate([lib_o['Rzphi'], lib_o['surface_density'], lib_o['h1'], lib_o['h2'],
                          lib_o['h3'], lib_o['h4'], lib_o['h0']], axis=1)
    assert np.array_equal(libd['valid'].cpu().numpy(), per['valid'])
    assert np.max(np.abs(libd['t_final'].cpu().numpy() - per['t_final']) / per['t_final']) < 1e-9
    col_err = np.abs(got - ref).max(axis=1) / np.maximum(np.abs(ref).max(axis=1), 1e-300)
    print(f"library columns: 90% {np.quantile(col_err, 0.9):.2e} max {col_err.max():.2e}")
    assert np.quantile(col_err, 0.9) < 1e-8, np.sort(col_err)[-5:]
    # the block structure: P density cells, then surface density, h1..h4, h0
    for blk in range(7):
        lo, hi = (0, P) if blk == 0 else (P + (blk - 1) * B, P + blk * B)
        assert np.any(ref[:, lo:hi] != 0.0), blk
