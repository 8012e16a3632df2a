"""Gradient parity (GPU): the adjoint kernels against central differences of
the FP64 oracle with the discrete decisions held fixed (the derivative
jax.grad takes of likelihood.py:96-244; see oracle.model.model_likelihood_grad_fd).

Tolerances: the library and IC derivatives agree to 1e-6 relative (fourth-
order differences of smooth FP64 maps); the full 12-component gradient to
1e-4 of its largest component, the oracle's own accuracy (its logL carries
~1e-10 relative roundoff from the nested finite differences of the Jeans
moments, jeans.py:37-60, which a difference step of 5e-5 turns into ~1e-5).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NAP = 16


def _tiny_galaxy(seed=5, n=12, N=100):
    from galaxies import observed_galaxy
    theta, dd, sched, _, _ = observed_galaxy(seed=seed, n=n, n_ap=NAP, grid=(6, 4, 4), K=3, nxy=12,
                                             N_max=N)
    return theta, dd, sched


@pytest.fixture(scope="module")
def galaxy():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return _tiny_galaxy()


def _d4(f, h):
    return (8.0 * (f(h) - f(-h)) - (f(2 * h) - f(-2 * h))) / (12.0 * h)


def _field_jacobian(theta):
    """d(fields, Omega_bar, rot row 2)/dtheta by torch autograd on the host map."""
    import torch
    from schwarmax_amd import grad as GR

    def f(t):
        fields, Om, rot, lm = GR.theta_fields(t)
        return torch.cat([fields, Om.reshape(1), rot[2]])
    return torch.autograd.functional.jacobian(f, torch.tensor(theta, dtype=torch.float64)).numpy()


def _oracle_library(theta, dd, sched, wb, Tt, dt0, tape):
    """Library (n, P + 6B) from given ICs by the oracle, rows [Rzphi|xy|h1..h4|h0]."""
    from schwarmax_amd import likelihood as L
    from oracle import model as OM, integrate as OI
    ph, pb = OM.theta_to_params(theta)
    s = OM.schedule_of(sched)
    Rl, Rn, xyl, xyn = L.grid_args(dd)
    rot = OM.make_rotation_matrix(pb['alpha'], pb['beta'], pb['gamma'])
    grids = OI.Grids(NAP, dd['bin_mapping'], dd['num_per_bin'], Rl, xyl, Rn, xyn, dd['v0'], dd['s'],
                     rot)
    acc_fn, pot_fn = OI.make_acc_pot(OM.baryon_params(pb), ph)
    n, r = wb.shape[:2]
    per = OI.integrate_chunked(wb.reshape(-1, 6), acc_fn, pot_fn, s['N_max'], Tt.reshape(-1),
                               dt0.reshape(-1), -pb['Omega_bar'], s['atol'], s['rtol'], s['dt_min'],
                               s['dt_max'], grids, chunk_size=s['chunk'], tape=tape)
    lib = OI.realization_average(per, n, r)
    return np.concatenate([lib['Rzphi'], lib['surface_density'], lib['h1'], lib['h2'], lib['h3'],
                           lib['h4'], lib['h0']], axis=1)


def _gpu_library_vjp(theta, dd, sched, wb, Tt, dt0, gbar, every=16):
    """smx_build_library_ckpt + smx_library_vjp on given ICs."""
    import torch
    from schwarmax_amd import _lib, likelihood as L, model as M, potentials as POT
    from schwarmax_amd import grad as GR
    ph, pb = L.theta_to_params(theta)
    s = M.schedule_of(sched)
    Rl, Rn, xyl, xyn = L.grid_args(dd)
    rot = M.makeRotationMatrix(pb['alpha'], pb['beta'], pb['gamma'])
    grids = M.make_grids(NAP, Rn, Rl, xyl, xyn, rot, True)
    pot = POT.pack_potential(M.baryon_params(pb), ph)
    fields, Om, rotT, _ = GR.theta_fields(torch.tensor(theta, dtype=torch.float64))
    wb_t = torch.tensor(wb, device='cuda')
    dt0_t = torch.tensor(dt0, device='cuda')
    Tt_t = torch.tensor(Tt, device='cuda')
    fields = fields.detach().requires_grad_(True)
    om = (-Om).detach().requires_grad_(True)
    r2 = rotT[2].detach().requires_grad_(True)
    wb_t.requires_grad_(True)
    dt0_t.requires_grad_(True)
    lib = GR._Library.apply(fields, om, r2, wb_t, dt0_t, Tt_t, pot, grids,
                            M.make_schedule(s, -pb['Omega_bar']), dd, every)
    g = torch.tensor(gbar, device='cuda')
    (lib * g).sum().backward()
    return (lib.detach().cpu().numpy(), wb_t.grad.cpu().numpy(), dt0_t.grad.cpu().numpy(),
            np.concatenate([fields.grad.numpy(), -om.grad.numpy().reshape(1), r2.grad.numpy()]))


@pytest.fixture(scope="module")
def ics(galaxy):
    from schwarmax_amd import likelihood as L
    from oracle import model as OM
    theta, dd, sched = galaxy
    ph, pb = L.theta_to_params(theta)
    s = OM.schedule_of(sched)
    wb, Tt, dt0 = OM.initial_conditions(dd, OM.baryon_params(pb), ph, 4, s['N_step_per_orb'],
                                        s['N_dynamical_time'])
    return wb, Tt, dt0


def test_library_vjp_state_directions(galaxy, ics):
    """<wb_bar, V> + <dt0_bar, u> equals the oracle's directional derivative of
    sum(gbar * lib) in (wb, dt0) with decisions fixed."""
    theta, dd, sched = galaxy
    wb, Tt, dt0 = ics
    rng = np.random.default_rng(3)
    lib0 = _oracle_library(theta, dd, sched, wb, Tt, dt0, None)
    gbar = rng.standard_normal(lib0.shape)
    lib_g, wb_bar, dt0_bar, _ = _gpu_library_vjp(theta, dd, sched, wb, Tt, dt0, gbar)
    # same decisions: the GPU library equals the oracle's
    assert np.max(np.abs(lib_g - lib0)) <= 1e-8 * np.max(np.abs(lib0)), 'trajectories diverged'
    tape = {'mode': 'record'}
    _oracle_library(theta, dd, sched, wb, Tt, dt0, tape)
    tape['mode'] = 'replay'
    # positions only, velocities only, dt0 only; tolerance relative to the
    # sum of |terms| of the directional derivative (the terms cancel)
    dirs = []
    for sel in ((0, 1, 2), (3, 4, 5)):
        V = np.zeros(wb.shape)
        V[..., sel] = rng.standard_normal(wb[..., sel].shape)
        dirs.append((V * (0.01 if sel[0] == 0 else 1.0), np.zeros(dt0.shape)))
    dirs.append((np.zeros(wb.shape), rng.standard_normal(dt0.shape) * 1e-3))
    for k, (V, u) in enumerate(dirs):
        # the replayed map is smooth in (wb, dt0): a large step keeps roundoff / h small
        fd = _d4(lambda h: np.sum(gbar * _oracle_library(theta, dd, sched, wb + h * V, Tt,
                                                          dt0 + h * u, tape)), 1e-3)
        ad = np.sum(wb_bar * V) + np.sum(dt0_bar * u)
        mag = np.sum(np.abs(wb_bar * V)) + np.sum(np.abs(dt0_bar * u))
        assert abs(ad - fd) <= 1e-6 * mag, (k, ad, fd, mag)


def test_library_vjp_parameter_directions(galaxy, ics):
    """d/dtheta_i of sum(gbar * lib) at fixed ICs: field / Omega / rotation
    cotangents chained through the host map, against oracle differences."""
    theta, dd, sched = galaxy
    wb, Tt, dt0 = ics
    rng = np.random.default_rng(4)
    lib0 = _oracle_library(theta, dd, sched, wb, Tt, dt0, None)
    gbar = rng.standard_normal(lib0.shape)
    _, _, _, pbar = _gpu_library_vjp(theta, dd, sched, wb, Tt, dt0, gbar)
    J = _field_jacobian(theta)            # (11 + 1 + 3, 12)
    ad = pbar @ J
    tape = {'mode': 'record'}
    _oracle_library(theta, dd, sched, wb, Tt, dt0, tape)
    tape['mode'] = 'replay'
    fd = np.zeros(12)
    for i in range(12):
        e = np.zeros(12)
        e[i] = 1.0
        fd[i] = _d4(lambda h: np.sum(gbar * _oracle_library(theta + h * e, dd, sched, wb, Tt, dt0,
                                                            tape)), 1e-5)
    err = np.abs(ad - fd) / np.max(np.abs(fd))
    assert np.all(err < 1e-6), np.stack([ad, fd, err])


def test_initial_conditions_jacobian(galaxy):
    """Jeans ICs and dt0: field Jacobian times d(fields)/dtheta_i against oracle
    differences, to 1e-5: the mean rotation v_phi^2 = sigma^2 + (R/nu) dnu/dR +
    R dPhi/dR cancels large terms of nested finite differences (jeans.py:37-79),
    so the oracle's ICs carry roundoff that a step of 1e-4 turns into ~1e-6."""
    import torch
    from schwarmax_amd import _lib, likelihood as L, model as M, potentials as POT
    from oracle import model as OM
    theta, dd, sched = galaxy
    ph, pb = L.theta_to_params(theta)
    s = M.schedule_of(sched)
    pot = POT.pack_potential(M.baryon_params(pb), ph)
    w0 = _lib.dev(np.asarray(dd['w0'])[:, :3])
    n = w0.shape[0]
    uni = _lib.dev(dd['jitter_uniforms'])
    jw = torch.empty((n, 4, 6, 11), dtype=torch.float64, device='cuda')
    jd = torch.empty((n, 4, 11), dtype=torch.float64, device='cuda')
    lib = _lib.lib()
    _lib.check(lib.smx_initial_conditions_jac(pot, _lib.ptr(w0), _lib.ptr(_lib.dev(dd['jeans_normals'])),
                                              _lib.ptr(uni), n, 4, uni.shape[1],
                                              float(s['N_step_per_orb']), _lib.ptr(jw), _lib.ptr(jd),
                                              _lib.stream_ptr()), 'ic_jac')
    J = _field_jacobian(theta)[:11]       # fields only
    jw = jw.cpu().numpy() @ J             # (n, 4, 6, 12)
    jd = jd.cpu().numpy() @ J
    h = 1e-4
    for i in (0, 1, 2, 3, 4, 5, 6):
        e = np.zeros(12)
        e[i] = h
        out = {}
        for k in (-2, -1, 1, 2):
            php, pbp = OM.theta_to_params(theta + k * e)
            out[k] = OM.initial_conditions(dd, OM.baryon_params(pbp), php, 4, s['N_step_per_orb'],
                                           s['N_dynamical_time'])
        fdw = (8.0 * (out[1][0] - out[-1][0]) - (out[2][0] - out[-2][0])) / (12.0 * h)
        fdd = (8.0 * (out[1][2] - out[-1][2]) - (out[2][2] - out[-2][2])) / (12.0 * h)
        sw = np.max(np.abs(fdw), axis=(0, 1)) + 1e-300
        ew = np.max(np.abs(jw[..., i] - fdw), axis=(0, 1))
        assert np.all(ew <= 1e-5 * sw + 1e-12), (i, ew, sw, np.max(np.abs(jw[..., i]), axis=(0, 1)))
        ed = np.max(np.abs(jd[..., i] - fdd))
        assert ed <= 1e-5 * np.max(np.abs(fdd)) + 1e-15, (i, ed, np.max(np.abs(fdd)))


def test_density_target_jacobian(galaxy):
    import torch
    from schwarmax_amd import _lib, likelihood as L, potentials as POT
    from oracle import model as OM
    theta, dd, sched = galaxy
    ph, pb = L.theta_to_params(theta)
    pot = POT.pack_potential(pb)
    Rg, zg, pg = (_lib.dev(dd[k]) for k in ('R_grid', 'z_grid', 'phi_grid'))
    smp = _lib.dev(dd['sample_for_integration'])
    P = Rg.numel()
    jac = torch.empty((P, 11), dtype=torch.float64, device='cuda')
    _lib.check(_lib.lib().smx_density_target_jac(pot, _lib.ptr(Rg), _lib.ptr(zg), _lib.ptr(pg), P,
                                                 float(dd['dR']), float(dd['dz']), float(dd['dphi']),
                                                 _lib.ptr(smp), smp.shape[0], _lib.ptr(jac),
                                                 _lib.stream_ptr()), 'density_jac')
    J = _field_jacobian(theta)[:11]
    ad = jac.cpu().numpy() @ J
    for i in (1, 2, 4, 5, 6):
        e = np.zeros(12)
        e[i] = 1.0
        fd = _d4(lambda h: OM.density_target(dd, OM.theta_to_params(theta + h * e)[1]), 1e-5)
        assert np.max(np.abs(ad[:, i] - fd)) <= 1e-6 * np.max(np.abs(fd)), i


def test_model_likelihood_gradient(galaxy):
    """Full 12-component gradient against the oracle's frozen-decision
    central differences, both with the GPU's NNLS weights."""
    import torch
    from schwarmax_amd import grad as GR, likelihood as L
    from oracle import model as OM
    theta, dd, sched = galaxy
    th = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    logl, aux = GR.loglike_graph(th, dd, NAP, dd['Rzphi_n_tot'], schedule=sched, nnls_maxiter=30,
                                 ckpt_every=16)
    logl.backward()
    g = th.grad.numpy()
    W = aux['W'].cpu().numpy()
    f0, fd, _ = OM.model_likelihood_grad_fd(theta, dd, NAP, dd['Rzphi_n_tot'], schedule=sched,
                                           nnls_maxiter=30, W_fixed=W)
    assert abs(float(logl) - f0) <= 1e-9 * abs(f0)
    # the plain path gives the same value
    plain = float(L.model_likelihood(theta, dd, NAP, dd['Rzphi_n_tot'], schedule=sched,
                                     nnls_maxiter=30))
    assert abs(plain - float(logl)) <= 1e-12 * abs(plain)
    err = np.abs(g - fd) / np.max(np.abs(fd))
    assert np.all(err < 1e-4), np.stack([g, fd, err])


def test_gradient_batch_matches_per_model(galaxy):
    """The batched forward + adjoint returns, per theta and orbit count, the
    single-model logL and gradient, and both agree with the oracle: logL to
    1e-4 and the gradient to 1e-4 of its largest component against the
    oracle's frozen-decision central differences (with the oracle's own NNLS
    weights, whose difference from the GPU's is ~1e-5)."""
    from schwarmax_amd import grad as GR, synth
    from oracle import model as OM
    from galaxies import oracle_likelihood
    theta, dd, sched = galaxy
    rng = np.random.default_rng(5)
    thetas = np.stack([theta] + synth.batched_thetas(1, rng))
    n_all = np.asarray(dd['w0']).shape[0]
    n_orbits = [n_all, n_all - 7]
    v, g = GR.model_likelihood_and_grad_batch(thetas, dd, NAP, dd['Rzphi_n_tot'],
                                              n_orbits=n_orbits, schedule=sched, nnls_maxiter=20,
                                              ckpt_every=16)
    for m in range(2):
        dm = dict(dd)
        for k in ('w0', 'jeans_normals', 'jitter_uniforms'):
            dm[k] = np.asarray(dd[k])[:n_orbits[m]]
        v1, g1 = GR.model_likelihood_and_grad(thetas[m], dm, NAP, dd['Rzphi_n_tot'],
                                              schedule=sched, nnls_maxiter=20, ckpt_every=16)
        assert float(v[m]) == float(v1), m
        assert np.array_equal(g[m].numpy(), g1.numpy()), m
        frac = oracle_likelihood(thetas[m], dm, NAP, sched, 20)[3]
        f0, fd, _ = OM.model_likelihood_grad_fd(thetas[m], dm, NAP, dd['Rzphi_n_tot'],
                                                schedule=sched, nnls_maxiter=20)
        err = np.abs(g[m].numpy() - fd) / np.max(np.abs(fd))
        print(f"model {m}: valid {frac:.3f} logL rel {abs(float(v[m]) - f0) / abs(f0):.2e} "
              f"gradient max rel {err.max():.2e}", flush=True)
        assert abs(float(v[m]) - f0) <= 1e-4 * abs(f0), (m, float(v[m]), f0)
        assert np.all(err < 1e-4), (m, np.stack([g[m].numpy(), fd, err]))


def test_gradient_finite_where_jeans_moments_clamp():
    """A parameter vector of the batched forward + adjoint workload whose
    Jeans moments clamp (sqrt(max(x, 0)) at x <= 0) for some stars.  Forward
    mode must give those stars zero tangents there, as jax.grad's reverse
    mode does (maximum passes no cotangent to the clamped side); 0 * inf
    would turn the whole gradient into NaN."""
    import torch
    import bench
    from schwarmax_amd import grad as GR, synth
    theta0, dd, sched = bench.make_workload('batched-grad-16', 0, n=8000)
    thetas = np.stack(synth.batched_thetas(16, np.random.default_rng(1)))
    v, g = GR.model_likelihood_and_grad(thetas[14], dd, int(dd['total_bins']),
                                        int(dd['Rzphi_n_tot']), schedule=sched, nnls_maxiter=200)
    assert np.isfinite(float(v))
    assert torch.isfinite(g).all(), g
