"""The two forms of smx_build_library's binning give the same bits.

Form 1 (default): the integrator stores every attempt's state and
bin_library_kernel bins one orbit column per workgroup into LDS histograms.
Form 2: the fused integrator bins into per-integration slabs through its
register caches and finalize_library_kernel averages them.  Both add the same
64-bit fixed-point contributions, so library, validity flags, T and t_final
must be identical, with and without the fourfold symmetry, and for a grid
whose rows (cells + 6 x apertures) exceed one LDS window of 20,480, which
bin_library_kernel covers in two passes.  The checkpointed build for the
gradient (the binning kernel also writes each integration's slab for the
adjoint's T-cotangent pre-pass) gives the same logL and gradient bit for bit
in both forms.  Each case also checks that the
library is not trivial (most integrations valid, non-zero surface density and
h terms).  The default form against the FP64 oracle (integrates.py:115-238):
test_gpu_stages.py::test_library_matches_oracle and every library parity test.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(n, grid, symmetry, N_max=300, seed=5, n_ap=24):
    import torch
    from schwarmax_amd import synth, model as M, likelihood as L, potentials as POT, _lib
    theta, dd, sched = synth.make_galaxy('plain-2k', seed=seed, n=n, n_ap=n_ap, grid=grid, K=2,
                                         nxy=16, N_max=N_max)
    ph, pb = L.theta_to_params(theta)
    s = M.schedule_of(sched)
    pot = POT.pack_potential(M.baryon_params(pb), ph)
    wb, Tt, dt0 = M.initial_conditions(pot, dd, 4, s)
    Rl, Rn, xyl, xyn = L.grid_args(dd)
    rot = M.makeRotationMatrix(pb['alpha'], pb['beta'], pb['gamma'])
    g = M.make_grids(n_ap, Rn, Rl, xyl, xyn, rot, symmetry)
    sc = M.make_schedule(s, -pb['Omega_bar'])
    dev = lambda k: torch.as_tensor(np.asarray(dd[k], dtype=np.float64), device='cuda')  # noqa: E731
    bm = _lib.bin_map(dd['bin_mapping'], g.nX, g.nY)
    return (theta, dd, sched), (pot, g, sc, wb, Tt, dt0, bm, dev('num_per_bin'), dev('v0'), dev('s'))


def _build(args, form):
    import torch
    from schwarmax_amd import _lib, tune
    L = _lib.lib()
    pot, g, sc, wb, Tt, dt0, bm, npb, v0, sd = args
    n, r = wb.shape[0], wb.shape[1]
    P, B = g.nR * g.nz * g.nphi, g.num_vbin
    mk = lambda *shape: torch.full(shape, float('nan'), dtype=torch.float64, device='cuda')  # noqa: E731
    lib, valid, T_int, t_final = mk(n, P + 6 * B), mk(n * r), mk(n * r), mk(n * r)
    ws = torch.empty(int(L.smx_library_build_bytes(g, sc, n * r)), dtype=torch.uint8, device='cuda')
    f64 = torch.float64
    before = tune.get_knob(tune.KNOB_LIB_BINNING)
    tune.set_knob(tune.KNOB_LIB_BINNING, form)
    try:
        print(f"binning form {form}: n={n} P={P} B={B} symmetry={g.symmetry}", flush=True)
        _lib.check(L.smx_build_library(pot, g, sc, _lib.ptr(wb, f64), _lib.ptr(Tt, f64),
                                       _lib.ptr(dt0, f64), n, r, _lib.ptr(bm, torch.int32),
                                       _lib.ptr(npb, f64), _lib.ptr(v0, f64), _lib.ptr(sd, f64),
                                       _lib.ptr(ws), ws.numel(), _lib.ptr(lib, f64),
                                       _lib.ptr(valid, f64),
                                       _lib.ptr(T_int, f64), _lib.ptr(t_final, f64),
                                       _lib.stream_ptr()), 'smx_build_library')
        torch.cuda.synchronize()
    finally:
        tune.set_knob(tune.KNOB_LIB_BINNING, before)
    return [t.cpu().numpy() for t in (lib, valid, T_int, t_final)], P, B


def _compare(args):
    a, P, B = _build(args, 1)
    b, _, _ = _build(args, 2)
    lib, valid = a[0], a[1]
    assert valid.mean() >= 0.8, valid.mean()
    assert np.all(np.isfinite(lib))
    assert np.any(lib[:, P:P + B] != 0.0) and np.any(lib[:, P + B:P + 5 * B] != 0.0)
    for name, x, y in zip(('lib', 'valid', 'T', 't_final'), a, b):
        assert np.array_equal(x, y), (name, np.nanmax(np.abs(x - y)))
    return a


@pytest.mark.parametrize("symmetry", [True, False])
def test_sample_binning_equals_fused(symmetry):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    _, args = _case(40, (8, 4, 6), symmetry)
    _compare(args)


def test_sample_binning_two_windows():
    """25,600 cells + 6 x 24 aperture rows: two LDS windows."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    _, args = _case(16, (40, 32, 20), True, N_max=200)
    g = args[1]
    assert g.nR * g.nz * g.nphi + 6 * g.num_vbin > 20480
    _compare(args)


@pytest.mark.parametrize("grid,n", [((8, 4, 6), 40), ((40, 32, 20), 16)])
def test_gradient_equal_in_both_forms(grid, n):
    """One LDS window, and two (25,600 cells): the slabs the binning kernel
    writes for the adjoint equal the fused integrator's."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from schwarmax_amd import grad as GR, tune
    from galaxies import observed_galaxy
    theta, dd, sched = observed_galaxy(seed=11, n=n, n_ap=24, grid=grid, N_max=500)[:3]
    before = tune.get_knob(tune.KNOB_LIB_BINNING)
    out = {}
    try:
        for form in (1, 2):
            tune.set_knob(tune.KNOB_LIB_BINNING, form)
            print(f"gradient, binning form {form}", flush=True)
            v, g = GR.model_likelihood_and_grad(theta, dd, 24, dd['Rzphi_n_tot'], schedule=sched,
                                                nnls_maxiter=30, ckpt_every=16)
            out[form] = (float(v), g.numpy())
    finally:
        tune.set_knob(tune.KNOB_LIB_BINNING, before)
    assert np.isfinite(out[1][0]) and np.all(np.isfinite(out[1][1]))
    assert np.any(out[1][1] != 0.0)
    assert out[1][0] == out[2][0]
    assert np.array_equal(out[1][1], out[2][1])
