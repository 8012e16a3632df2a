"""smx_build_library_noalloc (the orbit library without device allocation)
against smx_build_library on the same inputs: two calls in a row, the second
smaller, so that the shared scratch array's leftovers from the first call
would show.  The no-allocation entry point bins in the fused integrator's
slabs, smx_build_library stores samples and bins them in LDS histograms; both
hold the sums as 64-bit fixed point, so the results must be identical.
A call whose slabs exceed the scratch array is refused before any launch."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _inputs(n, N_max=200, seed=17):
    import torch
    from schwarmax_amd import synth, model as M, likelihood as L, potentials as POT
    theta, dd, sched = synth.make_galaxy('plain-2k', seed=seed, n=n, n_ap=24, grid=(8, 4, 6), K=2,
                                         nxy=16, N_max=N_max)
    ph, pb = L.theta_to_params(theta)
    s = M.schedule_of(sched)
    pot = POT.pack_potential(M.baryon_params(pb), ph)
    wb, Tt, dt0 = M.initial_conditions(pot, dd, 4, s)
    Rl, Rn, xyl, xyn = L.grid_args(dd)
    rot = M.makeRotationMatrix(pb['alpha'], pb['beta'], pb['gamma'])
    g = M.make_grids(24, Rn, Rl, xyl, xyn, rot, True)
    sc = M.make_schedule(s, -pb['Omega_bar'])
    dev = lambda k: torch.as_tensor(np.asarray(dd[k], dtype=np.float64), device='cuda')  # noqa: E731
    from schwarmax_amd import _lib
    bm = _lib.bin_map(dd['bin_mapping'], g.nX, g.nY)
    return pot, g, sc, wb, Tt, dt0, bm, dev('num_per_bin'), dev('v0'), dev('s')


def _outputs(n, r, P, B):
    import torch
    mk = lambda *shape: torch.full(shape, float('nan'), dtype=torch.float64, device='cuda')  # noqa: E731
    return mk(n, P + 6 * B), mk(n * r), mk(n * r), mk(n * r)


def _run(noalloc, args):
    import torch
    from schwarmax_amd import _lib
    L = _lib.lib()
    pot, g, sc, wb, Tt, dt0, bm, npb, v0, sd = args
    n, r = wb.shape[0], wb.shape[1]
    P, B = g.nR * g.nz * g.nphi, g.num_vbin
    lib, valid, T_int, t_final = _outputs(n, r, P, B)
    f64 = torch.float64
    common = (pot, g, sc, _lib.ptr(wb, f64), _lib.ptr(Tt, f64), _lib.ptr(dt0, f64), n, r,
              _lib.ptr(bm, torch.int32), _lib.ptr(npb, f64), _lib.ptr(v0, f64), _lib.ptr(sd, f64))
    outs = (_lib.ptr(lib, f64), _lib.ptr(valid, f64), _lib.ptr(T_int, f64), _lib.ptr(t_final, f64))
    if noalloc:
        rc = L.smx_build_library_noalloc(*common, *outs, _lib.stream_ptr())
        _lib.check(rc, 'smx_build_library_noalloc')
    else:
        ws = torch.empty(int(L.smx_library_build_bytes(g, sc, n * r)), dtype=torch.uint8,
                         device='cuda')
        _lib.check(L.smx_build_library(*common, _lib.ptr(ws), ws.numel(), *outs, _lib.stream_ptr()),
                   'smx_build_library')
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in (lib, valid, T_int, t_final)]


def test_noalloc_matches_workspace_entry_point():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    for n in (48, 20):                     # the second call is smaller
        args = _inputs(n)
        ref = _run(False, args)
        got = _run(True, args)
        print(f"n={n}: valid {ref[1].mean():.3f}, library max {np.abs(ref[0]).max():.3e}", flush=True)
        assert ref[1].mean() >= 0.8 and np.any(ref[0] != 0.0)
        for a, b, name in zip(got, ref, ('lib', 'valid', 'T_int', 't_final')):
            assert a.shape == b.shape, name
            assert np.all(np.isfinite(a)), name
            assert np.array_equal(a, b), (name, np.max(np.abs(a - b)))


def test_noalloc_rejects_oversized_calls():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from schwarmax_amd import _lib, model as M
    L = _lib.lib()
    g = M.make_grids(1000, (32, 16, 24), [[0, 10], [-3, 3], [-np.pi, np.pi]], [[-5, 5], [-5, 5]],
                     (60, 60), np.eye(3), True)
    n, r = 20000, 4                         # 12.9 GB of slabs
    assert L.smx_library_workspace_bytes(g, n * r) > (64 << 20)
    sc = _lib.SmxSchedule()
    sc.N_max = 10
    pot = _lib.SmxPotential()
    dummy = ctypes.c_void_p(1)              # never dereferenced: the size check comes first
    rc = L.smx_build_library_noalloc(pot, g, sc, dummy, dummy, dummy, n, r, dummy, dummy, dummy,
                                     dummy, dummy, dummy, dummy, dummy, _lib.stream_ptr())
    assert rc == 2
    assert 'scratch array' in _lib.last_error()
