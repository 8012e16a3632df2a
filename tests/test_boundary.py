"""The C-ABI library loads and exports every symbol include/smx.h declares (CPU)."""
import ctypes
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, 'include', 'smx.h')).read()
    return sorted(set(re.findall(r'\b(smx_[a-z0-9_]+)\s*\(', txt)))


def test_library_exports_every_declared_symbol():
    from schwarmax_amd import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = _declared()
    assert len(names) >= 15
    for nm in names:
        assert hasattr(lib, nm), nm
    assert set(names) == set(_lib.EXPORTED)


def test_struct_layouts_match_header():
    """ctypes structs have the header's field order and sizes."""
    from schwarmax_amd import _lib
    assert ctypes.sizeof(_lib.SmxPotential) == 8 * (5 + 3 + 9 + 4 + 3 + 9 + 5 + 5 + 1)
    assert ctypes.sizeof(_lib.SmxGrids) == 4 * 8 + 8 * (3 + 3 + 2 + 2 + 9 + 1)
    assert ctypes.sizeof(_lib.SmxSchedule) == 8 + 8 * 5
    assert ctypes.sizeof(_lib.SmxNnlsProblem) == 8 * 8
    assert _lib.SmxNnlsProblem.w_init.offset == 5 * 8


def test_split_gemm_refuses_bad_arguments_without_gpu():
    """smx_gemm_nt_f64_split checks its arguments before touching the device:
    no workspace, a leading dimension below K."""
    from schwarmax_amd import _lib
    L = _lib.load_library()
    fake = ctypes.c_void_p(256)                     # never dereferenced: refused first
    assert L.smx_gemm_nt_f64_split(4, 8, 16, 1.0, fake, 16, fake, 16, fake, 8, None, 0, None) == 2
    assert 'smx_gemm_nt_f64_split' in _lib.last_error()
    assert L.smx_gemm_nt_f64_split(4, 8, 16, 1.0, fake, 15, fake, 16, fake, 8, fake, 64, None) == 2


def test_loglike_entry_points_refuse_bad_arguments_without_gpu():
    """smx_model_epilogue and smx_bootstrap_loglike check pointers and sizes
    before touching the device."""
    from schwarmax_amd import _lib
    L = _lib.load_library()
    fake = ctypes.c_void_p(256)                     # never dereferenced: refused first
    assert L.smx_model_epilogue(fake, 10, 2, 4, 3, None, fake, fake, 1.0, 1.0, fake, fake, fake,
                                fake, None) == 2
    assert 'smx_model_epilogue' in _lib.last_error()
    # leading dimension below P + 5B
    assert L.smx_model_epilogue(fake, 18, 2, 4, 3, fake, fake, fake, 1.0, 1.0, fake, fake, fake,
                                fake, None) == 2
    model, data = _lib.SmxBootstrapModel(), _lib.SmxBootstrapData()
    assert L.smx_bootstrap_loglike(model, 4, 5, 3, data, 1.0, fake, fake, fake, None) == 2
    assert 'null model map' in _lib.last_error()
    model.d3 = model.y_R = model.surface_density = 256
    data.xy = data.xy_err = data.xy_normal = 256
    for j in range(4):
        model.h[j] = data.h[j] = data.h_err[j] = data.h_normal[j] = 256
    model.ld3 = 4                                   # below P = 5
    assert L.smx_bootstrap_loglike(model, 4, 5, 3, data, 1.0, fake, fake, fake, None) == 2
    assert 'bad sizes' in _lib.last_error()


def test_library_build_refuses_small_workspace_without_gpu():
    """smx_build_library and smx_build_library_ckpt refuse, before any launch,
    a workspace that holds neither the sample rows nor the fused form's slabs."""
    from schwarmax_amd import _lib, model as M
    L = _lib.load_library()
    g = M.make_grids(100, (8, 4, 6), [[0, 10], [-3, 3], [-np.pi, np.pi]], [[-5, 5], [-5, 5]],
                     (20, 20), np.eye(3), True)
    sc = _lib.SmxSchedule(N_max=50, Omega=0.0, atol=1e-6, rtol=1e-4, dt_min=1e-5, dt_max=0.3)
    pot = _lib.SmxPotential()
    fake = ctypes.c_void_p(256)                     # never dereferenced: refused first
    small = min(L.smx_library_workspace_bytes(g, 10 * 4), 10 * 4 * 50 * 64) - 8
    args = (pot, g, sc, fake, fake, fake, 10, 4, fake, fake, fake, fake, fake, small, fake, fake,
            fake, fake)
    assert L.smx_build_library(*args, None) == 2
    assert 'workspace of' in _lib.last_error()
    assert L.smx_build_library_ckpt(*args, fake, 16, None) == 2
    assert 'workspace of' in _lib.last_error()


def test_grouped_nnls_refuses_bad_arguments_without_gpu():
    """smx_nnls_admm_group checks its arguments before touching the device."""
    from schwarmax_amd import _lib
    L = _lib.load_library()
    probs = (_lib.SmxNnlsProblem * 1)()            # null buffers
    fake = ctypes.c_void_p(256)                     # never dereferenced: refused first
    assert L.smx_nnls_admm_group(1, probs, 100, 4, 1.0, 10, fake, None) == 2
    assert 'null pointer in problem 0' in _lib.last_error()
    assert L.smx_nnls_admm_group(0, probs, 100, 4, 1.0, 10, fake, None) == 2
    assert L.smx_nnls_admm_group(_lib.NNLS_GROUP_MAX + 1, probs, 100, 4, 1.0, 10, fake, None) == 2
    assert L.smx_nnls_group_state_bytes(0, 1, 1) == -1
    assert L.smx_nnls_group_scratch_bytes(1000, 700, 20) > 8 * 1000 * 1000
    # a model's state plus the shared scratch is the single solve's workspace
    # (and the scratch's room for the models' descriptors)
    assert (L.smx_nnls_group_state_bytes(1000, 700, 20) + L.smx_nnls_group_scratch_bytes(1000, 700, 20)
            >= L.smx_nnls_workspace_bytes(1000, 700, 20))


def test_workspace_queries_without_gpu():
    from schwarmax_amd import _lib, model as M
    L = _lib.load_library()
    g = M.make_grids(100, (8, 4, 6), [[0, 10], [-3, 3], [-np.pi, np.pi]], [[-5, 5], [-5, 5]],
                     (20, 20), np.eye(3), True)
    P = 8 * 4 * 6
    assert L.smx_library_workspace_bytes(g, 10) == 10 * 8 * (((P + 7) // 8) * 8 + 6 * 100)
    # smx_build_library: the larger of the slabs and the sample rows (64 bytes
    # per attempt)
    slabs = 10 * 8 * (((P + 7) // 8) * 8 + 6 * 100)
    for N_max in (5, 50, 5000):
        sc = _lib.SmxSchedule(N_max=N_max, Omega=0.0, atol=1e-6, rtol=1e-4, dt_min=1e-5, dt_max=0.3)
        assert L.smx_library_build_bytes(g, sc, 10) == max(slabs, 10 * N_max * 64)
    assert L.smx_library_build_bytes(g, sc, -1) == -1
    # the checkpointed build: the slabs (which the adjoint reads), then the rows
    assert L.smx_library_ckpt_bytes(g, sc, 10) == slabs + 10 * 5000 * 64
    assert L.smx_library_ckpt_bytes(g, sc, -1) == -1
    assert L.smx_nnls_workspace_bytes(1000, 700, 20) > 4 * 1000 * 1000
    assert L.smx_nnls_workspace_bytes(0, 1, 1) == -1


def test_product_path_refuses_without_gpu():
    import pytest
    import torch
    from schwarmax_amd import _lib
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(RuntimeError, match="needs a ROCm GPU"):
        _lib.lib()


def test_noncanonical_potential_is_rejected():
    import pytest
    from schwarmax_amd import model as M
    with pytest.raises(NotImplementedError):
        M._canonical(lambda *a: 0, lambda *a: 0)


def test_pointer_dtype_is_checked():
    """A float32 / int32 buffer handed to a double* parameter raises instead of
    being read as doubles (the reference's JAX buffers are float32 without x64)."""
    import pytest
    import torch
    from schwarmax_amd import _lib
    f32 = torch.zeros(8, dtype=torch.float32)
    with pytest.raises(TypeError, match="float64"):
        _lib.ptr(f32, torch.float64)
    with pytest.raises(TypeError, match="int32"):
        _lib.ptr(torch.zeros(4, dtype=torch.int64), torch.int32)
    assert _lib.ptr(f32.to(torch.float64), torch.float64).value is not None
    assert _lib.ptr(None, torch.float64) is None
    # the Python drop-ins convert user arrays of any float type to float64
    got = _lib.dev(np.arange(5, dtype=np.float32), device='cpu')
    assert got.dtype == torch.float64 and np.array_equal(got.numpy(), np.arange(5.0))


def test_aperture_moments_are_validated():
    """The binned Gauss-Hermite terms are fixed point scaled by 1/s: the host
    refuses non-positive or non-finite dispersions and non-finite v0 before
    any launch, and passes good ones through as float64."""
    import pytest
    from schwarmax_amd import _lib
    good = _lib.dispersions(np.array([10.0, 150.0], dtype=np.float32), device='cpu')
    assert good.dtype == torch_float64() and np.array_equal(good.numpy(), [10.0, 150.0])
    for bad in ([10.0, 0.0], [10.0, -1.0], [np.nan, 5.0], [np.inf, 5.0]):
        with pytest.raises(ValueError, match="dispersions"):
            _lib.dispersions(np.array(bad), device='cpu')
    assert np.array_equal(_lib.velocities([-30.0, 0.0, 12.5], device='cpu').numpy(), [-30.0, 0.0, 12.5])
    with pytest.raises(ValueError, match="v0"):
        _lib.velocities([0.0, np.nan], device='cpu')


def torch_float64():
    import torch
    return torch.float64


def test_bin_map_with_and_without_sentinel():
    """Off-field samples read entry nX*nY: the sentinel when present; without
    it, JAX's clamped gather of index nX*nY, i.e. the last pixel's aperture."""
    import pytest
    from schwarmax_amd import _lib
    from oracle import integrate as OI
    nX, nY = 4, 3
    bm = np.arange(nX * nY, dtype=np.int64) % 5
    padded = _lib.bin_map(bm, nX, nY, device='cpu').numpy()
    assert padded.dtype == np.int32 and padded.size == nX * nY + 1
    assert np.array_equal(padded[:-1], bm) and padded[-1] == bm[-1]
    with_sent = np.concatenate([bm, [-1]])
    assert np.array_equal(_lib.bin_map(with_sent, nX, nY, device='cpu').numpy(), with_sent)
    with pytest.raises(ValueError):
        _lib.bin_map(bm[:-1], nX, nY, device='cpu')
    # the oracle's gather: an off-field sample with and without the sentinel
    args = (5, None, np.ones(5), [[0, 10], [-3, 3], [-np.pi, np.pi]], [[-1, 1], [-1, 1]],
            [2, 2, 2], [nX, nY], np.zeros(5), np.ones(5), np.eye(3))
    ys = np.zeros((6, 2))
    ys[0] = [0.1, 5.0]                      # on the field, off the field
    for m, off in ((with_sent, -1), (bm, bm[-1])):
        g = OI.Grids(args[0], m, *args[2:])
        _, vb, _ = g.indices(ys)
        assert vb[1] == off, (m.size, vb)
        assert vb[0] == m[2 + 1 * nX]         # x=0.1 -> ix=2, y=0 -> iy=1


def test_chunked_schedule_rounds_to_whole_chunks():
    """integrates.py:63 runs N_max // chunk_size chunks; the trajectory path
    runs all N_max attempts."""
    from schwarmax_amd import model as M
    s = M.schedule_of(dict(N_max=250, chunk=100))
    assert M.make_schedule(s, 1.0).N_max == 200
    assert M.make_schedule(s, 1.0, chunked=False).N_max == 250
    assert M.make_schedule(M.schedule_of(None), 1.0).N_max == 5000


def test_tuning_knobs_without_gpu():
    """smx_tune_set validates each knob's range and 0 restores the default."""
    from schwarmax_amd import _lib, tune
    L = _lib.load_library()
    v = ctypes.c_int64(0)
    defaults = {tune.KNOB_ADMM_OCCUPANCY: 2, tune.KNOB_GRAM_WAVES: 8,
                tune.KNOB_INV_SYRKX_MIN: 16384, tune.KNOB_GEMM64_SPLIT_TARGET: 2048,
                tune.KNOB_ADMM_LOADS: 0, tune.KNOB_GEMM64_KERNEL: 1, tune.KNOB_LIB_BINNING: 1}
    bad = {tune.KNOB_ADMM_OCCUPANCY: 9, tune.KNOB_GRAM_WAVES: 6, tune.KNOB_INV_SYRKX_MIN: 100,
           tune.KNOB_GEMM64_SPLIT_TARGET: 8192, tune.KNOB_ADMM_LOADS: 3, tune.KNOB_GEMM64_KERNEL: 3,
           tune.KNOB_LIB_BINNING: 3}
    good = {tune.KNOB_ADMM_OCCUPANCY: 3, tune.KNOB_GRAM_WAVES: 4, tune.KNOB_INV_SYRKX_MIN: 4096,
            tune.KNOB_GEMM64_SPLIT_TARGET: 512, tune.KNOB_ADMM_LOADS: 2, tune.KNOB_GEMM64_KERNEL: 2,
            tune.KNOB_LIB_BINNING: 2}
    for k in defaults:
        assert L.smx_tune_get(k, ctypes.byref(v)) == 0
        before = v.value
        assert L.smx_tune_set(k, bad[k]) == 2
        assert L.smx_tune_set(k, good[k]) == 0
        assert L.smx_tune_get(k, ctypes.byref(v)) == 0 and v.value == good[k]
        assert L.smx_tune_set(k, 0) == 0
        assert L.smx_tune_get(k, ctypes.byref(v)) == 0 and v.value == defaults[k]
        assert L.smx_tune_set(k, before) == 0
    assert L.smx_tune_set(len(defaults), 1) == 2 and L.smx_tune_get(-1, ctypes.byref(v)) == 2
    # the occupancy entry point is the same knob
    o = ctypes.c_int32(0)
    assert L.smx_admm_set_occupancy(5) == 0 and L.smx_admm_get_occupancy(ctypes.byref(o)) == 0
    assert o.value == 5 and L.smx_tune_get(tune.KNOB_ADMM_OCCUPANCY, ctypes.byref(v)) == 0
    assert v.value == 5
    assert L.smx_admm_set_occupancy(0) == 0


def test_nnls_groups_respect_budget_and_order():
    """The batched path's NNLS groups: consecutive runs of the models, each
    within the state budget unless a single model exceeds it, at most
    NNLS_GROUP_MAX per run (size queries only, no GPU)."""
    from schwarmax_amd import _lib, batch, nnls
    L = _lib.load_library()
    m, K = 17288, 100
    rng = np.random.default_rng(3)
    n_orbits = [int(v) for v in rng.integers(5000, 40001, size=70)] + [100] * 80
    models = list(range(len(n_orbits)))
    state = {i: L.smx_nnls_group_state_bytes(n_orbits[i], m, K) for i in models}
    for budget in (1, 16 << 30, 80 << 30, 1 << 50):
        groups = batch._nnls_groups(models, n_orbits, m, K, budget)
        assert [i for g in groups for i in g] == models
        for g in groups:
            assert 1 <= len(g) <= _lib.NNLS_GROUP_MAX
            assert len(g) == 1 or sum(state[i] for i in g) <= budget
        if budget == 1:
            assert all(len(g) == 1 for g in groups)
        if budget == 1 << 50:
            assert [len(g) for g in groups] == [64, 64, 22]
    # a group's device bytes: its states plus the scratch of its largest model
    g = [0, 1, 2]
    need = nnls.group_bytes([n_orbits[i] for i in g], m, K)
    assert need >= sum(state[i] for i in g) + L.smx_nnls_group_scratch_bytes(
        max(n_orbits[i] for i in g), m, K)
