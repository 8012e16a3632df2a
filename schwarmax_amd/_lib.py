"""ctypes binding of libsmx.so (the C ABI declared in include/smx.h).

Device memory is owned by torch tensors; every call passes raw device pointers
and the current HIP stream.  There is no CPU fallback: loading fails loudly when
the library or a GPU is missing.
"""
import ctypes
import os

import numpy as np
import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'libsmx.so')

c_double_p = ctypes.POINTER(ctypes.c_double)


class SmxPotential(ctypes.Structure):
    _fields_ = [
        ('nfw_mGM', ctypes.c_double), ('nfw_Rs', ctypes.c_double), ('nfw_a', ctypes.c_double),
        ('nfw_b', ctypes.c_double), ('nfw_c', ctypes.c_double),
        ('nfw_origin', ctypes.c_double * 3), ('nfw_rot', ctypes.c_double * 9),
        ('mn_mGM', ctypes.c_double), ('mn_rho0', ctypes.c_double), ('mn_Rs', ctypes.c_double),
        ('mn_Hs', ctypes.c_double), ('mn_origin', ctypes.c_double * 3), ('mn_rot', ctypes.c_double * 9),
        ('t3_M', ctypes.c_double), ('t3_a', ctypes.c_double), ('t3_b', ctypes.c_double),
        ('t3_L', ctypes.c_double), ('t3_gamma', ctypes.c_double),
        ('v4_M', ctypes.c_double), ('v4_a', ctypes.c_double), ('v4_b', ctypes.c_double),
        ('v4_L', ctypes.c_double), ('v4_gamma', ctypes.c_double),
        ('G', ctypes.c_double),
    ]


class SmxGrids(ctypes.Structure):
    _fields_ = [
        ('nR', ctypes.c_int32), ('nz', ctypes.c_int32), ('nphi', ctypes.c_int32),
        ('nX', ctypes.c_int32), ('nY', ctypes.c_int32), ('num_vbin', ctypes.c_int32),
        ('symmetry', ctypes.c_int32), ('pad_', ctypes.c_int32),
        ('rz_min', ctypes.c_double * 3), ('rz_max', ctypes.c_double * 3),
        ('xy_min', ctypes.c_double * 2), ('xy_max', ctypes.c_double * 2),
        ('rot', ctypes.c_double * 9), ('area_pixel', ctypes.c_double),
    ]


class SmxSchedule(ctypes.Structure):
    _fields_ = [
        ('N_max', ctypes.c_int32), ('pad_', ctypes.c_int32), ('Omega', ctypes.c_double),
        ('atol', ctypes.c_double), ('rtol', ctypes.c_double), ('dt_min', ctypes.c_double),
        ('dt_max', ctypes.c_double),
    ]


class SmxBootstrapData(ctypes.Structure):
    _fields_ = [('xy', ctypes.c_void_p), ('xy_err', ctypes.c_void_p), ('xy_normal', ctypes.c_void_p),
                ('h', ctypes.c_void_p * 4), ('h_err', ctypes.c_void_p * 4),
                ('h_normal', ctypes.c_void_p * 4)]


class SmxBootstrapModel(ctypes.Structure):
    _fields_ = [('d3', ctypes.c_void_p), ('ld3', ctypes.c_int64), ('y_R', ctypes.c_void_p),
                ('surface_density', ctypes.c_void_p), ('h', ctypes.c_void_p * 4)]


class SmxNnlsProblem(ctypes.Structure):
    _fields_ = [
        ('lib', ctypes.c_void_p), ('ld_lib', ctypes.c_int64), ('n', ctypes.c_int64),
        ('row_scale', ctypes.c_void_p), ('Y', ctypes.c_void_p), ('w_init', ctypes.c_double),
        ('state', ctypes.c_void_p), ('W', ctypes.c_void_p),
    ]


NNLS_GROUP_MAX = 64


class SmxCylspline(ctypes.Structure):
    _fields_ = [
        ('u', ctypes.c_void_p), ('v', ctypes.c_void_p), ('coef', ctypes.c_void_p),
        ('nR', ctypes.c_int32), ('nZ', ctypes.c_int32), ('n_modes', ctypes.c_int32),
        ('n_modes_eval', ctypes.c_int32), ('R0', ctypes.c_double), ('G', ctypes.c_double),
    ]


CYLSPLINE_MAX_MODES = 21

# (name, restype, argtypes) of every entry point in include/smx.h
_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I32 = ctypes.c_int32
_D = ctypes.c_double
_SIGS = [
    ('smx_last_error', ctypes.c_int, [ctypes.c_char_p, _I64]),
    ('smx_version', ctypes.c_int, []),
    ('smx_timing_enable', ctypes.c_int, [ctypes.c_int]),
    ('smx_timing_read', ctypes.c_int, [_P, _P, _I32]),
    ('smx_potential_eval', ctypes.c_int, [ctypes.POINTER(SmxPotential), _P, _I64, _P, _P, _P]),
    ('smx_density_eval', ctypes.c_int, [ctypes.POINTER(SmxPotential), _P, _I64, _P, _P]),
    ('smx_initial_conditions', ctypes.c_int,
     [ctypes.POINTER(SmxPotential), _P, _P, _P, _I64, _I32, _I32, _D, _D, _P, _P, _P, _P]),
    ('smx_library_workspace_bytes', _I64, [ctypes.POINTER(SmxGrids), _I64]),
    ('smx_library_build_bytes', _I64,
     [ctypes.POINTER(SmxGrids), ctypes.POINTER(SmxSchedule), _I64]),
    ('smx_library_ckpt_bytes', _I64,
     [ctypes.POINTER(SmxGrids), ctypes.POINTER(SmxSchedule), _I64]),
    ('smx_build_library', ctypes.c_int,
     [ctypes.POINTER(SmxPotential), ctypes.POINTER(SmxGrids), ctypes.POINTER(SmxSchedule),
      _P, _P, _P, _I64, _I32, _P, _P, _P, _P, _P, _I64, _P, _P, _P, _P, _P]),
    ('smx_build_library_noalloc', ctypes.c_int,
     [ctypes.POINTER(SmxPotential), ctypes.POINTER(SmxGrids), ctypes.POINTER(SmxSchedule),
      _P, _P, _P, _I64, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    ('smx_build_library_ckpt', ctypes.c_int,
     [ctypes.POINTER(SmxPotential), ctypes.POINTER(SmxGrids), ctypes.POINTER(SmxSchedule),
      _P, _P, _P, _I64, _I32, _P, _P, _P, _P, _P, _I64, _P, _P, _P, _P, _P, _I32, _P]),
    ('smx_library_vjp_workspace_bytes', _I64, [_I64, _I32]),
    ('smx_library_vjp', ctypes.c_int,
     [ctypes.POINTER(SmxPotential), ctypes.POINTER(SmxGrids), ctypes.POINTER(SmxSchedule),
      _P, _P, _I64, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _I32, _P, _P, _P, _P, _P, _P, _P, _P]),
    ('smx_initial_conditions_jac', ctypes.c_int,
     [ctypes.POINTER(SmxPotential), _P, _P, _P, _I64, _I32, _I32, ctypes.c_double, _P, _P, _P]),
    ('smx_density_target_jac', ctypes.c_int,
     [ctypes.POINTER(SmxPotential), _P, _P, _P, _I64, ctypes.c_double, ctypes.c_double,
      ctypes.c_double, _P, _I64, _P, _P]),
    ('smx_nnls_operator_bytes', _I64, [_I64, _I64, _I64]),
    ('smx_nnls_operator', ctypes.c_int, [_P, _I64, _I64, _I64, _P, ctypes.c_double, _I64, _P, _P, _P]),
    ('smx_nnls_operator_apply', ctypes.c_int, [_I64, _I64, _I64, _P, _P, _I64, _P, _P]),
    ('smx_nnls_operator_apply_hi', ctypes.c_int, [_I64, _I64, _I64, _P, _P, _I64, _P, _P]),
    ('smx_jackknife_admm_step', ctypes.c_int, [_I64, _I64, _I64, _I64, _P, _P, _P, _P, _P, _P, _P,
                                                _D, _P, _P, _P, _P]),
    ('smx_cylspline_phi_m', ctypes.c_int, [_I32, _P, _I32, _P, _I32, _P, _P, _P, _I64, _P, _D, _P, _P]),
    ('smx_cylspline_eval', ctypes.c_int, [ctypes.POINTER(SmxCylspline), _P, _I64, _P, _P, _P, _P]),
    ('smx_legendre_q_half', ctypes.c_int, [_I32, _P, _I64, _P, _P]),
    ('smx_batch_workspace_bytes', _I64, [ctypes.POINTER(SmxGrids), _I32, _I64]),
    ('smx_build_library_batch', ctypes.c_int,
     [_I32, ctypes.POINTER(SmxPotential), ctypes.POINTER(SmxGrids), ctypes.POINTER(SmxSchedule),
      _P, _P, _P, ctypes.POINTER(_I64), _I32, _I32, ctypes.POINTER(ctypes.c_double),
      ctypes.POINTER(ctypes.c_double), _P, _P, _P, _P, ctypes.POINTER(_P), ctypes.POINTER(_P),
      ctypes.POINTER(_P), ctypes.POINTER(_P), _P, _I64, _P]),
    ('smx_build_library_traj', ctypes.c_int,
     [ctypes.POINTER(SmxPotential), ctypes.POINTER(SmxGrids), ctypes.POINTER(SmxSchedule),
      _P, _P, _P, _I64, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    ('smx_integrator_schedule_check', ctypes.c_int, [_I64, _I64, _I64, _P]),
    ('smx_adjoint_schedule_check', ctypes.c_int, [_I64, _I64, _I64, _P]),
    ('smx_bin_samples', ctypes.c_int,
     [ctypes.POINTER(SmxGrids), _P, _P, _I64, _P, _P, _P, _P, _P, _P, _P, _P]),
    ('smx_density_target', ctypes.c_int,
     [ctypes.POINTER(SmxPotential), _P, _P, _P, _I64, _D, _D, _D, _P, _I64, _P, _P]),
    ('smx_nnls_workspace_bytes', _I64, [_I64, _I64, _I64]),
    ('smx_nnls_admm', ctypes.c_int,
     [_P, _I64, _I64, _I64, _P, _P, _I64, _D, _I32, _D, _D, _P, _P, ctypes.POINTER(_I32), _P]),
    ('smx_nnls_admm_full_products', ctypes.c_int, [_I64, _I64, _I64, _P, ctypes.POINTER(_I32), _P]),
    ('smx_tune_set', ctypes.c_int, [_I32, _I64]),
    ('smx_tune_get', ctypes.c_int, [_I32, ctypes.POINTER(_I64)]),
    ('smx_nnls_group_state_bytes', _I64, [_I64, _I64, _I64]),
    ('smx_nnls_group_scratch_bytes', _I64, [_I64, _I64, _I64]),
    ('smx_nnls_admm_group', ctypes.c_int,
     [_I32, ctypes.POINTER(SmxNnlsProblem), _I64, _I64, ctypes.c_double, _I32, _P, _P]),
    ('smx_nnls_group_full_products', ctypes.c_int, [_I64, _I64, _I64, _P, ctypes.POINTER(_I32), _P]),
    ('smx_admm_set_occupancy', ctypes.c_int, [_I32]),
    ('smx_admm_get_occupancy', ctypes.c_int, [ctypes.POINTER(_I32)]),
    ('smx_admm_set_delta_threshold', ctypes.c_int, [ctypes.c_double]),
    ('smx_admm_get_delta_threshold', ctypes.c_int, [ctypes.POINTER(ctypes.c_double)]),
    ('smx_admm_product_time', ctypes.c_int,
     [_I64, _I64, _I64, _P, _I32, _I32, ctypes.POINTER(ctypes.c_float), _P]),
    ('smx_gemm_nt_f32', ctypes.c_int, [_I64, _I64, _I64, ctypes.c_float, _P, _I64, _P, _I64,
                                       ctypes.c_float, _P, _I64, _P]),
    ('smx_gemm_nt_f64', ctypes.c_int, [_I64, _I64, _I64, _D, _P, _I64, _P, _I64, _D, _P, _I64, _P]),
    ('smx_gemm_nt_f64_split', ctypes.c_int, [_I64, _I64, _I64, _D, _P, _I64, _P, _I64, _P, _I64, _P, _I64, _P]),
    ('smx_spd_inverse_f64', ctypes.c_int, [_P, _I64, _P, _P]),
    ('smx_model_maps', ctypes.c_int, [_P, _I64, _I64, _I64, _P, _I64, _P, _P]),
    ('smx_model_epilogue', ctypes.c_int,
     [_P, _I64, _I64, _I32, _I32, _P, _P, _P, _D, _D, _P, _P, _P, _P, _P]),
    ('smx_bootstrap_loglike', ctypes.c_int,
     [ctypes.POINTER(SmxBootstrapModel), _I64, _I32, _I32, ctypes.POINTER(SmxBootstrapData), _D,
      _P, _P, _P, _P]),
]

EXPORTED = [s[0] for s in _SIGS]

_lib = None


def load_library(path=LIB_PATH):
    """Load libsmx.so and bind signatures (does not touch the GPU)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"libsmx.so not built at {path}; run __graft_entry__.build()")
    lib = ctypes.CDLL(path)
    for name, res, args in _SIGS:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def lib():
    """The loaded library, after checking a GPU is present (no CPU fallback)."""
    if not torch.cuda.is_available():
        raise RuntimeError("schwarmax_amd needs a ROCm GPU: torch.cuda.is_available() is False")
    return load_library()


def last_error():
    buf = ctypes.create_string_buffer(2048)
    load_library().smx_last_error(buf, 2048)
    return buf.value.decode(errors='replace')


def check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed (code {rc}): {last_error()}")


def ptr(t, dtype=None):
    """Device pointer of a contiguous tensor (or None).

    ``dtype``: the element type the C entry point reads through this pointer
    (torch.float64 for ``const double*``, torch.int32 for ``const int32_t*``,
    ...); a tensor of another type raises TypeError instead of being read as
    the wrong type or past its end."""
    if t is None:
        return None
    if dtype is not None and t.dtype != dtype:
        raise TypeError(f"tensor has dtype {t.dtype}, the entry point reads {dtype}; "
                        f"convert it first (tensor.to({dtype}))")
    if not t.is_contiguous():
        raise ValueError("tensor must be contiguous")
    return ctypes.c_void_p(t.data_ptr())


def bin_map(bin_mapping, nX, nY, device='cuda'):
    """Pixel -> aperture map as the kernels read it: int32, nX*nY + 1 entries.

    Entry nX*nY is read for samples off the sky field.  The reference indexes
    bin_mapping with id nX*nY for such samples (integrates.py:132-143,
    utils.py:40-42) and accepts maps of nX*nY or nX*nY + 1 entries (utils.py:429):
    with the sentinel the entry there is used (normally -1, dropped); without it
    JAX clamps the out-of-range gather to the last pixel's aperture, so that
    aperture's id is appended here."""
    bm = np.asarray(bin_mapping.cpu() if isinstance(bin_mapping, torch.Tensor) else bin_mapping)
    bm = bm.astype(np.int32).reshape(-1)
    npix = int(nX) * int(nY)
    if bm.size < npix:
        raise ValueError(f"bin_mapping has {bm.size} entries, the {nX} x {nY} field needs "
                         f"{npix} (plus an optional sentinel)")
    if bm.size == npix:
        bm = np.concatenate([bm, bm[-1:]])
    return torch.as_tensor(np.ascontiguousarray(bm[:npix + 1]), dtype=torch.int32, device=device)


def stream_ptr(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _host(a):
    h = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    return np.asarray(h, dtype=np.float64).reshape(-1)


def dispersions(s, device='cuda'):
    """The apertures' Gauss-Hermite dispersions s as a device FP64 tensor.

    The library kernels hold the binned Gauss-Hermite terms as 64-bit fixed
    point scaled by a bound proportional to 1/s, so every s must be positive
    and finite; the reference's gh weights (ghmoments.py:6-36) would be NaN or
    infinite otherwise.  Raises ValueError."""
    h = _host(s)
    if not np.all(np.isfinite(h) & (h > 0.0)):
        raise ValueError("aperture dispersions 's' must be positive and finite")
    return dev(h, device=device)


def velocities(v0, device='cuda'):
    """The apertures' Gauss-Hermite mean velocities v0 as a device FP64 tensor;
    they must be finite, as the fixed-point binned terms cannot carry a NaN.
    Raises ValueError."""
    h = _host(v0)
    if not np.all(np.isfinite(h)):
        raise ValueError("aperture velocities 'v0' must be finite")
    return dev(h, device=device)


def dev(a, dtype=torch.float64, device='cuda'):
    """Host or device array -> contiguous device tensor of dtype."""
    if isinstance(a, torch.Tensor):
        return a.to(device=device, dtype=dtype).contiguous()
    a = np.ascontiguousarray(a)
    if not a.flags.writeable:
        a = a.copy()
    return torch.as_tensor(a, dtype=dtype, device=device)


STAGES = ('ics', 'integrate', 'finalize', 'gram', 'factor', 'rhs', 'admm', 'maps', 'density',
          'admm_product', 'adjoint', 'jacobian')


TIMING = {'on': False}


def timing_enable(on=True):
    check(load_library().smx_timing_enable(1 if on else 0), 'smx_timing_enable')
    TIMING['on'] = bool(on)


def timing_read():
    """{stage: (ms, launches)} summed since the last enable/read."""
    ms = (ctypes.c_double * len(STAGES))()
    ln = (ctypes.c_int64 * len(STAGES))()
    check(load_library().smx_timing_read(ms, ln, len(STAGES)), 'smx_timing_read')
    return {s: (ms[i], ln[i]) for i, s in enumerate(STAGES)}
