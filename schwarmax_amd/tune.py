"""Autotuner for the ADMM product's occupancy.

The ADMM product (one launch per iteration, the largest share of a model
evaluation) splits its reduction so that the grid holds a chosen number of
product workgroups per CU (and at least one workgroup per CU).  ``autotune_admm`` times the product kernel at each
candidate setting on a synthetic problem of the workload's shape, full and
delta launches weighted by the share of full products a solve takes, and
keeps the fastest setting in the library (``smx_admm_set_occupancy``).  The
setting changes only the order of the split-K partial sums (results move by
FP32 rounding), never the iterate definition.

``ensure_tuned`` tunes once per process and shape bucket; the batched path
(``batch.model_likelihood_batch(autotune=True)``) calls it.
"""
import ctypes

import torch

from . import _lib

# Share of full products among the split products of a 1,000-iteration solve
# at the plain-20k shapes (675 of 990; the other 10 iterations take FP64
# anchors); the rest are delta products.
FULL_SHARE = 0.68
CANDIDATES = (1, 2, 3, 4, 5, 6, 7, 8)

_RESULTS = {}


def get_occupancy():
    v = ctypes.c_int32(0)
    _lib.check(_lib.lib().smx_admm_get_occupancy(ctypes.byref(v)), 'smx_admm_get_occupancy')
    return int(v.value)


def set_occupancy(workgroups_per_cu):
    """0 restores the default."""
    _lib.check(_lib.lib().smx_admm_set_occupancy(int(workgroups_per_cu)), 'smx_admm_set_occupancy')


def _synthetic_workspace(n, m, K, seed):
    """A filled NNLS workspace of shape (n, m, K): a sparse nonnegative
    library, unit row scales, random data, two ADMM iterations."""
    L = _lib.lib()
    g = torch.Generator(device='cuda')
    g.manual_seed(seed)
    lib = torch.rand((n, m), dtype=torch.float64, device='cuda', generator=g)
    lib *= (torch.rand((n, m), device='cuda', generator=g) < 0.05)
    scale = torch.ones(m, dtype=torch.float64, device='cuda')
    Y = torch.rand((K, m), dtype=torch.float64, device='cuda', generator=g)
    ws = torch.empty(int(L.smx_nnls_workspace_bytes(n, m, K)), dtype=torch.uint8, device='cuda')
    W = torch.empty((K, n), dtype=torch.float64, device='cuda')
    iters = ctypes.c_int32(0)
    _lib.check(L.smx_nnls_admm(_lib.ptr(lib), m, n, m, _lib.ptr(scale), _lib.ptr(Y), K, 1.0, 2,
                               1.0, 0.0, _lib.ptr(ws), _lib.ptr(W), ctypes.byref(iters),
                               _lib.stream_ptr()), 'smx_nnls_admm')
    del lib, W
    return ws


def time_product(n, m, K, ws, full, reps):
    ms = ctypes.c_float(0.0)
    _lib.check(_lib.lib().smx_admm_product_time(n, m, K, _lib.ptr(ws), 1 if full else 0, int(reps),
                                                ctypes.byref(ms), _lib.stream_ptr()),
               'smx_admm_product_time')
    return float(ms.value)


def autotune_admm(n=20000, m=17288, K=100, candidates=CANDIDATES, reps=10, full_share=FULL_SHARE,
                  seed=0, apply=True):
    """Time the ADMM product at each occupancy candidate; keep the fastest.

    Returns {'occupancy': best, 'ms': {occ: (full, delta, weighted)}, 'default': ...}.
    """
    default = 2
    before = get_occupancy()
    ws = _synthetic_workspace(n, m, K, seed)
    results = {}
    try:
        for occ in candidates:
            set_occupancy(occ)
            time_product(n, m, K, ws, True, 1)              # first launch at this grid
            t_full = time_product(n, m, K, ws, True, reps)
            t_delta = time_product(n, m, K, ws, False, reps)
            results[int(occ)] = (t_full, t_delta, full_share * t_full + (1 - full_share) * t_delta)
    finally:
        set_occupancy(before)
    del ws
    torch.cuda.empty_cache()
    best = min(results, key=lambda o: results[o][2])
    if apply:
        set_occupancy(best)
    return {'occupancy': best, 'ms': results, 'default': default, 'shape': (n, m, K)}


def _bucket(n, m, K):
    # one tuning per power-of-two bucket of n and of K, per m
    return (int(n).bit_length(), int(m), int(K).bit_length())


def ensure_tuned(n, m, K, **kw):
    """Tune once per process and shape bucket; returns that tuning's result
    (and applies its setting)."""
    key = _bucket(n, m, K)
    if key not in _RESULTS:
        _RESULTS[key] = autotune_admm(n, m, K, **kw)
    else:
        set_occupancy(_RESULTS[key]['occupancy'])
    return _RESULTS[key]
