"""Autotuner for the launch-shape knobs of the NNLS stages.

The knobs (``smx_tune_set``; process-wide, read by each launch) choose how
work is laid out, never what is computed:

* ADMM product occupancy: the workgroups per CU that the product's split-K
  rule targets (the product is the largest share of a model evaluation);
* Gram waves: 4 waves of 128 x 64 outputs or 8 waves of 128 x 32 per
  workgroup;
* inverse syrkx threshold: the node size from which the FP64 inverse's two
  symmetric updates use rocBLAS syrkx rather than full dgemm products;
* FP64 GEMM split target: the workgroups the split-K rule of the FP64 GEMM
  (right-hand sides c = -Y U and the ADMM anchors) fills;
* ADMM loads: how the product loads the inverse's planes (non-temporal,
  cached, or by the automatic rule on the planes' size and K).

``autotune_admm`` times the product kernel at each occupancy on a synthetic
problem of the workload's shape, full and delta launches weighted by the
share of full products a solve takes.  ``autotune_factor`` times a one-
iteration solve of the same shape (Gram, inverse, c = -Y U and one FP64
anchor) under each candidate of the other three knobs, one knob at a time,
with the library's stage timing.  Both keep the fastest settings.  The
settings change only layouts and the order of partial sums (results move by
rounding), never the iterate definition.

``ensure_tuned`` tunes once per process and shape bucket; the batched paths
(``batch.model_likelihood_batch`` and ``grad.model_likelihood_and_grad_batch``
with autotune=True) call it.
"""
import ctypes

import torch

from . import _lib

# Share of full products among the split products of a 1,000-iteration solve
# at the plain-20k shapes (675 of 990; the other 10 iterations take FP64
# anchors); the rest are delta products.
FULL_SHARE = 0.68
CANDIDATES = (1, 2, 3, 4, 5, 6, 7, 8)

KNOB_ADMM_OCCUPANCY = 0
KNOB_GRAM_WAVES = 1
KNOB_INV_SYRKX_MIN = 2
KNOB_GEMM64_SPLIT_TARGET = 3
KNOB_ADMM_LOADS = 4
KNOB_GEMM64_KERNEL = 5
KNOB_LIB_BINNING = 6
GRAM_WAVES = (4, 8)
GEMM64_KERNELS = (1, 2)   # 1 thin kernel, 2 tiled kernel with split-K
SYRKX_MIN = (4096, 8192, 16384, 32768)
SPLIT_TARGETS = (256, 512, 1024, 2048, 4096)

_RESULTS = {}


def get_knob(knob):
    v = ctypes.c_int64(0)
    _lib.check(_lib.lib().smx_tune_get(int(knob), ctypes.byref(v)), 'smx_tune_get')
    return int(v.value)


def set_knob(knob, value):
    """0 restores the knob's default."""
    _lib.check(_lib.lib().smx_tune_set(int(knob), int(value)), 'smx_tune_set')


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
    """Time the ADMM product at each occupancy candidate (loads of M by the
    automatic policy); keep the fastest.  Then time the two explicit load
    policies at that occupancy and keep one only if it beats the automatic
    policy by more than 3% (the automatic one also adapts to other shapes).

    Returns {'occupancy': best, 'ms': {occ: (full, delta, weighted)},
    'loads': 0 | 1 | 2, 'loads_ms': {policy: weighted}, 'default': ...}.
    """
    default = 2
    before = get_occupancy()
    loads_before = get_knob(KNOB_ADMM_LOADS)
    ws = _synthetic_workspace(n, m, K, seed)
    results, lres = {}, {}

    def weighted():
        time_product(n, m, K, ws, True, 1)                  # first launch at this setting
        t_full = time_product(n, m, K, ws, True, reps)
        t_delta = time_product(n, m, K, ws, False, reps)
        return (t_full, t_delta, full_share * t_full + (1 - full_share) * t_delta)

    try:
        set_knob(KNOB_ADMM_LOADS, 0)
        for occ in candidates:
            set_occupancy(occ)
            results[int(occ)] = weighted()
        best = min(results, key=lambda o: results[o][2])
        set_occupancy(best)
        lres[0] = results[best][2]
        for v in (1, 2):
            set_knob(KNOB_ADMM_LOADS, v)
            lres[v] = weighted()[2]
    finally:
        set_occupancy(before)
        set_knob(KNOB_ADMM_LOADS, loads_before)
    del ws
    torch.cuda.empty_cache()
    loads = min((1, 2), key=lres.get)
    loads = loads if lres[loads] < 0.97 * lres[0] else 0
    if apply:
        set_occupancy(best)
        set_knob(KNOB_ADMM_LOADS, loads)
    return {'occupancy': best, 'ms': results, 'loads': loads, 'loads_ms': lres,
            'default': default, 'shape': (n, m, K)}


def _solve_stages(n, m, K, lib, scale, Y, ws, W):
    """{stage: (ms, launches)} of a one-iteration solve on the given buffers."""
    L = _lib.lib()
    iters = ctypes.c_int32(0)
    _lib.timing_enable(True)
    _lib.check(L.smx_nnls_admm(_lib.ptr(lib), m, n, m, _lib.ptr(scale), _lib.ptr(Y), K, 1.0, 1, 1.0,
                               0.0, _lib.ptr(ws), _lib.ptr(W), ctypes.byref(iters),
                               _lib.stream_ptr()), 'smx_nnls_admm')
    torch.cuda.synchronize()
    return _lib.timing_read()


def autotune_factor(n=20000, m=17288, K=100, seed=0, apply=True, repeats=2):
    """Time the Gram, inverse and FP64-GEMM knobs on a one-iteration solve of
    shape (n, m, K); keep the fastest value of each (one knob at a time, the
    others at their current values; the tiled FP64 GEMM's split target only
    when the tiled kernel wins).  Returns {'gram_waves': best, 'syrkx_min':
    best, 'gemm64_kernel': best, 'split_target': best, 'ms': {knob name:
    {value: ms}}}."""
    L = _lib.lib()
    g = torch.Generator(device='cuda')
    g.manual_seed(seed)
    lib = torch.rand((n, m), dtype=torch.float64, device='cuda', generator=g)
    lib *= (torch.rand((n, m), device='cuda', generator=g) < 0.05)
    scale = torch.ones(m, dtype=torch.float64, device='cuda')
    Y = torch.rand((K, m), dtype=torch.float64, device='cuda', generator=g)
    ws = torch.empty(int(L.smx_nnls_workspace_bytes(n, m, K)), dtype=torch.uint8, device='cuda')
    W = torch.empty((K, n), dtype=torch.float64, device='cuda')
    was_on = _lib.TIMING['on']
    knobs = {'gram_waves': (KNOB_GRAM_WAVES, GRAM_WAVES, ('gram',)),
             # a threshold above n leaves every node on dgemm; below the
             # smallest node size every node uses syrkx
             'syrkx_min': (KNOB_INV_SYRKX_MIN,
                           tuple(sorted({v for v in SYRKX_MIN if v <= n} | {max(n + 1, 256)})),
                           ('factor',)),
             'gemm64_kernel': (KNOB_GEMM64_KERNEL, GEMM64_KERNELS, ('rhs', 'admm')),
             'split_target': (KNOB_GEMM64_SPLIT_TARGET, SPLIT_TARGETS, ('rhs', 'admm'))}
    results, best = {}, {}
    try:
        _solve_stages(n, m, K, lib, scale, Y, ws, W)          # warm-up
        for name, (knob, values, stages) in knobs.items():
            before = get_knob(knob)
            if name == 'split_target' and best.get('gemm64_kernel') == 1:
                best[name] = before                           # the thin kernel has no split target
                continue
            res = {}
            try:
                for v in values:
                    set_knob(knob, v)
                    t = []
                    for _ in range(repeats):
                        st = _solve_stages(n, m, K, lib, scale, Y, ws, W)
                        t.append(sum(st[s][0] for s in stages))
                    res[int(v)] = min(t)
            finally:
                set_knob(knob, before)
            results[name] = res
            best[name] = min(res, key=res.get)
            if apply:
                set_knob(knob, best[name])
    finally:
        _lib.timing_enable(was_on)
    del lib, ws, W
    torch.cuda.empty_cache()
    return dict(best, ms=results, shape=(n, m, K))


def _bucket(n, m, K):
    # one tuning per power-of-two bucket of n and of K, per m
    return (int(n).bit_length(), int(m), int(K).bit_length())


def ensure_tuned(n, m, K, **kw):
    """Tune once per process and shape bucket; returns that tuning's result
    (and applies its settings)."""
    key = _bucket(n, m, K)
    if key not in _RESULTS:
        res = autotune_admm(n, m, K, **kw)
        res['factor'] = autotune_factor(n, m, K)
        _RESULTS[key] = res
    else:
        res = _RESULTS[key]
        set_occupancy(res['occupancy'])
        set_knob(KNOB_ADMM_LOADS, res['loads'])
        f = res['factor']
        set_knob(KNOB_GRAM_WAVES, f['gram_waves'])
        set_knob(KNOB_INV_SYRKX_MIN, f['syrkx_min'])
        set_knob(KNOB_GEMM64_KERNEL, f['gemm64_kernel'])
        set_knob(KNOB_GEMM64_SPLIT_TARGET, f['split_target'])
    return _RESULTS[key]
