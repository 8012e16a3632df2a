"""JAX's default PRNG restated in NumPy, for the draws the reference takes from
fixed keys.

The reference (JAX 0.7.2, float32 defaults) draws
* the Jeans velocity normals from keys 42, 109 and 2026 (model.py:112,
  jeans.py:89),
* the jitter uniforms from ``split(PRNGKey(911), 6)`` (model.py:128-138), and
* the loader's initial radii and heights from keys 10086 and 10010
  (utils.py:671, 679, through sample_from_logP utils.py:489-509).

JAX is a third-party dependency (jax/jaxlib 0.7.2) whose generator is the
counter-based Threefry-2x32 with 20 rounds (Salmon, Moraes, Dror and Shaw,
"Parallel random numbers: as easy as 1, 2, 3", SC'11).  Its layering, restated
from JAX's published behaviour:

* ``PRNGKey(seed)`` is the uint32 pair (seed >> 32, seed & 0xffffffff); for a
  32-bit seed the high word is 0.
* Partitionable layout (JAX's default since 0.5.0, so the reference's): a draw
  of shape S hashes the 64-bit row-major element index i as the counter pair
  (i >> 32, i & 0xffffffff); 32-bit draws take y0 ^ y1 of the hash;
  ``split(key, num)`` returns the pairs (y0, y1) of counters (0, j).
* Original layout (JAX before 0.5.0): a draw of m words hashes the counters
  0..m-1 paired as (j, j + ceil(m/2)) and returns all y0 then all y1;
  ``split`` draws 2*num words.
* ``uniform`` keeps the top 23 bits as the mantissa of a float in [1, 2) and
  subtracts 1; ``normal`` maps a uniform on (-1, 1) through sqrt(2) erfinv,
  with erfinv in float32 by Giles' single-precision approximation
  (M. Giles, "Approximating the erfinv function", GPU Computing Gems Jade,
  2011), the method XLA's float32 erf_inv uses.

The core, the original layout's split and its uniform/normal maps are
checked against published answers (the Random123 known-answer vectors and
JAX's documented ``split(PRNGKey(0))`` and ``normal(PRNGKey(0), (1,))``).
"""
import functools

import numpy as np

_U32 = np.uint32
_ROT = ((13, 15, 26, 6), (17, 29, 16, 24))
_PARITY = 0x1BD11BDA


def _rotl(x, r):
    return (x << _U32(r)) | (x >> _U32(32 - r))


def threefry2x32(k0, k1, x0, x1):
    """Threefry-2x32-20 of counters (x0, x1) under key (k0, k1), elementwise."""
    k0, k1 = _U32(k0), _U32(k1)
    ks = (k0, k1, _U32(k0 ^ k1 ^ _U32(_PARITY)))
    x0 = np.asarray(x0, dtype=_U32)
    x1 = np.asarray(x1, dtype=_U32)
    with np.errstate(over='ignore'):
        x0 = x0 + ks[0]
        x1 = x1 + ks[1]
        for j in range(5):
            for r in _ROT[j % 2]:
                x0 = x0 + x1
                x1 = _rotl(x1, r) ^ x0
            x0 = x0 + ks[(j + 1) % 3]
            x1 = x1 + ks[(j + 2) % 3] + _U32(j + 1)
    return x0, x1


def PRNGKey(seed):
    """jax.random.PRNGKey for the threefry implementation."""
    seed = int(seed)
    if not 0 <= seed < 2 ** 64:
        raise ValueError("seed must lie in [0, 2**64)")
    return np.array([seed >> 32, seed & 0xFFFFFFFF], dtype=_U32)


def _counters(size):
    idx = np.arange(size, dtype=np.uint64)
    return (idx >> np.uint64(32)).astype(_U32), (idx & np.uint64(0xFFFFFFFF)).astype(_U32)


def _words_original(key, m):
    """m hashed words of the original layout: counters 0..m-1 as (j, j + h)."""
    h = (m + 1) // 2
    c = np.zeros(2 * h, dtype=_U32)
    c[:m] = np.arange(m, dtype=_U32)
    y0, y1 = threefry2x32(key[0], key[1], c[:h], c[h:])
    return np.concatenate([y0, y1])[:m]


def split(key, num=2, partitionable=True):
    """jax.random.split: (num, 2) uint32 keys."""
    key = np.asarray(key, dtype=_U32)
    if partitionable:
        hi, lo = _counters(num)
        y0, y1 = threefry2x32(key[0], key[1], hi, lo)
        return np.stack([y0, y1], axis=-1)
    return _words_original(key, 2 * num).reshape(num, 2)


def random_bits(key, shape, partitionable=True):
    """32-bit random words of the given shape (jax.random.bits, uint32)."""
    key = np.asarray(key, dtype=_U32)
    shape = tuple(int(s) for s in np.atleast_1d(shape)) if np.ndim(shape) else (int(shape),)
    size = int(np.prod(shape))
    if partitionable:
        hi, lo = _counters(size)
        y0, y1 = threefry2x32(key[0], key[1], hi, lo)
        return (y0 ^ y1).reshape(shape)
    return _words_original(key, size).reshape(shape)


def uniform(key, shape, minval=0.0, maxval=1.0, partitionable=True):
    """jax.random.uniform in float32."""
    bits = random_bits(key, shape, partitionable)
    f = ((bits >> _U32(9)) | _U32(0x3F800000)).view(np.float32) - np.float32(1.0)
    lo, hi = np.float32(minval), np.float32(maxval)
    return np.maximum(lo, f * (hi - lo) + lo).astype(np.float32)


_GILES_LT5 = (2.81022636e-08, 3.43273939e-07, -3.5233877e-06, -4.39150654e-06, 0.00021858087,
              -0.00125372503, -0.00417768164, 0.246640727, 1.50140941)
_GILES_GE5 = (-0.000200214257, 0.000100950558, 0.00134934322, -0.00367342844, 0.00573950773,
              -0.0076224613, 0.00943887047, 1.00167406, 2.83297682)


def erfinv_f32(x):
    """Giles' single-precision erfinv, evaluated in float32."""
    x = np.asarray(x, dtype=np.float32)
    with np.errstate(divide='ignore', invalid='ignore'):
        w = -np.log1p(-x * x)
        wa = (w - np.float32(2.5)).astype(np.float32)
        wb = (np.sqrt(w) - np.float32(3.0)).astype(np.float32)
    pa = np.full_like(x, _GILES_LT5[0])
    pb = np.full_like(x, _GILES_GE5[0])
    for ca, cb in zip(_GILES_LT5[1:], _GILES_GE5[1:]):
        pa = (np.float32(ca) + pa * wa).astype(np.float32)
        pb = (np.float32(cb) + pb * wb).astype(np.float32)
    out = np.where(w < np.float32(5.0), pa, pb) * x
    return np.where(np.abs(x) == 1, np.copysign(np.float32(np.inf), x), out).astype(np.float32)


def normal(key, shape, partitionable=True):
    """jax.random.normal in float32."""
    lo = np.nextafter(np.float32(-1.0), np.float32(0.0))
    u = uniform(key, shape, lo, 1.0, partitionable)
    return (np.float32(np.sqrt(2.0)) * erfinv_f32(u)).astype(np.float32)


@functools.lru_cache(maxsize=4)
def _reference_draws(n, r, partitionable):
    normals = np.stack([normal(PRNGKey(s), (n,), partitionable) for s in (42, 109, 2026)],
                       axis=-1).astype(np.float64)
    keys = split(PRNGKey(911), 6, partitionable)
    uni = np.stack([uniform(keys[c], (n, r), partitionable=partitionable) for c in range(6)],
                   axis=-1).astype(np.float64)
    normals.setflags(write=False)
    uni.setflags(write=False)
    return normals, uni


def reference_draws(n, r=4, partitionable=True):
    """The build_model draws for n orbits and r realizations, as float64:
    Jeans normals (n, 3), columns from keys 42, 109, 2026 (model.py:112,
    jeans.py:89), and jitter uniforms (n, r, 6), component c from
    split(PRNGKey(911), 6)[c] (model.py:128-138).  Read-only, cached."""
    return _reference_draws(int(n), int(r), bool(partitionable))


def sample_from_logP(x_grid, logP, N, key, partitionable=True):
    """utils.py:489-509 in float32: inverse CDF of exp(logP) on x_grid at N
    uniforms, by linear interpolation (numpy.interp's rule, which jnp.interp
    follows for increasing x)."""
    x = np.asarray(x_grid, dtype=np.float32)
    lp = np.asarray(logP, dtype=np.float32)
    P = np.exp(lp - lp.max()).astype(np.float32)
    P = (P / P.sum(dtype=np.float32)).astype(np.float32)
    cdf = np.cumsum(P, dtype=np.float32)
    u = uniform(key, (int(N),), partitionable=partitionable)
    return _interp_f32(u, cdf, x)


def _interp_f32(u, xp, fp):
    i = np.clip(np.searchsorted(xp, u, side='right'), 1, len(xp) - 1)
    df = fp[i] - fp[i - 1]
    dx = xp[i] - xp[i - 1]
    delta = u - xp[i - 1]
    tiny = np.spacing(np.finfo(np.float32).eps)
    dx0 = np.abs(dx) <= tiny
    f = np.where(dx0, fp[i - 1], fp[i - 1] + (delta / np.where(dx0, np.float32(1), dx)) * df)
    f = np.where(u < xp[0], fp[0], f)
    f = np.where(u > xp[-1], fp[-1], f)
    return f.astype(np.float32)
