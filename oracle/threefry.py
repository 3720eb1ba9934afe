"""Threefry-2x32-20 and the JAX partitionable key tree, restated in NumPy.

TEST INFRASTRUCTURE ONLY (oracle).  The reference draws every random number
through ``jax.random`` with the default threefry2x32 implementation in
partitionable mode; the call sites are pypomp/core/algorithms/pfilter.py:23,105,160,
mop.py:29,89,122, mif.py:22,113,123,136,174, core/model_mechanics.py:586,904 and
core/algorithms/helpers.py:28.  JAX itself is an un-vendored dependency
(pyproject.toml:28, ``jax>=0.5.3``), so its published algorithm is restated here:

* hash: Threefry-2x32 with 20 rounds (Salmon, Moraes, Dror, Shaw, "Parallel
  random numbers: as easy as 1, 2, 3", SC'11), rotation constants
  13,15,26,6 / 17,29,16,24, key-schedule parity 0x1BD11BDA, key injection
  after every 4 rounds;
* ``key(seed)``      -> (seed >> 32, seed & 0xffffffff);
* ``split(k, n)[i]`` -> threefry(k, (i >> 32, i & 0xffffffff)), both words;
* ``random_bits(k, shape)[i]`` -> b0 ^ b1 of threefry(k, (hi_i, lo_i));
* ``uniform``        -> bitcast((bits >> 9) | 0x3f800000) - 1, then * (hi-lo) + lo,
  then max(lo, .);
* ``normal``         -> sqrt(2) * erfinv(uniform(lo=nextafter(-1, 0), hi=1)),
  with erfinv from M. Giles, "Approximating the erfinv function" (GPU Computing
  Gems Jade, 2011), single-precision branch, with his w = -log((1-x)(1+x))
  evaluated by the engine's deterministic log (oracle/detmath.py).  (XLA
  forms w = -log1p(-x*x); the values differ in the last bits only.)

Pinned by the Random123 known-answer vectors (tests/test_oracle_threefry.py) and,
end to end, by the reference regression CSVs (tests/test_oracle_regression.py).
"""

import numpy as np

from .detmath import d_log

U32 = np.uint32
_ROT0 = (13, 15, 26, 6)
_ROT1 = (17, 29, 16, 24)
_PARITY = U32(0x1BD11BDA)


def _rotl(x, r):
    return (x << U32(r)) | (x >> U32(32 - r))


def threefry2x32(k0, k1, x0, x1):
    """Threefry-2x32-20 of counter (x0, x1) under key (k0, k1); uint32 arrays."""
    with np.errstate(over="ignore"):
        k0 = np.asarray(k0, dtype=U32)
        k1 = np.asarray(k1, dtype=U32)
        x0 = np.asarray(x0, dtype=U32).copy()
        x1 = np.asarray(x1, dtype=U32).copy()
        ks = (k0, k1, k0 ^ k1 ^ _PARITY)
        x0 = x0 + ks[0]
        x1 = x1 + ks[1]
        for blk in range(5):
            for r in (_ROT0 if blk % 2 == 0 else _ROT1):
                x0 = x0 + x1
                x1 = _rotl(x1, r)
                x1 = x1 ^ x0
            x0 = x0 + ks[(blk + 1) % 3]
            x1 = x1 + ks[(blk + 2) % 3] + U32(blk + 1)
    return x0, x1


def key(seed):
    """``jax.random.key(seed)`` for a non-negative integer seed: shape (2,) uint32."""
    seed = int(seed)
    return np.array([(seed >> 32) & 0xFFFFFFFF, seed & 0xFFFFFFFF], dtype=U32)


def _counters(n):
    i = np.arange(n, dtype=np.uint64)
    return (i >> np.uint64(32)).astype(U32), (i & np.uint64(0xFFFFFFFF)).astype(U32)


def split(k, n=2):
    """``jax.random.split(k, n)`` (partitionable): array (n, 2) uint32."""
    k = np.asarray(k, dtype=U32)
    hi, lo = _counters(n)
    b0, b1 = threefry2x32(k[0], k[1], hi, lo)
    return np.stack([b0, b1], axis=-1)


def split_many(keys, n):
    """split each key of an (m, 2) array into n: returns (m, n, 2)."""
    keys = np.asarray(keys, dtype=U32)
    hi, lo = _counters(n)
    b0, b1 = threefry2x32(keys[:, 0:1], keys[:, 1:2], hi[None, :], lo[None, :])
    return np.stack([b0, b1], axis=-1)


def random_bits(k, size):
    """32-bit ``random_bits`` of a flat shape ``size`` for one key."""
    k = np.asarray(k, dtype=U32)
    hi, lo = _counters(size)
    b0, b1 = threefry2x32(k[0], k[1], hi, lo)
    return b0 ^ b1


def random_bits_many(keys, size):
    """random bits for each key of an (m, 2) array: returns (m, size)."""
    keys = np.asarray(keys, dtype=U32)
    hi, lo = _counters(size)
    b0, b1 = threefry2x32(keys[:, 0:1], keys[:, 1:2], hi[None, :], lo[None, :])
    return b0 ^ b1


def bits_to_unit_float(bits):
    """Mantissa trick: float32 in [0, 1) from 32 random bits."""
    fb = (np.asarray(bits, dtype=U32) >> U32(9)) | U32(0x3F800000)
    return fb.view(np.float32) - np.float32(1.0)


def uniform_from_bits(bits, lo=0.0, hi=1.0):
    lo = np.float32(lo)
    hi = np.float32(hi)
    f = bits_to_unit_float(bits)
    span = np.float32(hi - lo)
    return np.maximum(lo, (f * span + lo).astype(np.float32)).astype(np.float32)


def uniform(k, size=None):
    """``jax.random.uniform(k, shape)``; ``size=None`` gives the scalar draw."""
    if size is None:
        return uniform_from_bits(random_bits(k, 1))[0]
    return uniform_from_bits(random_bits(k, size))


# Giles (2011) single-precision erfinv coefficients, Horner order (highest first).
_ERFINV_CENTRAL = np.array(
    [2.81022636e-08, 3.43273939e-07, -3.5233877e-06, -4.39150654e-06,
     0.00021858087, -0.00125372503, -0.00417768164, 0.246640727, 1.50140941],
    dtype=np.float32)
_ERFINV_TAIL = np.array(
    [-0.000200214257, 0.000100950558, 0.00134934322, -0.00367342844,
     0.00573950773, -0.0076224613, 0.00943887047, 1.00167406, 2.83297682],
    dtype=np.float32)


def erfinv_f32(x):
    """Giles' single-precision erfinv evaluated in float32 arithmetic."""
    x = np.asarray(x, dtype=np.float32)
    with np.errstate(divide="ignore", invalid="ignore"):
        w = -d_log(((np.float32(1.0) - x).astype(np.float32) * (np.float32(1.0) + x).astype(np.float32)).astype(np.float32))
        lt = w < np.float32(5.0)
        wc = np.where(lt, w - np.float32(2.5), np.sqrt(w) - np.float32(3.0)).astype(np.float32)
        p = np.where(lt, _ERFINV_CENTRAL[0], _ERFINV_TAIL[0]).astype(np.float32)
        for i in range(1, 9):
            c = np.where(lt, _ERFINV_CENTRAL[i], _ERFINV_TAIL[i]).astype(np.float32)
            p = (c + p * wc).astype(np.float32)
        r = (p * x).astype(np.float32)
        r = np.where(np.abs(x) == np.float32(1.0), x * np.finfo(np.float32).max, r)
    return r.astype(np.float32)


_NORMAL_LO = np.nextafter(np.float32(-1.0), np.float32(0.0))
_SQRT2 = np.float32(np.sqrt(2.0))


def normal_from_bits(bits):
    u = uniform_from_bits(bits, lo=_NORMAL_LO, hi=1.0)
    return (_SQRT2 * erfinv_f32(u)).astype(np.float32)


def normal(k, size):
    """``jax.random.normal(k, (size,))`` in float32."""
    return normal_from_bits(random_bits(k, size))


def normal_many(keys, size):
    """normal draws of a flat shape ``size`` for each key of an (m, 2) array."""
    return normal_from_bits(random_bits_many(keys, size))
