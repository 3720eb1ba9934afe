"""JAX-compatible threefry keys on the host (key derivation only).

Keys are uint32 arrays of shape (..., 2), bit-identical to ``jax.random.key``
data in its default threefry2x32 implementation.  The hashing runs in the
native library (``mp_host_split``); device-side draws use the same functions.
"""

import ctypes as C

import numpy as np

from . import _native


def key(seed):
    """``jax.random.key(seed)`` for a non-negative integer seed."""
    seed = int(seed)
    if seed < 0:
        seed &= 0xFFFFFFFFFFFFFFFF
    return np.array([(seed >> 32) & 0xFFFFFFFF, seed & 0xFFFFFFFF], dtype=np.uint32)


def split(k, num=2):
    """``jax.random.split(k, num)`` (partitionable mode): (num, 2) uint32."""
    k = np.ascontiguousarray(np.asarray(k, dtype=np.uint32).reshape(2))
    out = np.empty((int(num), 2), dtype=np.uint32)
    _native.check(_native.lib().mp_host_split(k.ctypes.data_as(C.c_void_p), int(num),
                                              out.ctypes.data_as(C.c_void_p)))
    return out


def threefry2x32(keys, counters):
    """Elementwise Threefry-2x32-20: keys, counters (n, 2) uint32 -> (n, 2)."""
    keys = np.ascontiguousarray(np.asarray(keys, dtype=np.uint32).reshape(-1, 2))
    ctr = np.ascontiguousarray(np.asarray(counters, dtype=np.uint32).reshape(-1, 2))
    out = np.empty_like(ctr)
    _native.check(_native.lib().mp_host_threefry2x32(
        keys.ctypes.data_as(C.c_void_p), ctr.ctypes.data_as(C.c_void_p),
        out.ctypes.data_as(C.c_void_p), ctr.shape[0]))
    return out


def _device_draw(fn, keys, n, dtype):
    import torch
    dev = _native.require_device()
    keys = np.ascontiguousarray(np.asarray(keys, dtype=np.uint32).reshape(-1, 2))
    kt = torch.from_numpy(keys.view(np.int32)).to(dev)
    if dtype == "key":
        out = torch.empty((keys.shape[0], n, 2), dtype=torch.int32, device=dev)
    else:
        out = torch.empty((keys.shape[0], n), dtype=dtype, device=dev)
    _native.check(getattr(_native.lib(), fn)(_native.ptr(kt), keys.shape[0], int(n),
                                             _native.ptr(out), _native.stream_ptr()))
    return out


def bits(keys, n):
    """Device ``jax.random.bits(k, (n,))`` for each key: (nkeys, n) as int32 bit patterns."""
    import torch
    return _device_draw("mp_random_bits", keys, n, torch.int32)


def uniform(keys, n):
    import torch
    return _device_draw("mp_uniform", keys, n, torch.float32)


def normal(keys, n):
    import torch
    return _device_draw("mp_normal", keys, n, torch.float32)


def split_device(keys, n):
    return _device_draw("mp_split", keys, n, "key")
