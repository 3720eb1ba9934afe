"""GPU parity of the key-tree draws and the standalone resampler, through the C ABI."""

import numpy as np
import pytest

from oracle import resample as rs
from oracle import threefry as tf

pytestmark = pytest.mark.gpu
F32 = np.float32


def _keys(n, seed=5):
    return tf.split(tf.key(seed), n)


def test_bits_bit_exact():
    import pypomp_amd.random as prandom
    keys = _keys(16)
    got = prandom.bits(keys, 1000).cpu().numpy().view(np.uint32)
    assert np.array_equal(got, tf.random_bits_many(keys, 1000))


def test_uniform_bit_exact():
    import pypomp_amd.random as prandom
    keys = _keys(8, 9)
    got = prandom.uniform(keys, 4096).cpu().numpy()
    exp = tf.uniform_from_bits(tf.random_bits_many(keys, 4096))
    assert np.array_equal(got.view(np.uint32), exp.view(np.uint32))


def test_split_bit_exact():
    import pypomp_amd.random as prandom
    keys = _keys(4, 2)
    got = prandom.split_device(keys, 33).cpu().numpy().view(np.uint32)
    assert np.array_equal(got, tf.split_many(keys, 33))


def test_normal_bit_exact():
    # erfinv with the deterministic log1p: the device reproduces the oracle's bits
    import pypomp_amd.random as prandom
    keys = _keys(8, 4)
    got = prandom.normal(keys, 20000).cpu().numpy()
    exp = tf.normal_many(keys, 20000)
    assert np.array_equal(got.view(np.uint32), exp.view(np.uint32))


def _device_resample(lw, keys):
    import torch
    from pypomp_amd import _native
    dev = _native.require_device()
    R, J = lw.shape
    t_lw = torch.from_numpy(np.ascontiguousarray(lw)).to(dev)
    t_k = torch.from_numpy(np.ascontiguousarray(keys).view(np.int32)).to(dev)
    out = torch.empty((R, J), dtype=torch.int32, device=dev)
    nws = _native.lib().mp_resample_workspace_size(R, J)
    ws = torch.empty(nws, dtype=torch.uint8, device=dev)
    _native.check(_native.lib().mp_resample_systematic(_native.ptr(t_lw), _native.ptr(t_k), R, J,
                                                       _native.ptr(out), _native.ptr(ws), nws,
                                                       _native.stream_ptr()))
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("J", [1, 2, 7, 64, 1000, 4099, 65536])
def test_resample_bit_exact(J):
    rng = np.random.default_rng(J)
    R = 6
    lw = (rng.normal(size=(R, J)) * np.array([0.01, 1, 5, 50, 500, 1])[:, None]).astype(F32)
    if J > 4:
        lw[5, : J // 2] = -np.inf      # half the particles dead
        lw[4, 1] = np.nan               # NaN weight: no offspring
    keys = _keys(R, 31 + J)
    got = _device_resample(lw, keys)
    for r in range(R):
        u = tf.uniform(keys[r])
        exp = rs.resample_fixed_fast(lw[r], u)
        assert np.array_equal(got[r], exp), (J, r)


def test_resample_all_dead_is_identity():
    J = 513
    lw = np.full((2, J), -np.inf, dtype=F32)
    got = _device_resample(lw, _keys(2, 3))
    assert np.array_equal(got[0], np.arange(J)) and np.array_equal(got[1], np.arange(J))
