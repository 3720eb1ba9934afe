"""mp_pfilter_noalloc (the entry point with its scratch in a fixed library
array) returns what mp_pfilter with a caller workspace returns, through the
C ABI, on consecutive calls whose second input is smaller than the first."""

import ctypes as C

import numpy as np
import pytest

from oracle import threefry as tf
from tests.helpers import workloads

pytestmark = pytest.mark.gpu
F32 = np.float32


def lg2_case(T):
    import pypomp_amd.functional as F
    from pypomp_amd.models import LGModel
    from pypomp_amd.series import Series
    w = workloads()
    ys = w["lg2_seed1_ys"][:T]
    struct = F.PompStruct(model=LGModel(2, 2, separate_init=True),
                          units=[Series.build(ys, np.arange(1, ys.shape[0] + 1.0), 0.0, nstep=1)])
    return struct, w["lg2_seed1_theta"]


def run(struct, theta, R, J, seed, form, noalloc, flags):
    """One filter call through the C ABI; returns (loglik (R,), cll (R, T))."""
    import torch
    from pypomp_amd import _native
    L = _native.lib()
    dev = _native.require_device()
    ser = struct.device_series()
    md = struct.model.desc()
    T = ser.Tmax
    th = torch.from_numpy(np.repeat(theta[None], R, axis=0).astype(F32)).to(dev)
    keys = torch.from_numpy(tf.split(tf.key(seed), R).view(np.int32)).to(dev)
    ll = torch.full((R,), np.nan, dtype=torch.float64, device=dev)
    cll = torch.full((R, T), np.nan, dtype=torch.float32, device=dev)
    ptr, st = _native.ptr, _native.stream_ptr()
    common = (md, ser.desc, ptr(th), struct.model.d, ptr(keys), None, R, J, 0.0, flags, form, ptr(ll), ptr(cll),
              None, None, None)
    if noalloc:
        rc = L.mp_pfilter_noalloc(*common, st)
    else:
        n = L.mp_filter_workspace_size(_native.MP_ALGO_PF, md, ser.desc, R, J, form)
        ws = torch.empty(int(n), dtype=torch.uint8, device=dev)
        rc = L.mp_pfilter(*common, ptr(ws), C.c_size_t(n), st)
    _native.check(rc)
    torch.cuda.synchronize()
    return ll.cpu().numpy(), cll.cpu().numpy()


@pytest.mark.parametrize("form", [0, 1])
def test_noalloc_equals_the_workspace_entry_point(form):
    from pypomp_amd import _native
    flags = _native.MP_CLL
    # the second call is smaller (fewer filters, particles and observations),
    # so it runs over scratch the first call left behind
    for T, R, J, seed in ((40, 4, 20000, 11), (12, 2, 3000, 12)):
        print(f"noalloc: form={form} T={T} R={R} J={J}", flush=True)
        struct, theta = lg2_case(T)
        ref = run(struct, theta, R, J, seed, form, False, flags)
        got = run(struct, theta, R, J, seed, form, True, flags)
        assert np.all(np.isfinite(ref[0]))
        np.testing.assert_array_equal(got[0], ref[0])
        np.testing.assert_array_equal(got[1], ref[1])


def test_noalloc_rejects_calls_beyond_its_scratch():
    import torch
    from pypomp_amd import _native
    L = _native.lib()
    struct, theta = lg2_case(8)
    ser = struct.device_series()
    md = struct.model.desc()
    R, J = 64, 200000
    need = L.mp_filter_workspace_size(_native.MP_ALGO_PF, md, ser.desc, R, J, _native.MP_FORM_BATCHED)
    assert need > _native.MP_NOALLOC_SCRATCH_BYTES
    dev = _native.require_device()
    th = torch.zeros((R, struct.model.d), dtype=torch.float32, device=dev)
    keys = torch.zeros((R, 2), dtype=torch.int32, device=dev)
    ll = torch.zeros(R, dtype=torch.float64, device=dev)
    rc = L.mp_pfilter_noalloc(md, ser.desc, _native.ptr(th), struct.model.d, _native.ptr(keys), None, R, J, 0.0, 0,
                              _native.MP_FORM_BATCHED, _native.ptr(ll), None, None, None, None,
                              _native.stream_ptr())
    assert rc == 1
    assert b"no-allocation entry point" in L.mp_last_error()
