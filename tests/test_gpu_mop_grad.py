"""GPU parity of the MOP(alpha) adjoint gradient against the oracle (torch autograd
of the reference's estimator with fixed ancestry) and central finite differences."""

import numpy as np
import pytest

from oracle import mop_grad as mg
from oracle import sv as osv, threefry as tf
from tests.helpers import workloads

pytestmark = pytest.mark.gpu


def sv_struct(T):
    import pypomp_amd.functional as F
    from pypomp_amd.models import SVModel
    from pypomp_amd.series import Series
    ys = workloads()["sv_seed4_ys"][:T]
    return F.PompStruct(model=SVModel(), units=[Series.build(ys, np.arange(1, T + 1.0), 0.0, nstep=1)]), ys


@pytest.mark.parametrize("K", [7, 50])
def test_sv_gradient_matches_autograd_oracle(K):
    import pypomp_amd.functional as F
    struct, ys = sv_struct(120)
    th = osv.SVModel().to_est(osv.default_theta())
    keys = tf.split(tf.key(11), 2)
    ll, g = F.mop_grad(struct, np.stack([th, th]), 300, 0.97, keys, K=K)
    for r in range(2):
        val_np, val_t, grad_ref, trace = mg.sv_mop_reference_grad(th, ys, keys[r], 300, 0.97)
        assert ll[r] == val_np  # same trajectories, exact fixed-point exp-sums
        scale = np.abs(grad_ref).max()
        np.testing.assert_allclose(g[r], grad_ref, rtol=1e-2, atol=1e-3 * scale)


def test_gradient_is_checkpoint_independent():
    import pypomp_amd.functional as F
    struct, _ = sv_struct(100)
    th = osv.SVModel().to_est(osv.default_theta())[None]
    k = tf.split(tf.key(3), 1)
    ll1, g1 = F.mop_grad(struct, th, 256, 0.5, k, K=1)
    ll2, g2 = F.mop_grad(struct, th, 256, 0.5, k, K=33)
    ll3, g3 = F.mop_grad(struct, th, 256, 0.5, k, K=100)
    assert ll1[0] == ll2[0] == ll3[0]
    np.testing.assert_array_equal(g1, g2)
    np.testing.assert_array_equal(g1, g3)


def test_value_matches_mop_entry_point():
    import pypomp_amd.functional as F
    struct, _ = sv_struct(80)
    th = osv.SVModel().to_est(osv.default_theta())
    keys = tf.split(tf.key(21), 3)
    ll, _ = F.mop_grad(struct, np.stack([th] * 3), 200, 0.8, keys, K=16)
    v = F.mop(struct, np.stack([th] * 3), 200, 0.8, keys, return_float64=True)
    np.testing.assert_allclose(ll, -v, rtol=1e-12)


def test_plain_form_equals_batched():
    import pypomp_amd.functional as F
    struct, _ = sv_struct(60)
    th = osv.SVModel().to_est(osv.default_theta())
    keys = tf.split(tf.key(4), 3)
    a = F.mop_grad(struct, np.stack([th] * 3), 128, 0.97, keys, K=10, form="batched")
    b = F.mop_grad(struct, np.stack([th] * 3), 128, 0.97, keys, K=10, form="plain")
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])


def test_autograd_function_backward():
    import torch
    import pypomp_amd.functional as F
    struct, _ = sv_struct(50)
    th = osv.SVModel().to_est(osv.default_theta())
    keys = tf.split(tf.key(9), 2)
    t = torch.tensor(np.stack([th, th]), dtype=torch.float32, requires_grad=True)
    ll = F.mop_loglik_autograd(t, struct, 128, 0.97, keys, 10)
    (-ll.sum() / 50).backward()
    _, g = F.mop_grad(struct, np.stack([th, th]), 128, 0.97, keys, K=10)
    np.testing.assert_allclose(t.grad.numpy(), -g / 50, rtol=1e-6)


def test_autograd_function_on_device_tensors():
    # theta and keys as device tensors: the forward stays on the GPU and the
    # gradient lands on theta's device
    import torch
    import pypomp_amd.functional as F
    struct, _ = sv_struct(50)
    th = osv.SVModel().to_est(osv.default_theta())
    keys = tf.split(tf.key(9), 2)
    t = torch.tensor(np.stack([th, th]), dtype=torch.float32, device="cuda", requires_grad=True)
    kd = torch.from_numpy(keys.view(np.int32)).cuda()
    ll = F.mop_loglik_autograd(t, struct, 128, 0.97, kd, 10)
    assert ll.device.type == "cuda" and ll.dtype == torch.float64
    (-ll.sum() / 50).backward()
    assert t.grad.device.type == "cuda"
    v, g = F.mop_grad(struct, np.stack([th, th]), 128, 0.97, keys, K=10)
    np.testing.assert_array_equal(ll.detach().cpu().numpy(), v)
    np.testing.assert_allclose(t.grad.cpu().numpy(), -g / 50, rtol=1e-6)


def test_workgroup_sizes_agree():
    # the value is bit-exact across workgroup sizes; the float64 gradient sums
    # follow the thread layout, so they agree to rounding
    import pypomp_amd.functional as F
    from pypomp_amd import autotune
    struct, _ = sv_struct(70)
    th = osv.SVModel().to_est(osv.default_theta())
    keys = tf.split(tf.key(8), 4)
    old = autotune.get("grad_block")
    try:
        autotune.set("grad_block", 256)
        a = F.mop_grad(struct, np.stack([th] * 4), 700, 0.97, keys, K=16)
        autotune.set("grad_block", 512)
        b = F.mop_grad(struct, np.stack([th] * 4), 700, 0.97, keys, K=16)
    finally:
        autotune.set("grad_block", old)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_allclose(a[1], b[1], rtol=1e-5, atol=1e-6 * np.abs(a[1]).max())
