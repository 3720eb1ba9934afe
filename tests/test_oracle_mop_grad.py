"""Pins the MOP gradient oracle (oracle/mop_grad.py)."""

import numpy as np
import pytest

from oracle import mop_grad as mg
from oracle import sv as osv, threefry as tf
from tests.helpers import workloads


@pytest.fixture(scope="module")
def sv_case():
    ys = workloads()["sv_seed4_ys"][:40]
    model = osv.SVModel()
    th_est = model.to_est(osv.default_theta())
    return ys, th_est


def test_torch_value_equals_numpy_oracle(sv_case):
    ys, th = sv_case
    val_np, val_t, grad, _ = mg.sv_mop_reference_grad(th, ys, tf.split(tf.key(3), 1)[0], 64, 0.97)
    assert abs(val_np - val_t) < 2e-5 * abs(val_np)
    assert np.all(np.isfinite(grad))


def test_autograd_equals_frozen_ancestry_fd(sv_case):
    ys, th = sv_case
    _, _, grad, trace = mg.sv_mop_reference_grad(th, ys, tf.split(tf.key(5), 1)[0], 64, 0.97)
    fd = mg.frozen_fd_grad(th, ys, trace, 0.97, h=1e-3)
    np.testing.assert_allclose(grad, fd, rtol=1e-3, atol=1e-4 * np.abs(fd).max())


def test_alpha_zero_drops_weight_history(sv_case):
    # With alpha = 0 the weights wP vanish every step: the gradient only flows
    # through the current step's measurement densities and the states.
    ys, th = sv_case
    _, _, g0, _ = mg.sv_mop_reference_grad(th, ys, tf.split(tf.key(7), 1)[0], 64, 0.0)
    _, _, g1, _ = mg.sv_mop_reference_grad(th, ys, tf.split(tf.key(7), 1)[0], 64, 1.0)
    assert not np.allclose(g0, g1)
