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


# ---- optimizers of the train loop (core/optimizer.py, core/algorithms/train.py) ----

def test_bfgs_update_satisfies_the_secant_equation():
    rng = np.random.default_rng(5)
    d = 4
    a = rng.normal(size=(d, d)).astype(np.float32)
    H = (a @ a.T + d * np.eye(d)).astype(np.float32)      # symmetric positive definite
    prev = rng.normal(size=d).astype(np.float32)
    g = rng.normal(size=d).astype(np.float32)
    eta = np.full(d, 0.05, np.float32)
    # the first two iterations take -grad and keep H
    for step in (0, 1):
        dr, Hn, pg = mg.bfgs_direction(g, H, prev, step, eta)
        np.testing.assert_array_equal(dr, -g)
        np.testing.assert_array_equal(Hn, H)
        np.testing.assert_array_equal(pg, g)
    dr, Hn, _ = mg.bfgs_direction(g, H, prev, 2, eta)
    s = np.float32(0.05) * (-prev)
    y = g - prev
    np.testing.assert_allclose(Hn @ y, s, rtol=2e-3, atol=1e-6)   # secant: H_new y = s
    np.testing.assert_allclose(Hn, Hn.T, rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(dr, -(Hn @ g), rtol=1e-5, atol=1e-6)
    # a non-finite curvature (y . s = 0) keeps H
    _, Hk, _ = mg.bfgs_direction(prev, H, prev, 3, eta)
    np.testing.assert_array_equal(Hk, H)


def test_full_matrix_adam_is_the_inverse_square_root():
    rng = np.random.default_rng(6)
    d = 3
    m1 = np.zeros(d, np.float32)
    v1 = np.zeros((d, d), np.float32)
    for step in range(4):
        g = rng.normal(size=d).astype(np.float32)
        dr, m_new, v_new = mg.fmadam_direction(g, m1, v1, step)
        np.testing.assert_allclose(v_new, 0.999 * v1 + 0.001 * np.outer(g, g), rtol=1e-5, atol=1e-9)
        mh = m_new.astype(np.float64) / (1 - 0.9 ** (step + 1))
        A = v_new.astype(np.float64) / (1 - 0.999 ** (step + 1)) + 1e-4 * np.eye(d)
        # d = -A^(-1/2) mh  <=>  d' A d = |mh|^2  and  |d|^2 = mh' A^-1 mh
        dd = dr.astype(np.float64)
        np.testing.assert_allclose(dd @ A @ dd, mh @ mh, rtol=1e-3)
        np.testing.assert_allclose(dd @ dd, mh @ np.linalg.solve(A, mh), rtol=1e-3)
        assert dd @ mh < 0
        m1, v1 = m_new, v_new
    # one parameter: the scalar closed form
    g = np.array([0.3], np.float32)
    dr, _, _ = mg.fmadam_direction(g, np.zeros(1, np.float32), np.zeros((1, 1), np.float32), 0)
    np.testing.assert_allclose(dr, [-0.3 / np.sqrt(0.09 + 1e-4)], rtol=1e-5)


def test_armijo_backtracking():
    q = lambda x: float(np.sum(x.astype(np.float64) ** 2))   # noqa: E731
    x = np.array([1.0, -2.0], np.float32)
    g = (2 * x).astype(np.float32)
    dr = -g
    eta = mg.armijo(q, np.float32(q(x)), x, g, dr, 4.0, c=0.1)
    # halved from 4 until sufficient decrease: 4, 2, 1 fail, 0.5 passes (q(x - 0.5 g) = 0)
    assert eta == np.float32(0.5)
    assert q(x + eta * dr) <= q(x) + eta * 0.1 * float(g @ dr)
    # an objective that never decreases (or is NaN) stops after max_itn halvings
    eta = mg.armijo(lambda _: np.nan, np.float32(1.0), x, g, dr, 1.0, max_itn=10)
    assert eta == np.float32(0.5 ** 10)


# ---- Hessian (jax.hessian of _mop_internal_mean, ad_helpers.py:8, 56-75) ----

@pytest.mark.parametrize("alpha", [0.97, 0.0])
def test_sv_hessian_is_symmetric_and_the_derivative_of_the_gradient(sv_case, alpha):
    ys, th = sv_case
    obj = mg.sv_hessian_case(th, ys, tf.split(tf.key(9), 1)[0], 64, alpha)
    val, g, H, fd = mg.hessian_and_fd(obj, th, len(ys))
    np.testing.assert_allclose(H, H.T, rtol=1e-10, atol=1e-10 * np.abs(H).max())
    np.testing.assert_allclose(H, fd, rtol=1e-5, atol=1e-6 * np.abs(H).max())
    # its gradient is the oracle's MOP gradient
    _, _, grad, _ = mg.sv_mop_reference_grad(th, ys, tf.split(tf.key(9), 1)[0], 64, alpha)
    np.testing.assert_allclose(g, grad, rtol=1e-9, atol=1e-12)


def test_lg_hessian_is_the_derivative_of_the_gradient():
    from oracle import lg as olg
    ys = workloads()["lg2_seed1_ys"][:30]
    model = olg.LGModel(olg.LGSpec(2, 2, separate_init=True))
    th = model.to_est(workloads()["lg2_seed1_theta"])
    obj = mg.lg_hessian_case(model, th, ys, tf.split(tf.key(10), 1)[0], 48, 0.9)
    val, g, H, fd = mg.hessian_and_fd(obj, th, len(ys))
    assert H.shape == (19, 19)
    np.testing.assert_allclose(H, H.T, rtol=1e-10, atol=1e-10 * np.abs(H).max())
    np.testing.assert_allclose(H, fd, rtol=1e-5, atol=1e-6 * np.abs(H).max())


def test_hessian_differs_from_differencing_the_moving_stop_gradient(sv_case):
    # the stop-gradient values are fixed at theta: re-freezing them at each
    # shifted point (differences of the MOP gradient itself) gives another
    # matrix when alpha > 0 -- the property the Hessian kernel must respect
    import torch
    ys, th = sv_case
    key = tf.split(tf.key(11), 1)[0]
    obj = mg.sv_hessian_case(th, ys, key, 64, 0.97)
    _, _, H, _ = mg.hessian_and_fd(obj, th, len(ys))
    h = 1e-4
    moving = np.zeros_like(H)
    for j in range(len(th)):
        cols = []
        for sgn in (1, -1):
            t = torch.tensor(np.asarray(th, np.float64) + sgn * h * np.eye(len(th))[j], requires_grad=True)
            obj(t).backward()
            cols.append(t.grad.numpy().copy())
        moving[:, j] = (cols[0] - cols[1]) / (2 * h)
    assert not np.allclose(H, moving, rtol=1e-3, atol=1e-3 * np.abs(H).max())


def test_panel_mop_reduces_to_the_unit_objective():
    # one unit: the panel objective is the unit's -loglik / T and its gradient
    # the unit's; two copies of a unit with one key: the same value, the
    # shared gradient summed over both and each unit gradient halved
    from oracle import lg as olg
    from tests.helpers import lg1_data, lg1_theta
    model = olg.LGModel(olg.LGSpec(1, 1))
    ys = np.concatenate([lg1_data(0), lg1_data(1)])
    th = model.to_est(lg1_theta())
    key = tf.split(tf.key(3), 1)[0]

    def vg(t, k, u):
        val, _, g = mg.lg_mop_value_and_grad(model, t, ys, k, 50, 0.9)
        return val, g

    perm = np.array([0, 1, 2, 3, 4])
    ll, g = vg(th, key, 0)
    v1, s1, u1 = mg.panel_mop(vg, th[:2], th[None, 2:], perm, key[None], 8)
    assert v1 == np.float32(np.float32(-ll) / np.float32(8))
    np.testing.assert_allclose(s1, -g[:2] / 8, rtol=1e-6)
    np.testing.assert_allclose(u1[0], -g[2:] / 8, rtol=1e-6)
    v2, s2, u2 = mg.panel_mop(vg, th[:2], np.stack([th[2:], th[2:]]), perm, np.stack([key, key]), 8, chunk_size=2)
    assert v2 == v1
    np.testing.assert_allclose(s2, s1, rtol=1e-6)
    np.testing.assert_allclose(u2, np.stack([u1[0], u1[0]]) / 2, rtol=1e-6)
