"""The C restatement (CPU baseline) agrees with the NumPy oracle."""

import numpy as np

from oracle import cpu, filters as fl, lg as olg, sv as osv, threefry as tf
from tests.helpers import workloads


def test_c_lg_pfilter_matches_numpy_oracle():
    w = workloads()
    ys, theta = w["lg2_seed1_ys"], w["lg2_seed1_theta"]
    keys = tf.split(tf.key(12), 3)
    got = cpu.lg_pfilter(theta, 2, ys, 500, keys)
    model = olg.LGModel(olg.LGSpec(2, 2, separate_init=True))
    ser = fl.make_series(ys, np.arange(1, ys.shape[0] + 1.0), 0.0, nstep=1)
    for r in range(3):
        exp = float(fl.pfilter(model, theta, ser, keys[r], 500)["logLik"])
        assert abs(got[r] - exp) <= 2e-5 * abs(exp), (got[r], exp)


def test_c_sv_mop_matches_numpy_oracle():
    w = workloads()
    ys = w["sv_seed4_ys"][:100]
    model = osv.SVModel()
    th = model.to_est(osv.default_theta())
    keys = tf.split(tf.key(6), 2)
    got = cpu.sv_mop(th, ys, 300, 0.97, keys)
    ser = fl.make_series(ys, np.arange(1, 101.0), 0.0, nstep=1)
    for r in range(2):
        exp = float(fl.mop(model, th, ser, keys[r], 300, 0.97))
        assert abs(got[r] - exp) <= 2e-5 * abs(exp), (got[r], exp)


def test_c_sv_mop_gradient_matches_the_autograd_oracle():
    # the compiled hand adjoint (the config-4 CPU baseline) against torch
    # float64 autograd of the reference's MOP objective on the same ancestry
    from oracle import mop_grad as mg
    ys = workloads()["sv_seed4_ys"][:60]
    th = osv.SVModel().to_est(osv.default_theta())
    th = (th + np.float32(0.03) * np.arange(6, dtype=np.float32)).astype(np.float32)
    keys = tf.split(tf.key(8), 2)
    for alpha in (0.97, 0.0):
        ll, g = cpu.sv_mop_grad(th, ys, 128, alpha, keys)
        mop = cpu.sv_mop(th, ys, 128, alpha, keys)
        np.testing.assert_array_equal(ll, mop)  # same forward, same value
        for r in range(2):
            val, _, grad, _ = mg.sv_mop_reference_grad(th, ys, keys[r], 128, alpha)
            assert abs(ll[r] - val) <= 1e-12 * abs(val)
            np.testing.assert_allclose(g[r], grad, rtol=1e-4, atol=1e-5 * np.abs(grad).max())
