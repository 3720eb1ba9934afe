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
