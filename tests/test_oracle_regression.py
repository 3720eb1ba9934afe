"""Pins the oracle filters against the reference's locked regression baselines
(tests/test_regression/*.csv in the reference, copied into tests/golden)."""

import numpy as np
import pytest

from oracle import filters as fl, lg, threefry as tf
from tests import helpers
from tests.helpers import assert_close_tol, lg1_data, lg1_theta

F32 = np.float32


@pytest.fixture(scope="module")
def lg1():
    model = lg.LGModel(lg.LGSpec(1, 1))
    ser = fl.make_series(lg1_data(0), np.arange(1, 5, dtype=float), 0.0, nstep=1)
    return model, ser


def test_simulate(golden):
    model = lg.LGModel(lg.LGSpec(1, 1))
    ser = fl.make_series(np.zeros((4, 1)), np.arange(1, 5, dtype=float), 0.0, nstep=1)
    X, Y = fl.simulate(model, lg1_theta(), ser, tf.split(tf.key(20260811), 1)[0])
    g = golden["test_simulate_regression"]
    assert_close_tol(X[-1], g["X_final"])
    assert_close_tol(Y, g["Y"])


@pytest.mark.parametrize("resampler", ["reference", "fixed"])
def test_pfilter(golden, lg1, resampler):
    model, ser = lg1
    out = fl.pfilter(model, lg1_theta(), ser, tf.split(tf.key(20260811), 1)[0], 12, resampler=resampler)
    assert_close_tol(out["logLik"], golden["test_pfilter_regression"]["logLik"])


def test_pfilter_diagnostics(golden, lg1):
    model, ser = lg1
    g = golden["test_pfilter_diagnostics_regression"]
    thetas = [lg1_theta(), (lg1_theta() * F32(1.05)).astype(F32)]
    keys = tf.split(tf.key(20260812), 2)
    for r in range(2):
        o = fl.pfilter(model, thetas[r], ser, keys[r], 12, diagnostics=True)
        assert_close_tol(o["logLik"], g["logLik"][r])
        for name in ("CLL", "ESS", "filter_mean", "prediction_mean"):
            vals = np.ravel(o[name])
            for t in range(4):
                assert_close_tol(vals[t], g[f"{name}_rep{r}_t{t}"])


def test_mop(golden, lg1):
    model, ser = lg1
    v = fl.mop(model, lg1_theta(), ser, tf.split(tf.key(20260811), 1)[0], 12, 0.5)
    assert_close_tol(-v, golden["test_mop_regression"]["mop"])


def test_mif(golden, lg1):
    model, ser = lg1
    rw = fl.RWSigma(np.full(5, 0.02), np.zeros(5), a=0.5)
    lls, tr, _ = fl.mif(model, np.repeat(lg1_theta()[None], 12, 0), ser, tf.split(tf.key(20260811), 1)[0],
                        12, 3, rw)
    g = golden["test_mif_regression"]
    assert_close_tol(lls[-1], g["final_loglik"])
    assert_close_tol(tr[-1], g["final_theta"])


def test_mif_full_trace(golden, lg1):
    model, ser = lg1
    g = golden["test_mif_full_trace_regression"]
    rw = fl.RWSigma(np.full(5, 0.02), np.zeros(5), a=0.5)
    thetas = [lg1_theta(), (lg1_theta() * F32(1.05)).astype(F32)]
    keys = tf.split(tf.key(20260812), 2)
    for r in range(2):
        lls, tr, _ = fl.mif(model, np.repeat(thetas[r][None], 12, 0), ser, keys[r], 12, 4, rw, n_monitors=1)
        for m in range(4):
            assert_close_tol(lls[m], g[f"loglik_rep{r}_m{m}"])
        for m in range(5):
            for p, name in enumerate(model.spec.param_names):
                assert_close_tol(tr[m, p], g[f"{name}_rep{r}_m{m}"])


def test_panel_pfilter_units(golden):
    # units are independent when every parameter is unit-specific or shared-and-fixed
    model = lg.LGModel(lg.LGSpec(1, 1))
    keys = tf.split(tf.key(20260811), 2)
    got = [fl.pfilter(model, lg1_theta(), fl.make_series(lg1_data(s), np.arange(1, 5.0), 0.0, nstep=1),
                      keys[u], 12)["logLik"] for u, s in enumerate((0, 1))]
    assert_close_tol(got, golden["test_panel_pfilter_regression"]["logLik"])
