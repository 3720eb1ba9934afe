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


def _lg_panel():
    model = lg.LGModel(lg.LGSpec(1, 1))
    sers = [fl.make_series(lg1_data(s), np.arange(1, 5.0), 0.0, nstep=1) for s in (0, 1)]
    th = lg1_theta()
    J = 12
    sh0 = np.repeat(th[None, :2], J, 0)
    un0 = np.repeat(np.stack([th[2:], th[2:]])[None], J, 0)
    rw = fl.RWSigma(np.full(5, 0.02), np.zeros(5), a=0.5)
    return model, sers, sh0, un0, rw, J


def test_panel_mif(golden):
    # lg_panel_struct: shared (A11, C11), unit-specific (Q11, R11, X0_1); the
    # panel's ParTrans() is the identity while the filter applies the LG model's
    # own transforms (bound when the unit Pomps were built)
    model, sers, sh0, un0, rw, J = _lg_panel()
    st, ut, _, _ = fl.panel_mif(model, sh0, un0, sers, tf.split(tf.key(20260811), 1)[0], J, 3, rw,
                                np.arange(5), par_trans=fl.IdentityTrans(model))
    g = golden["test_panel_mif_regression"]
    assert_close_tol(st[-1], g["final_shared"])
    assert_close_tol(ut[-1].ravel(), g["final_unit"])


def test_perfilter_ancestry_maps_final_to_initial_particles():
    # with every random-walk sigma zero the final swarm is the initial swarm
    # gathered through the composed resampling ancestry
    model, sers, _, _, _, J = _lg_panel()
    rng = np.random.default_rng(0)
    th = (lg1_theta()[None] + F32(0.01) * rng.standard_normal((J, 5)).astype(F32)).astype(F32)
    th_est = model.to_est(th)
    rw0 = fl.RWSigma(np.zeros(5), np.zeros(5), a=0.5)
    out, _, anc = fl.perfilter(model, 0, th_est, sers[0], tf.key(7), J, rw0, return_ancestry=True)
    assert anc.shape == (J,) and anc.min() >= 0 and anc.max() < J
    assert len(np.unique(anc)) < J  # resampling happened
    np.testing.assert_array_equal(out, th_est[anc])


def test_panel_mif_single_unit_block_false_equals_block_true():
    model, sers, sh0, un0, rw, J = _lg_panel()
    k = tf.key(99)
    a = fl.panel_mif(model, sh0, un0[:, :1], sers[:1], k, J, 2, rw, np.arange(5), block=True)
    b = fl.panel_mif(model, sh0, un0[:, :1], sers[:1], k, J, 2, rw, np.arange(5), block=False)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


def test_panel_mif_all_unit_specific_equals_per_unit_mif():
    # with no shared parameters (block=True) unit u's chain is the single-unit
    # mif of that unit driven by the unit keys split(iter_key, U)[u]
    model, sers, sh0, un0, rw, J = _lg_panel()
    full = np.concatenate([np.repeat(sh0[:, None, :], 2, 1), un0], axis=2)  # (J, U, 5)
    key = tf.key(5)
    st, ut, _, fu = fl.panel_mif(model, np.zeros((J, 0), F32), full, sers, key, J, 2, rw, np.arange(5))
    mkeys = tf.split(key, 3)[1:]
    for u in range(2):
        th = model.to_est(full[:, u, :])
        for m in range(2):
            th, ll = fl.perfilter(model, m, th, sers[u], tf.split(mkeys[m], 2)[u], J, rw)
            assert_close_tol(ut[m + 1, u, 0], ll)
        np.testing.assert_array_equal(fu[:, u, :], model.from_est(th))


def test_train(golden, lg1):
    # the reference's F.train on the lg_struct fixture: 3 Adam steps,
    # eta 0.01, alpha 0, one monitor; theta0 (natural values) used as the
    # estimation-scale start, as the reference passes it straight through
    from oracle import mop_grad as mg
    model, ser = lg1
    ys = lg1_data(0)

    def vg(th, k, alpha):
        val, _, g = mg.lg_mop_value_and_grad(model, th, ys, k, 12, alpha)
        return val, g

    eta = np.full((3, 5), 0.01, np.float32)
    nll, tr = mg.train(vg, lg1_theta(), tf.split(tf.key(20260811), 1)[0], 4, 3, eta, 0.0)
    g = golden["test_train_regression"]
    assert_close_tol(nll[-1], g["final_neg_loglik"])
    assert_close_tol(tr[-1], g["final_theta"])


def test_panel_train(golden, lg1):
    # the reference's F.panel_train on the lg_panel_struct fixture: units on
    # data seeds 0 and 1, shared A11 and C11, three Adam steps, eta 0.01,
    # alpha 0.97, chunk_size 1, keys split(key(20260811), M U); the panel
    # transform is the identity, so the natural values are the estimation-scale
    # start and the unit mechanics still apply the model's transforms
    from oracle import mop_grad as mg
    model, _ = lg1
    data = [lg1_data(0), lg1_data(1)]

    def vg(th, k, alpha, u):
        val, _, g = mg.lg_mop_value_and_grad(model, th, data[u], k, 12, alpha)
        return val, g

    th = lg1_theta()
    keys = tf.split(tf.key(20260811), 6).reshape(3, 2, 2)
    nll, sh, un = mg.panel_train(vg, th[:2], np.stack([th[2:], th[2:]]), np.arange(5), keys, 4, 3,
                                 np.full((3, 2), 0.01, np.float32), np.full((3, 3), 0.01, np.float32), 0.97)
    g = golden["test_panel_train_regression"]
    assert np.isnan(nll[0])
    assert_close_tol(nll[-1], g["final_neg_loglik"])
    assert_close_tol(sh[-1], g["final_shared"])
    assert_close_tol(un[-1], g["final_unit"])
