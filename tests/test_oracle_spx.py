"""The oracle's SPX model and covariate interpolation, pinned by the
reference's own covariate tests (tests/test_core/test_helpers.py:10-31,
221-292, 374-409), scipy's normal density and the model's closed form."""

import numpy as np
import pytest
from scipy import stats

from oracle import filters as fl
from oracle import spx as ospx
from oracle import threefry as tf

F32 = np.float32


def test_interp_covars_reference_values():
    # tests/test_core/test_helpers.py:221-285
    ct = np.array([0.0, 1.0, 2.0])
    cv = np.array([10.0, 20.0, 30.0])
    t = np.array([-0.5, 0.0, 0.5, 1.0, 1.5, 2.0, 2.5])
    np.testing.assert_allclose(fl.interp_covars(t, ct, cv, "linear"), [5, 10, 15, 20, 25, 30, 35])
    np.testing.assert_allclose(fl.interp_covars(t, ct, cv, "constant"), [10, 10, 10, 20, 20, 30, 30])
    cv2 = np.array([[10.0, 100.0], [20.0, 200.0], [30.0, 300.0]])
    np.testing.assert_allclose(fl.interp_covars(t, ct, cv2, "constant"),
                               [[10, 100], [10, 100], [10, 100], [20, 200], [20, 200], [30, 300], [30, 300]])
    with pytest.raises(ValueError, match="Unsupported interpolation order"):
        fl.interp_covars(0.5, ct, cv, "spline")


def test_host_interpolation_equals_the_oracle():
    from pypomp_amd.series import interp_covars
    rng = np.random.default_rng(2)
    ct = np.sort(rng.uniform(-3, 40, 23))
    cv = rng.normal(size=(23, 3))
    t = np.concatenate([rng.uniform(-10, 50, 200), ct])
    for order in ("linear", "constant"):
        np.testing.assert_array_equal(interp_covars(t, ct, cv, order), fl.interp_covars(t, ct, cv, order))


def test_grid_covariates_follow_the_reference_cases():
    from pypomp_amd.series import Series
    # tests/test_core/test_helpers.py:374-409: nstep = 2, constant order
    kw = dict(ctimes=np.array([0.0, 2.0]), covars=np.array([[10.0], [20.0]]), order="constant")
    for s in (Series.build(np.zeros(2), [1.0, 2.0], 0.0, nstep=2, **kw),
              fl.make_series(np.zeros(2), [1.0, 2.0], 0.0, nstep=2, **kw)):
        np.testing.assert_array_equal(s.covars.ravel(), [10, 10, 10, 10, 20])
    # tests/test_core/test_helpers.py:10-31: one row per Euler step plus one
    times = np.array([0, 0.5, 1, 1.5, 2, 2.5, 3, 3.5, 4, 4.5, 5])
    ct = np.array([0, 1, 2, 3, 4, 5.0])
    cv = np.array([[1, 2], [3, 4], [5, 6], [7, 8], [9, 10], [11, 12.0]])
    h = Series.build(np.zeros(11), times, -1.0, dt=0.5, ctimes=ct, covars=cv)
    o = fl.make_series(np.zeros(11), times, -1.0, dt=0.5, ctimes=ct, covars=cv)
    assert h.covars.shape == (h.dt_ext.shape[0] + 1, 2)
    np.testing.assert_array_equal(h.covars, o.covars)
    # linear extrapolation below the table at t0 = -1: (1, 2) - (2, 2)
    np.testing.assert_allclose(h.covars[0], [-1.0, 0.0])


def test_spx_y_prev_is_the_previous_return():
    ys = np.array([0.01, -0.02, 0.005, 0.03])
    times = np.array([1.0, 2.0, 4.0, 5.0])
    ct, cv = ospx.covariate_table(ys, times)
    s = fl.make_series(ys, times, 0.0, nstep=1, ctimes=ct, covars=cv)
    np.testing.assert_array_equal(s.covars[:, 0], np.array([0.0, 0.01, -0.02, 0.005, 0.03], F32))


def test_spx_dmeas_matches_scipy():
    m = ospx.SPXModel()
    th = ospx.default_theta()
    rng = np.random.default_rng(3)
    V = rng.uniform(1e-6, 5e-4, 64).astype(F32)
    X = np.stack([V, np.full(64, F32(1105))], axis=1)
    for y in (-0.03, 0.0, 0.012):
        got = m.dmeas(np.array([y], F32), X, th, 0.0)
        exp = stats.norm.logpdf(np.float32(y), th[0] - 0.5 * V.astype(np.float64), np.sqrt(V.astype(np.float64)))
        np.testing.assert_allclose(got, exp, rtol=1e-5)


def test_spx_step_without_vol_noise_is_the_closed_form_in_both_key_modes():
    # xi = 0: dZ never reaches the state, so the @vectorized and per-particle
    # draws give the same step (the reference's own equivalence argument,
    # tests/test_core/test_vectorized_rproc.py:1-8), equal to the recursion in float64
    th = ospx.default_theta()
    th[3] = 0.0
    J = 32
    rng = np.random.default_rng(4)
    V = rng.uniform(1e-5, 3e-4, J).astype(F32)
    X = np.stack([V, rng.uniform(900, 1300, J).astype(F32)], axis=1)
    sk = tf.split(tf.key(5))[1]
    cov = np.array([0.004], F32)
    a = ospx.SPXModel(vectorized=True).rproc_step(X, th, sk, 0.0, 1.0, cov=cov)
    b = ospx.SPXModel(vectorized=False).rproc_step(X, th, tf.split(sk, J), 0.0, 1.0, cov=cov)
    np.testing.assert_array_equal(a, b)
    mu, kappa, theta = (float(th[i]) for i in range(3))
    V64, S64 = V.astype(np.float64), X[:, 1].astype(np.float64)
    dWs = (0.004 - mu + 0.5 * V64) / np.sqrt(V64)
    np.testing.assert_allclose(a[:, 1], S64 + S64 * (mu + np.sqrt(V64) * dWs), rtol=1e-6)
    np.testing.assert_allclose(a[:, 0], V64 + kappa * (theta - V64), rtol=1e-6)


def test_spx_key_modes_draw_differently_with_noise():
    th = ospx.default_theta()
    J = 16
    X = np.stack([np.full(J, F32(1e-4)), np.full(J, F32(1105))], axis=1)
    sk = tf.split(tf.key(6))[1]
    cov = np.array([0.0], F32)
    a = ospx.SPXModel(vectorized=True).rproc_step(X, th, sk, 0.0, 1.0, cov=cov)
    b = ospx.SPXModel(vectorized=False).rproc_step(X, th, tf.split(sk, J), 0.0, 1.0, cov=cov)
    assert not np.array_equal(a[:, 0], b[:, 0])
    # the vectorized draw is element j of normal(sk, (J,))
    np.testing.assert_array_equal(ospx.SPXModel(True).draw(sk, J), tf.normal(sk, J))


def test_spx_transforms_invert():
    m = ospx.SPXModel()
    th = ospx.default_theta()
    np.testing.assert_allclose(m.from_est(m.to_est(th)), th, rtol=2e-6)


def test_host_spx_transform_equals_the_oracle():
    from pypomp_amd.models import SPXModel
    rng = np.random.default_rng(7)
    x = rng.normal(scale=2.0, size=(64, 6)).astype(F32)
    np.testing.assert_array_equal(SPXModel().from_est(x), ospx.SPXModel().from_est(x))
    np.testing.assert_allclose(SPXModel().to_est(ospx.default_theta()), ospx.SPXModel().to_est(ospx.default_theta()))


def test_pomp_covariate_frame_and_errors():
    import pandas as pd
    import pypomp_amd as pp
    from pypomp_amd.models import SPXModel, spx
    d = np.load("tests/golden/spx.npz")
    P = spx(d["y"][:30], d["time"][:30])
    assert P.covar_names == ["y_prev"] and P.series.covars.shape == (31, 1)
    np.testing.assert_array_equal(P.series.covars[1:, 0], d["y"][:30].astype(F32))
    m = SPXModel()
    ys = pd.DataFrame({"y": d["y"][:30]}, index=d["time"][:30])
    with pytest.raises(ValueError, match="reads the covariates"):
        pp.Pomp(ys=ys, theta=dict(ospx.THETA), t0=0.0, model=m, nstep=1)
    with pytest.raises(ValueError, match="lacks the columns"):
        pp.Pomp(ys=ys, theta=dict(ospx.THETA), t0=0.0, model=m, nstep=1,
                covars=pd.DataFrame({"other": [0.0, 1.0]}, index=[0.0, 40.0]))
    with pytest.raises(TypeError, match="DataFrame"):
        pp.Pomp(ys=ys, theta=dict(ospx.THETA), t0=0.0, model=m, nstep=1, covars=np.zeros((2, 1)))
    Q = spx(d["y"][:30], d["time"][:30])
    assert P == Q
    R = spx(d["y"][:30] * 2, d["time"][:30])
    assert P != R
