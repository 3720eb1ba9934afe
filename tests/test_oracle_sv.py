"""The stochastic-volatility oracle (oracle/sv.py): the measurement density is
the conditional normal law of the return given (v, eta), and simulated returns
have the model's leverage correlation."""

import numpy as np
from scipy import stats

from oracle import filters as fl, sv, threefry as tf

F32 = np.float32


def test_dmeas_is_conditional_normal():
    model = sv.SVModel()
    th = sv.default_theta()
    rng = np.random.default_rng(0)
    J = 256
    X = np.stack([rng.uniform(-11, -7, J), rng.normal(size=J)], axis=1).astype(F32)
    mu, phi, sig, rho, V0, S0 = (np.float64(v) for v in th)
    for y in (-0.03, 0.0, 0.012):
        got = model.dmeas(np.array([y], F32), X, th, 1.0).astype(np.float64)
        v, eta = X[:, 0].astype(np.float64), X[:, 1].astype(np.float64)
        ref = stats.norm.logpdf(y, loc=np.exp(v / 2) * rho * eta, scale=np.exp(v / 2) * np.sqrt(1 - rho * rho))
        assert np.max(np.abs(got - ref) / np.maximum(1, np.abs(ref))) <= 1e-5


def test_simulated_returns_have_leverage_correlation():
    model = sv.SVModel()
    th = sv.default_theta()
    ser = fl.make_series(np.zeros((400, 1)), np.arange(1, 401.0), 0.0, nstep=1)
    X, Y = fl.simulate(model, th, ser, tf.key(9), nsim=64)
    eta = X[1:, :, 1].astype(np.float64)
    eps = Y[:, :, 0].astype(np.float64) / np.exp(X[1:, :, 0].astype(np.float64) / 2)
    c = np.corrcoef(eta.ravel(), eps.ravel())[0, 1]
    assert abs(c - float(th[3])) < 0.03
    # log-variance is stationary around mu with sd sigma / sqrt(1 - phi^2)
    v = X[200:, :, 0].astype(np.float64)
    assert abs(v.mean() - float(th[0])) < 0.2
    assert abs(v.std() - float(th[2]) / np.sqrt(1 - float(th[1]) ** 2)) < 0.25
