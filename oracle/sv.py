"""Stochastic volatility with leverage restated in NumPy float32 -- TEST ORACLE ONLY.

Model defined here (the MOP workload; the reference's spx model,
pypomp/models/spx.py:66-126, is the pattern).  State (v, eta), theta =
(mu, phi, sigma, rho, V0, S0):
    v_0 = V0 + S0 z, eta_0 = 0;  v_t = mu + phi (v_{t-1} - mu) + sigma eta_t,
    r_t | v_t, eta_t ~ N(exp(v_t/2) rho eta_t, exp(v_t)(1 - rho^2)).
One normal per particle per step from its key.  Estimation scale:
phi = tanh, sigma = exp, rho = tanh, S0 = exp.

Pinned: dmeas against scipy.stats.norm.logpdf of the conditional law; the
marginal of r_t given v_t against a direct simulation of the correlated pair
(tests/test_oracle_sv.py).
"""

import numpy as np

from . import filters as fl
from . import threefry as tf
from .detmath import d_exp, d_log, d_tanh

F32 = np.float32
PARAM_NAMES = ["mu", "phi", "sigma", "rho", "V0", "S0"]
DEFAULTS = {"mu": -9.0, "phi": 0.98, "sigma": 0.2, "rho": -0.6, "V0": -9.0, "S0": 1.0}


class SVModel:
    S = 2
    accumvars = ()
    param_names = PARAM_NAMES

    def from_est(self, th):
        th = np.array(th, dtype=F32, copy=True)
        th[..., 1] = d_tanh(th[..., 1])
        th[..., 2] = d_exp(th[..., 2])
        th[..., 3] = d_tanh(th[..., 3])
        th[..., 5] = d_exp(th[..., 5])
        return th

    def to_est(self, th):
        th = np.array(th, dtype=F32, copy=True)
        th[..., 1] = np.arctanh(th[..., 1])
        th[..., 2] = np.log(th[..., 2])
        th[..., 3] = np.arctanh(th[..., 3])
        th[..., 5] = np.log(th[..., 5])
        return th

    @staticmethod
    def _cols(theta, J):
        th = np.asarray(theta, dtype=F32)
        if th.ndim == 1:
            th = np.broadcast_to(th, (J, th.shape[0]))
        return [th[:, i] for i in range(th.shape[1])]

    def rinit(self, theta, keys, t0=0.0):
        J = keys.shape[0]
        mu, phi, sig, rho, V0, S0 = self._cols(theta, J)
        z = tf.normal_many(keys, 1)[:, 0]
        return np.stack([(V0 + S0 * z).astype(F32), np.zeros(J, F32)], axis=1)

    def rproc_step(self, X, theta, keys, t, dt):
        J = X.shape[0]
        mu, phi, sig, rho, V0, S0 = self._cols(theta, J)
        z = tf.normal_many(keys, 1)[:, 0]
        v = (mu + phi * (X[:, 0] - mu) + sig * z).astype(F32)
        return np.stack([v, z], axis=1).astype(F32)

    def dmeas(self, y, X, theta, t):
        J = X.shape[0]
        mu, phi, sig, rho, V0, S0 = self._cols(theta, J)
        omr2 = (F32(1) - (rho * rho).astype(F32)).astype(F32)
        c0 = (F32(-0.918938533) - (F32(0.5) * d_log(omr2)).astype(F32)).astype(F32)
        inv = (F32(1) / omr2).astype(F32)
        yv = F32(np.asarray(y, dtype=F32)[0])
        u = ((yv * d_exp((F32(-0.5) * X[:, 0]).astype(F32))).astype(F32) - (rho * X[:, 1]).astype(F32)).astype(F32)
        return ((c0 - (F32(0.5) * X[:, 0]).astype(F32)).astype(F32)
                - (((F32(0.5) * u).astype(F32) * u).astype(F32) * inv).astype(F32)).astype(F32)

    def rmeas(self, X, theta, keys, t):
        J = X.shape[0]
        mu, phi, sig, rho, V0, S0 = self._cols(theta, J)
        e = tf.normal_many(keys, 1)[:, 0]
        sd = np.sqrt((F32(1) - (rho * rho).astype(F32)).astype(F32))
        r = (d_exp((F32(0.5) * X[:, 0]).astype(F32))
             * ((rho * X[:, 1]).astype(F32) + (sd * e).astype(F32)).astype(F32))
        return r.astype(F32)[:, None]


def default_theta():
    return np.array([DEFAULTS[n] for n in PARAM_NAMES], dtype=F32)


def simulate_data(seed, T, theta=None):
    model = SVModel()
    theta = default_theta() if theta is None else np.asarray(theta, dtype=F32)
    ser = fl.make_series(np.zeros((T, 1)), np.arange(1, T + 1, dtype=float), 0.0, nstep=1)
    k = tf.split(fl.pomp_new_key(tf.key(seed)), 1)[0]
    _, Y = fl.simulate(model, theta, ser, k)
    return Y[:, 0, :].astype(F32), theta
