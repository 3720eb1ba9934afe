"""The reference's S&P 500 volatility model restated in NumPy float32 -- TEST ORACLE ONLY.

Follows pypomp/models/spx.py: rinit (:56-63), the @vectorized rproc (:66-96)
and the per-particle rproc (:99-126), which differ only in how dZ is drawn
(normal(key, (J,)) from the sub-step key, or normal(key) from each particle's
key), dmeas (:129-137) in jax.scipy.stats.norm.logpdf's arithmetic
(log(2 pi s^2) + (y - loc)^2 / s^2) / -2, and the transforms (:140-151).  The
reference defines no rmeas; ``rmeas`` samples the measurement density.

Pinned: dmeas against scipy.stats.norm.logpdf; a step with xi = 0 (no noise
reaches the state) is the same in both key modes and equals the closed-form
recursion; the transforms invert each other (tests/test_oracle_spx.py).
"""

import numpy as np

from . import threefry as tf
from .detmath import d_exp, d_log

F32 = np.float32
PARAM_NAMES = ["mu", "kappa", "theta", "xi", "rho", "V_0"]
THETA = {"mu": 3.68e-4, "kappa": 3.14e-2, "theta": 1.12e-4, "xi": 2.27e-3, "rho": -7.38e-1,
         "V_0": 7.66e-3 ** 2}
_LOG = [0, 1, 2, 3, 5]


class SPXModel:
    S = 2
    accumvars = ()
    uses_covars = True
    param_names = PARAM_NAMES

    def __init__(self, vectorized=True):
        self.vectorized = bool(vectorized)

    def from_est(self, th):
        th = np.array(th, dtype=F32, copy=True)
        for i in _LOG:
            th[..., i] = d_exp(th[..., i])
        e = d_exp((-th[..., 4]).astype(F32))
        th[..., 4] = (F32(-1) + (F32(2) / (F32(1) + e).astype(F32)).astype(F32)).astype(F32)
        return th

    def to_est(self, th):
        th = np.array(th, dtype=F32, copy=True)
        th[..., _LOG] = np.log(th[..., _LOG])
        r = th[..., 4]
        th[..., 4] = np.log((F32(1) + r) / (F32(1) - r))
        return th.astype(F32)

    @staticmethod
    def _cols(theta, J):
        th = np.asarray(theta, dtype=F32)
        if th.ndim == 1:
            th = np.broadcast_to(th, (J, th.shape[0]))
        return [th[:, i] for i in range(th.shape[1])]

    def rinit(self, theta, keys, t0=0.0, cov=None):
        J = keys.shape[0]
        V0 = self._cols(theta, J)[5]
        return np.stack([V0.astype(F32), np.full(J, F32(1105))], axis=1)

    def draw(self, keys, J):
        """dZ of every particle: element j of normal(sk, (J,)) for the
        @vectorized rproc, normal(key_j) for the per-particle one."""
        if self.vectorized:
            return tf.normal(keys, J).astype(F32)
        return tf.normal_many(keys, 1)[:, 0].astype(F32)

    def step_z(self, X, theta, dZ, cov):
        J = X.shape[0]
        mu, kappa, th, xi, rho, V0 = self._cols(theta, J)
        V, S = X[:, 0], X[:, 1]
        y_prev = F32(np.asarray(cov, dtype=F32).reshape(-1)[0])
        sV = np.sqrt(V).astype(F32)
        dWs = ((((y_prev - mu).astype(F32) + (F32(0.5) * V).astype(F32)).astype(F32)) / sV).astype(F32)
        sq = np.sqrt((F32(1) - (rho * rho).astype(F32)).astype(F32)).astype(F32)
        dWv = ((rho * dWs).astype(F32) + (sq * dZ).astype(F32)).astype(F32)
        Sn = (S + (S * (mu + (np.sqrt(np.maximum(V, F32(1e-32))).astype(F32) * dWs).astype(F32)).astype(F32))
              .astype(F32)).astype(F32)
        Vn = ((V + (kappa * (th - V).astype(F32)).astype(F32)).astype(F32)
              + ((xi * sV).astype(F32) * dWv).astype(F32)).astype(F32)
        Vn = np.maximum(Vn, F32(1e-32)).astype(F32)
        return np.stack([Vn, Sn], axis=1).astype(F32)

    def rproc_step(self, X, theta, keys, t, dt, cov=None):
        return self.step_z(X, theta, self.draw(keys, X.shape[0]), cov)

    def dmeas(self, y, X, theta, t, cov=None):
        J = X.shape[0]
        mu = self._cols(theta, J)[0]
        V = X[:, 0]
        loc = (mu - (F32(0.5) * V).astype(F32)).astype(F32)
        s = np.sqrt(V).astype(F32)
        s2 = (s * s).astype(F32)
        r = (F32(np.asarray(y, dtype=F32).reshape(-1)[0]) - loc).astype(F32)
        ln = d_log((F32(6.2831855) * s2).astype(F32))
        q = ((r * r).astype(F32) / s2).astype(F32)
        return ((ln + q).astype(F32) / F32(-2)).astype(F32)

    def rmeas(self, X, theta, keys, t, cov=None):
        J = X.shape[0]
        mu = self._cols(theta, J)[0]
        V = X[:, 0]
        e = tf.normal_many(keys, 1)[:, 0].astype(F32)
        y = ((mu - (F32(0.5) * V).astype(F32)).astype(F32) + (np.sqrt(V).astype(F32) * e).astype(F32)).astype(F32)
        return y[:, None]


def default_theta():
    return np.array([THETA[n] for n in PARAM_NAMES], dtype=F32)


def covariate_table(ys, times):
    """The reference's covariate frame (spx.py:36-40) as (ctimes, values):
    the returns at their own times, with 0 one time unit before the first."""
    ys = np.asarray(ys, dtype=np.float64).reshape(-1)
    times = np.asarray(times, dtype=np.float64).reshape(-1)
    return np.concatenate([[times[0] - 1.0], times]), np.concatenate([[0.0], ys])[:, None]
