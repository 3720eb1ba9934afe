"""S&P 500 stochastic-volatility model, native form (pypomp/models/spx.py).

State (V, S); theta = (mu, kappa, theta, xi, rho, V_0); one covariate, y_prev
(the previous log return).  Per Euler sub-step, with dZ standard normal:
    dWs = (y_prev - mu + V/2) / sqrt(V),  dWv = rho dWs + sqrt(1 - rho^2) dZ,
    S <- S + S (mu + sqrt(max(V, 1e-32)) dWs),
    V <- max(V + kappa (theta - V) + xi sqrt(V) dWv, 1e-32);
y | V ~ N(mu - V/2, sqrt(V)).  ``vectorized=True`` is the reference's
``@vectorized`` rproc (one normal draw of shape (J,) per sub-step key),
``False`` its per-particle rproc (``spx(_pre_vectorized=False)``).
Estimation scale: log of mu, kappa, theta, xi, V_0; rho = -1 + 2/(1 + exp(-x)).
"""

import numpy as np
import pandas as pd

from .base import F32, NativeModel
from .. import _native

SPX_THETA = {"mu": 3.68e-4, "kappa": 3.14e-2, "theta": 1.12e-4, "xi": 2.27e-3, "rho": -7.38e-1,
             "V_0": 7.66e-3 ** 2}

_LOG = [0, 1, 2, 3, 5]


class SPXModel(NativeModel):
    kind = _native.MP_MODEL_SPX
    S = 2
    dy = 1
    param_names = ["mu", "kappa", "theta", "xi", "rho", "V_0"]
    statenames = ["V", "S"]
    obsnames = ["y"]
    covar_names = ["y_prev"]

    def __init__(self, vectorized=True):
        self.vectorized = bool(vectorized)

    def _to_est(self, theta):
        th = np.array(theta, dtype=F32, copy=True)
        th[..., _LOG] = np.log(th[..., _LOG])
        r = th[..., 4]
        th[..., 4] = np.log((F32(1) + r) / (F32(1) - r))
        return th.astype(F32)

    def _log_jac(self, th):
        # d rho / dx = 2 s (1 - s), s = logistic(x)
        return th[..., _LOG].sum(axis=-1) + np.log(2.0) + self._log_dlogistic(th[..., 4])


def spx_covars(ys, times):
    """The reference's covariate frame for the model (spx.py:36-40): the
    returns themselves as ``y_prev`` at their own times, with a 0 one time unit
    before the first, so the sub-step into observation k reads return k - 1."""
    ys = np.asarray(ys, dtype=np.float64).reshape(-1)
    times = np.asarray(times, dtype=np.float64).reshape(-1)
    first = float(times[0]) - 1.0
    return pd.DataFrame({"y_prev": np.concatenate([[0.0], ys])}, index=np.concatenate([[first], times]))


def spx(ys, times, theta=None, vectorized=True, t0=0.0):
    """A ``Pomp`` of the model over a return series ``ys`` at ``times``, built
    as the reference's ``spx()`` builds its own (spx.py:154-196: t0 = 0,
    nstep = 1, the covariate frame above)."""
    from ..pomp import Pomp
    m = SPXModel(vectorized=vectorized)
    ys = np.asarray(ys, dtype=np.float64).reshape(-1)
    frame = pd.DataFrame({"y": ys}, index=np.asarray(times, dtype=np.float64))
    return Pomp(ys=frame, theta=dict(SPX_THETA) if theta is None else theta, t0=t0, model=m, nstep=1,
                covars=spx_covars(ys, times))
