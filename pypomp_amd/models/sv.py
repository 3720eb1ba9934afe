"""Stochastic volatility with leverage, native form (the MOP workload).

State (v, eta); theta = (mu, phi, sigma, rho, V0, S0):
    v_0 = V0 + S0 z;  v_t = mu + phi (v_{t-1} - mu) + sigma eta_t;
    r_t | v_t, eta_t ~ N(exp(v_t/2) rho eta_t, exp(v_t)(1 - rho^2)),
so corr(eta_t, eps_t) = rho for r_t = exp(v_t/2) eps_t.
Estimation scale: phi = tanh(.), sigma = exp(.), rho = tanh(.), S0 = exp(.).
"""

import numpy as np

from .base import F32, NativeModel
from .. import _native

SV_DEFAULTS = {"mu": -9.0, "phi": 0.98, "sigma": 0.2, "rho": -0.6, "V0": -9.0, "S0": 1.0}


class SVModel(NativeModel):
    kind = _native.MP_MODEL_SV
    S = 2
    dy = 1
    param_names = ["mu", "phi", "sigma", "rho", "V0", "S0"]
    statenames = ["v", "eta"]
    obsnames = ["r"]

    def _to_est(self, theta):
        th = np.array(theta, dtype=F32, copy=True)
        th[..., 1] = np.arctanh(th[..., 1])
        th[..., 2] = np.log(th[..., 2])
        th[..., 3] = np.arctanh(th[..., 3])
        th[..., 5] = np.log(th[..., 5])
        return th

    def _log_jac(self, th):
        return self._log_dtanh(th[..., 1]) + th[..., 2] + self._log_dtanh(th[..., 3]) + th[..., 5]
