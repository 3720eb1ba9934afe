"""Stochastic SEIR with Euler-binomial transitions, native form (epidemic workload).

States S, E, I, R, H (H = E->I transitions since the last observation, reset
at each observation); theta = (beta, amp, sigma, gamma, rho, k, N, S0f, E0, I0);
reports y ~ NegBin(mean rho H, size k).  Sub-step draws: split(key, 3), one
fast_binomial uniform per transition (S->E, E->I, I->R).
Estimation scale: log for beta, sigma, gamma, k, N, E0, I0; logit for amp, rho, S0f.
"""

import numpy as np

from .base import F32, NativeModel
from .. import _native

SEIR_DEFAULTS = {"beta": 1.5, "amp": 0.2, "sigma": 1.0, "gamma": 0.7, "rho": 0.4, "k": 10.0,
                 "N": 1e6, "S0f": 0.95, "E0": 500.0, "I0": 500.0}

_LOG = [0, 2, 3, 5, 6, 8, 9]
_LOGIT = [1, 4, 7]


class SEIRModel(NativeModel):
    kind = _native.MP_MODEL_SEIR
    S = 5
    dy = 1
    accumvars = (4,)
    param_names = ["beta", "amp", "sigma", "gamma", "rho", "k", "N", "S0f", "E0", "I0"]
    statenames = ["S", "E", "I", "R", "H"]
    obsnames = ["reports"]

    def _to_est(self, theta):
        th = np.array(theta, dtype=F32, copy=True)
        th[..., _LOG] = np.log(th[..., _LOG])
        p = th[..., _LOGIT]
        th[..., _LOGIT] = np.log(p / (F32(1) - p))
        return th.astype(F32)

    def _log_jac(self, th):
        return th[..., _LOG].sum(axis=-1) + self._log_dlogistic(th[..., _LOGIT]).sum(axis=-1)
