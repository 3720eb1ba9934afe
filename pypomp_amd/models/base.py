"""Native model specifications: the host half of each device model functor.

A model carries its parameter/state/observation names, its estimation-scale
transforms (host mirror of the device ``from_est``, float32) and the C
descriptor handed to the library.  The plugin slots of ``Pomp``
(rinit/rproc/dmeas/rmeas, pypomp/core/model_mechanics.py:361-760) are filled
with ``Component`` handles of one native model.
"""

import numpy as np

from .. import _native

F32 = np.float32


class Component:
    """Handle for one plugin slot (rinit / rproc / dmeas / rmeas) of a native model."""

    def __init__(self, model, role):
        self.model = model
        self.role = role

    def __repr__(self):
        return f"<{type(self.model).__name__}.{self.role}>"

    def __eq__(self, other):
        return isinstance(other, Component) and self.role == other.role and self.model == other.model

    def __hash__(self):
        return hash((self.role, self.model))


class NativeModel:
    kind = 0
    S = 0
    dy = 1
    dx = 0
    lg_sep_init = 0
    accumvars = ()
    # True: parameters are estimated on their natural scale (the reference's
    # ParTrans() with no transforms, core/par_trans.py:70-111)
    identity_trans = False
    # True: the model's rproc is @vectorized (one draw per Euler sub-step for
    # all particles, core/model_mechanics.py:311-355); models that define it only
    vectorized = False
    # covariates the model reads, by name (the reference's covar_names)
    covar_names: list = []

    param_names: list = []
    statenames: list = []
    obsnames: list = []

    @property
    def d(self):
        return len(self.param_names)

    def __eq__(self, other):
        # one device model and the same sizes, names and transform choice
        return type(self) is type(other) and self.__dict__ == other.__dict__

    def __hash__(self):
        return hash((type(self).__name__, tuple(self.param_names)))

    @property
    def rinit(self):
        return Component(self, "rinit")

    @property
    def rproc(self):
        return Component(self, "rproc")

    @property
    def dmeas(self):
        return Component(self, "dmeas")

    @property
    def rmeas(self):
        return Component(self, "rmeas")

    def desc(self):
        mask = 0
        for a in self.accumvars:
            mask |= 1 << int(a)
        return _native.ModelDesc(kind=self.kind, S=self.S, d=self.d, dy=self.dy, dx=self.dx,
                                 lg_sep_init=self.lg_sep_init, accum_mask=mask,
                                 identity_trans=int(bool(self.identity_trans)),
                                 vectorized=int(bool(self.vectorized)))

    # host mirrors of the device transforms (float32)
    def to_est(self, theta):
        th = np.asarray(theta, dtype=F32)
        return th.copy() if self.identity_trans else self._to_est(th)

    def from_est(self, theta):
        """Estimation -> natural scale with the device's own arithmetic
        (``mp_host_from_est``), bit-identical to the kernels' transform."""
        th = np.array(theta, dtype=F32, copy=True, order="C")
        if self.identity_trans or th.size == 0:
            return th
        import ctypes as C
        flat = th.reshape(-1, self.d)
        _native.check(_native.lib().mp_host_from_est(self.desc(), flat.ctypes.data_as(C.c_void_p), flat.shape[0]))
        return flat.reshape(th.shape)

    def _to_est(self, theta):
        return np.asarray(theta, dtype=F32).copy()

    def log_jac_from_est(self, theta_est):
        """log |det d from_est / d theta_est| over the last axis, float32: the
        change-of-variables term the reference adds to an estimation-scale
        prior (core/model_mechanics.py:841-862).  The transforms act
        elementwise, so the Jacobian is diagonal."""
        th = np.asarray(theta_est, dtype=np.float64)
        if self.identity_trans:
            return np.zeros(th.shape[:-1], dtype=F32)
        return self._log_jac(th).astype(F32)

    def _log_jac(self, th):
        return np.zeros(th.shape[:-1])

    @staticmethod
    def _log_dtanh(x):
        # log(1 - tanh(x)^2) = 2 (log 2 - |x| - log1p(exp(-2|x|)))
        a = np.abs(x)
        return 2.0 * (np.log(2.0) - a - np.log1p(np.exp(-2.0 * a)))

    @staticmethod
    def _log_dlogistic(x):
        # log(s (1 - s)), s = 1 / (1 + exp(-x))
        a = np.abs(x)
        return -a - 2.0 * np.log1p(np.exp(-a))

    def theta_array(self, theta):
        """dict (or sequence of dicts) of natural-scale values -> (n, d) float32."""
        if isinstance(theta, dict):
            theta = [theta]
        if isinstance(theta, (list, tuple)) and theta and isinstance(theta[0], dict):
            missing = set(self.param_names) - set(theta[0])
            if missing:
                raise ValueError(f"theta is missing parameters {sorted(missing)}")
            extra = set(theta[0]) - set(self.param_names)
            if extra:
                raise ValueError(f"theta has unknown parameters {sorted(extra)}")
            return np.array([[float(t[n]) for n in self.param_names] for t in theta], dtype=F32)
        arr = np.asarray(theta, dtype=F32)
        if arr.ndim == 1:
            arr = arr[None, :]
        if arr.shape[-1] != self.d:
            raise ValueError(f"theta has {arr.shape[-1]} entries, model needs {self.d}")
        return arr


class ParTrans:
    """Identity parameter transform: the reference's ``ParTrans()`` with no
    transforms (core/par_trans.py:70-111), usable as a panel-level transform."""

    def to_est(self, theta):
        return np.array(theta, dtype=F32, copy=True)

    def from_est(self, theta):
        return np.array(theta, dtype=F32, copy=True)
