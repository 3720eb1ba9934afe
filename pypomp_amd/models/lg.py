"""Linear-Gaussian model (pypomp/models/linear_gaussian.py), native form.

Parameter naming, defaults and transforms follow the reference
(linear_gaussian.py:53-92 names, :221-280 defaults, :197-218 transforms).  The
device functor supports dx, dy <= 8.  ``P0`` optionally gives the initial
state its own covariance (Cholesky entries ``P{i}{j}`` appended to theta); the
reference draws x0 from N(X0, Q).
"""

import numpy as np

from .base import F32, NativeModel
from .. import _native


def _matrix_names(name, nrow, ncol):
    return [f"{name}{i}{j}" for i in range(1, nrow + 1) for j in range(1, ncol + 1)]


def _tril_names(name, d):
    return [f"{name}{i}{j}" for i in range(1, d + 1) for j in range(1, i + 1)]


class LGModel(NativeModel):
    kind = _native.MP_MODEL_LG

    def __init__(self, dx=2, dy=None, separate_init=False):
        dy = dx if dy is None else dy
        if not (1 <= dx <= 8 and 1 <= dy <= 8):
            raise ValueError("the native LG model supports dimensions 1..8")
        self.dx, self.dy, self.S = int(dx), int(dy), int(dx)
        self.lg_sep_init = int(bool(separate_init))
        self.param_names = (_matrix_names("A", dx, dx) + _matrix_names("C", dy, dx)
                            + _tril_names("Q", dx) + _tril_names("R", dy)
                            + [f"X0_{i}" for i in range(1, dx + 1)])
        if separate_init:
            self.param_names += _tril_names("P", dx)
        self.statenames = [f"X{i}" for i in range(1, dx + 1)]
        self.obsnames = [f"Y{j}" for j in range(1, dy + 1)]

    def _diag_mask(self):
        return np.array([len(n) == 3 and n[0] in ("Q", "R") and n[1] == n[2] for n in self.param_names])

    def _to_est(self, theta):
        th = np.array(theta, dtype=F32, copy=True)
        m = self._diag_mask()
        th[..., m] = np.log(th[..., m])
        return th

    def _log_jac(self, th):
        return th[..., self._diag_mask()].sum(axis=-1)


def _default_A(dx):
    A = np.zeros((dx, dx))
    c, s = float(np.cos(0.2)), float(np.sin(0.2))
    i = 0
    while i + 1 < dx:
        A[i:i + 2, i:i + 2] = np.array([[c, -s], [s, c]])
        i += 2
    if i < dx:
        A[i, i] = c
    return A


def _default_cov(d, diag, offdiag):
    M = np.eye(d) * diag
    if d > 1:
        M = M + (np.ones((d, d)) - np.eye(d)) * (offdiag / (d - 1))
    return M


def lg_theta(A=None, C=None, Q=None, R=None, X0=None, P0=None, dx=None, dy=None):
    """Natural-scale parameter dict with the reference's defaults."""
    if dx is None:
        dx = (np.asarray(A).shape[0] if A is not None else
              np.asarray(Q).shape[0] if Q is not None else
              np.asarray(X0).shape[0] if X0 is not None else 2)
    if dy is None:
        dy = (np.asarray(C).shape[0] if C is not None else
              np.asarray(R).shape[0] if R is not None else dx)
    A = _default_A(dx) if A is None else np.asarray(A, float)
    C = np.eye(dy, dx) if C is None else np.asarray(C, float)
    Q = _default_cov(dx, 1.0, 0.02) / 100 if Q is None else np.asarray(Q, float)
    R = _default_cov(dy, 1.0, 0.1) / 10 if R is None else np.asarray(R, float)
    X0 = np.zeros(dx) if X0 is None else np.asarray(X0, float)
    for name, mat in (("Q", Q), ("R", R)):
        if not np.allclose(mat, mat.T, atol=1e-8, rtol=1e-5):
            raise ValueError(f"Covariance matrix {name} must be symmetric.")
        try:
            np.linalg.cholesky(mat)
        except np.linalg.LinAlgError as e:
            raise ValueError(f"Covariance matrix {name} must be positive-definite.") from e
    LQ, LR = np.linalg.cholesky(Q), np.linalg.cholesky(R)
    theta = {}
    for n in _matrix_names("A", dx, dx):
        theta[n] = float(A[int(n[-2]) - 1, int(n[-1]) - 1])
    for n in _matrix_names("C", dy, dx):
        theta[n] = float(C[int(n[-2]) - 1, int(n[-1]) - 1])
    for n in _tril_names("Q", dx):
        theta[n] = float(LQ[int(n[-2]) - 1, int(n[-1]) - 1])
    for n in _tril_names("R", dy):
        theta[n] = float(LR[int(n[-2]) - 1, int(n[-1]) - 1])
    for i in range(dx):
        theta[f"X0_{i + 1}"] = float(X0[i])
    if P0 is not None:
        LP = np.linalg.cholesky(np.asarray(P0, float))
        for n in _tril_names("P", dx):
            theta[n] = float(LP[int(n[-2]) - 1, int(n[-1]) - 1])
    return theta, dx, dy


def LG(T=4, A=None, C=None, Q=None, R=None, X0=None, P0=None, ys=None, times=None, key=None):
    """A ``Pomp`` for the linear-Gaussian model (reference constructor:
    linear_gaussian.py:337-451).  Observations: ``ys`` (T, dy) if given;
    otherwise they are simulated on the device with the reference's key tree
    from ``key`` (default ``key(1)``): ``simulate(key, nsim=1, as_pomp=True)``
    of a model with zero observations (linear_gaussian.py:392, 425-449)."""
    from .. import random as prandom
    from ..pomp import Pomp
    theta, dx, dy = lg_theta(A=A, C=C, Q=Q, R=R, X0=X0, P0=P0)
    model = LGModel(dx, dy, separate_init=P0 is not None)
    if times is None:
        times = np.arange(1, T + 1, dtype=float) if ys is None else np.arange(1, len(ys) + 1, dtype=float)
    if ys is not None:
        return Pomp(ys=np.asarray(ys, dtype=np.float32).reshape(len(times), dy), theta=theta, t0=0.0,
                    times=times, nstep=1, model=model)
    tmp = Pomp(ys=np.zeros((len(times), dy), np.float32), theta=theta, t0=0.0, times=times, nstep=1, model=model)
    return tmp.simulate(key=prandom.key(1) if key is None else key, nsim=1, as_pomp=True)
