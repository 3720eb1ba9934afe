"""Linear-Gaussian model components restated in NumPy float32 -- TEST ORACLE ONLY.

Follows pypomp/models/linear_gaussian.py:139-191 (components), :221-280
(defaults), :337-451 (constructor and parameter naming).  State x has dimension
dx, observation y dimension dy:

    x0 = X0 + L0 z      (reference: L0 = L_Q; a separate initial Cholesky
                         factor P is also allowed, which workloads with
                         x0 ~ N(0, I) need -- see ``LGSpec.init_chol``)
    x' = A x + L_Q z
    y  ~ N(C x, L_R L_R^T)

Random draw order per particle (one key per particle): ``normal(key, (dx,))``
for rinit/rproc and ``normal(key, (dy,))`` for rmeas, exactly as the reference.
"""

from dataclasses import dataclass

import numpy as np

from . import threefry as tf
from .detmath import d_exp, d_log

F32 = np.float32


def matrix_names(name, nrow, ncol):
    return [f"{name}{i}{j}" for i in range(1, nrow + 1) for j in range(1, ncol + 1)]


def tril_names(name, d):
    return [f"{name}{i}{j}" for i in range(1, d + 1) for j in range(1, i + 1)]


@dataclass
class LGSpec:
    dx: int
    dy: int
    separate_init: bool = False  # adds P{i}{j} Cholesky entries for x0's covariance

    @property
    def param_names(self):
        names = (matrix_names("A", self.dx, self.dx) + matrix_names("C", self.dy, self.dx)
                 + tril_names("Q", self.dx) + tril_names("R", self.dy)
                 + [f"X0_{i}" for i in range(1, self.dx + 1)])
        if self.separate_init:
            names += tril_names("P", self.dx)
        return names

    @property
    def statenames(self):
        return [f"X{i}" for i in range(1, self.dx + 1)]

    @property
    def obsnames(self):
        return [f"Y{j}" for j in range(1, self.dy + 1)]

    def unpack(self, theta):
        """theta (..., d) float32 -> A, C, LQ, LR, X0, L0 with leading batch dims."""
        theta = np.asarray(theta, dtype=F32)
        dx, dy = self.dx, self.dy
        lead = theta.shape[:-1]
        o = 0
        A = theta[..., o:o + dx * dx].reshape(lead + (dx, dx)); o += dx * dx
        C = theta[..., o:o + dy * dx].reshape(lead + (dy, dx)); o += dy * dx

        def tril(o, d):
            L = np.zeros(lead + (d, d), dtype=F32)
            for i in range(d):
                for j in range(i + 1):
                    L[..., i, j] = theta[..., o]
                    o += 1
            return L, o

        LQ, o = tril(o, dx)
        LR, o = tril(o, dy)
        X0 = theta[..., o:o + dx]; o += dx
        if self.separate_init:
            L0, o = tril(o, dx)
        else:
            L0 = LQ
        return A, C, LQ, LR, X0, L0

    # -- parameter transforms (reference: linear_gaussian.py:_to_est/_from_est):
    #    log on the diagonal Cholesky entries of Q and R only.
    def _diag_mask(self):
        names = self.param_names
        return np.array([len(n) == 3 and n[0] in ("Q", "R") and n[1] == n[2] for n in names])

    def to_est(self, theta):
        theta = np.array(theta, dtype=F32, copy=True)
        m = self._diag_mask()
        theta[..., m] = np.log(theta[..., m])  # host-side transform, as the host layer
        return theta

    def from_est(self, theta):
        theta = np.array(theta, dtype=F32, copy=True)
        m = self._diag_mask()
        theta[..., m] = d_exp(theta[..., m])
        return theta


def chol_logdet(LR, d):
    """float32 Cholesky factor of LR LR^T and the sum of log of its diagonal, in the
    engine's operation order (sequential sums, deterministic log)."""
    lead = LR.shape[:-2]
    cov = np.zeros(lead + (d, d), dtype=F32)
    for i in range(d):
        for j in range(d):
            s = (LR[..., i, 0] * LR[..., j, 0]).astype(F32)
            for k in range(1, d):
                s = (s + (LR[..., i, k] * LR[..., j, k]).astype(F32)).astype(F32)
            cov[..., i, j] = s
    L = np.zeros(lead + (d, d), dtype=F32)
    logdet = np.zeros(lead, dtype=F32)
    for j in range(d):
        s = cov[..., j, j]
        for k in range(j):
            s = (s - (L[..., j, k] * L[..., j, k]).astype(F32)).astype(F32)
        djj = np.sqrt(s).astype(F32)
        L[..., j, j] = djj
        logdet = (logdet + d_log(djj)).astype(F32)
        for i in range(j + 1, d):
            t = cov[..., i, j]
            for k in range(j):
                t = (t - (L[..., i, k] * L[..., j, k]).astype(F32)).astype(F32)
            L[..., i, j] = (t / djj).astype(F32)
    return L, logdet


def _matvec(M, v):
    """M (..., r, c) @ v (..., c) in float32 with sequential accumulation."""
    r, c = M.shape[-2], M.shape[-1]
    out = np.zeros(np.broadcast_shapes(M.shape[:-2], v.shape[:-1]) + (r,), dtype=F32)
    for i in range(r):
        acc = (M[..., i, 0] * v[..., 0]).astype(F32)
        for k in range(1, c):
            acc = (acc + (M[..., i, k] * v[..., k]).astype(F32)).astype(F32)
        out[..., i] = acc
    return out


class LGModel:
    """Per-particle LG components over arrays of particles (J, dx)."""

    def __init__(self, spec: LGSpec):
        self.spec = spec
        self.S = spec.dx
        self.accumvars = ()

    @property
    def param_names(self):
        return self.spec.param_names

    def from_est(self, theta):
        return self.spec.from_est(theta)

    def to_est(self, theta):
        return self.spec.to_est(theta)

    def rinit(self, theta, keys, t0=0.0):
        A, C, LQ, LR, X0, L0 = self.spec.unpack(theta)
        z = tf.normal_many(keys, self.spec.dx)
        return (X0 + _matvec(L0, z)).astype(F32)

    def rproc_step(self, X, theta, keys, t, dt):
        A, C, LQ, LR, X0, L0 = self.spec.unpack(theta)
        z = tf.normal_many(keys, self.spec.dx)
        return (_matvec(A, X) + _matvec(LQ, z)).astype(F32)

    def dmeas(self, y, X, theta, t):
        """multivariate normal log density of y given C x, cov L_R L_R^T, through the
        float32 Cholesky factor of the covariance (the reference's
        jax.scipy.stats.multivariate_normal.logpdf path)."""
        A, C, LQ, LR, X0, L0 = self.spec.unpack(theta)
        dy = self.spec.dy
        mean = _matvec(C, X)
        L, logdet = chol_logdet(LR, dy)
        r = (np.asarray(y, dtype=F32) - mean).astype(F32)
        # forward substitution L s = r
        s = np.zeros_like(r)
        for i in range(dy):
            acc = r[..., i]
            for k in range(i):
                acc = (acc - L[..., i, k] * s[..., k]).astype(F32)
            s[..., i] = (acc / L[..., i, i]).astype(F32)
        quad = np.zeros(s.shape[:-1], dtype=F32)
        for i in range(dy):
            quad = (quad + (s[..., i] * s[..., i]).astype(F32)).astype(F32)
        hn = ((F32(0.5) * F32(dy)).astype(F32) * F32(1.83787707)).astype(F32)
        return (((F32(-0.5) * quad).astype(F32) - hn).astype(F32) - logdet).astype(F32)

    def rmeas(self, X, theta, keys, t):
        A, C, LQ, LR, X0, L0 = self.spec.unpack(theta)
        z = tf.normal_many(keys, self.spec.dy)
        return (_matvec(C, X) + _matvec(LR, z)).astype(F32)


def default_theta(dx=2, dy=None, A=None, C=None, Q=None, R=None, X0=None, P0=None):
    """Reference defaults (linear_gaussian.py:221-280) as a float64 parameter dict."""
    dy = dx if dy is None else dy
    if A is None:
        A = np.zeros((dx, dx))
        c, s = np.cos(0.2), np.sin(0.2)
        i = 0
        while i + 1 < dx:
            A[i:i + 2, i:i + 2] = [[c, -s], [s, c]]
            i += 2
        if i < dx:
            A[i, i] = c
    if C is None:
        C = np.eye(dy, dx)

    def dcov(d, diag, off):
        M = np.eye(d) * diag
        if d > 1:
            M = M + (np.ones((d, d)) - np.eye(d)) * (off / (d - 1))
        return M

    Q = dcov(dx, 1.0, 0.02) / 100 if Q is None else np.asarray(Q, float)
    R = dcov(dy, 1.0, 0.1) / 10 if R is None else np.asarray(R, float)
    X0 = np.zeros(dx) if X0 is None else np.asarray(X0, float)
    LQ, LR = np.linalg.cholesky(Q), np.linalg.cholesky(R)
    theta = {}
    for n in matrix_names("A", dx, dx):
        theta[n] = float(A[int(n[-2]) - 1, int(n[-1]) - 1])
    for n in matrix_names("C", dy, dx):
        theta[n] = float(C[int(n[-2]) - 1, int(n[-1]) - 1])
    for n in tril_names("Q", dx):
        theta[n] = float(LQ[int(n[-2]) - 1, int(n[-1]) - 1])
    for n in tril_names("R", dy):
        theta[n] = float(LR[int(n[-2]) - 1, int(n[-1]) - 1])
    for i in range(dx):
        theta[f"X0_{i + 1}"] = float(X0[i])
    if P0 is not None:
        LP = np.linalg.cholesky(np.asarray(P0, float))
        for n in tril_names("P", dx):
            theta[n] = float(LP[int(n[-2]) - 1, int(n[-1]) - 1])
    return theta
