"""Exact FP64 Kalman-filter log-likelihood of a linear-Gaussian POMP -- TEST ORACLE ONLY.

Multivariate form of the reference's 1-D ground truth
(tests/helpers/kalman.py:13-32 in the reference): the initial state is drawn
from N(x0, P0) at t0 and propagated once before the first observation.
Pinned in tests/test_oracle_kalman.py against that 1-D closed form and against
a dense joint-Gaussian evaluation (scipy.stats.multivariate_normal of the
stacked observations) on a short series.
"""

import numpy as np


def kalman_loglik(ys, A, C, Q, R, x0, P0):
    """ys (T, dy); A (dx,dx); C (dy,dx); Q, R covariances; x0 (dx,), P0 (dx,dx)."""
    ys = np.asarray(ys, dtype=np.float64)
    A, C, Q, R = (np.asarray(m, dtype=np.float64) for m in (A, C, Q, R))
    x = np.asarray(x0, dtype=np.float64).copy()
    P = np.asarray(P0, dtype=np.float64).copy()
    ll = 0.0
    dy = C.shape[0]
    for y in ys:
        x = A @ x
        P = A @ P @ A.T + Q
        v = y - C @ x
        S = C @ P @ C.T + R
        Lc = np.linalg.cholesky(S)
        s = np.linalg.solve(Lc, v)
        ll += -0.5 * (dy * np.log(2 * np.pi) + 2 * np.sum(np.log(np.diag(Lc))) + s @ s)
        K = P @ C.T @ np.linalg.inv(S)
        x = x + K @ v
        P = (np.eye(len(x)) - K @ C) @ P
    return float(ll)


def dense_loglik(ys, A, C, Q, R, x0, P0):
    """Same likelihood by building the joint covariance of (y_1..y_T) directly."""
    from scipy.stats import multivariate_normal
    ys = np.asarray(ys, dtype=np.float64)
    T, dy = ys.shape
    dx = A.shape[0]
    means, covs = [], {}
    # x_t = A^t x0 + sum_{s=1..t} A^{t-s} w_s,  w_s ~ N(0,Q); x0 ~ N(x0, P0)
    Apow = [np.eye(dx)]
    for _ in range(T):
        Apow.append(A @ Apow[-1])
    mu = [C @ Apow[t + 1] @ x0 for t in range(T)]
    big = np.zeros((T * dy, T * dy))
    for t in range(T):
        for u in range(T):
            # Cov(x_{t+1}, x_{u+1})
            cov = Apow[t + 1] @ P0 @ Apow[u + 1].T
            for s in range(1, min(t, u) + 2):
                cov = cov + Apow[t + 1 - s] @ Q @ Apow[u + 1 - s].T
            blk = C @ cov @ C.T
            if t == u:
                blk = blk + R
            big[t * dy:(t + 1) * dy, u * dy:(u + 1) * dy] = blk
    return float(multivariate_normal(mean=np.concatenate(mu), cov=big).logpdf(ys.ravel()))
