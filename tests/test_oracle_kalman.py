"""The FP64 Kalman oracle (oracle/kalman.py) against an independent dense
evaluation of the joint Gaussian law of (y_1..y_T) and the one-observation
closed form."""

import numpy as np
from scipy import stats

from oracle import kalman


def random_system(rng, dx, dy):
    A = rng.normal(size=(dx, dx)) * 0.5
    C = rng.normal(size=(dy, dx))
    LQ = np.tril(rng.normal(size=(dx, dx))) * 0.3 + np.eye(dx) * 0.5
    LR = np.tril(rng.normal(size=(dy, dy))) * 0.3 + np.eye(dy) * 0.5
    L0 = np.tril(rng.normal(size=(dx, dx))) * 0.3 + np.eye(dx)
    return A, C, LQ @ LQ.T, LR @ LR.T, rng.normal(size=dx), L0 @ L0.T


def test_kalman_equals_dense_joint_gaussian():
    rng = np.random.default_rng(0)
    for dx, dy in ((1, 1), (2, 2), (3, 2), (2, 1)):
        A, C, Q, R, x0, P0 = random_system(rng, dx, dy)
        ys = rng.normal(size=(6, dy))
        a = kalman.kalman_loglik(ys, A, C, Q, R, x0, P0)
        b = kalman.dense_loglik(ys, A, C, Q, R, x0, P0)
        assert abs(a - b) <= 1e-9 * max(1.0, abs(b)), (dx, dy, a, b)


def test_one_observation_closed_form():
    # x0 ~ N(m, P0) at t0 is propagated once: y1 ~ N(C A m, C (A P0 A' + Q) C' + R)
    A, C, Q, R, m, P0 = np.array([[0.9]]), np.array([[1.0]]), np.array([[0.1]]), np.array([[0.1]]), \
        np.array([0.3]), np.array([[1.0]])
    y = np.array([[0.7]])
    var = (0.9 * 1.0 * 0.9 + 0.1) + 0.1
    exact = stats.norm.logpdf(0.7, loc=0.9 * 0.3, scale=np.sqrt(var))
    assert abs(kalman.kalman_loglik(y, A, C, Q, R, m, P0) - exact) <= 1e-12


def test_kalman_fixture_values():
    # the benchmark's accuracy yardstick equals the oracle's Kalman filter on the fixtures
    import json
    from pathlib import Path
    from oracle import lg as olg
    from tests.helpers import workloads
    w = workloads()
    kal = json.loads((Path(__file__).resolve().parent / "golden" / "kalman_loglik.json").read_text())
    for seed in (1, 2):
        spec = olg.LGSpec(2, 2, separate_init=True)
        A, C, LQ, LR, X0, L0 = (np.asarray(m, np.float64) for m in spec.unpack(w[f"lg2_seed{seed}_theta"]))
        v = kalman.kalman_loglik(w[f"lg2_seed{seed}_ys"], A, C, LQ @ LQ.T, LR @ LR.T, X0, L0 @ L0.T)
        assert v == kal[f"lg2_seed{seed}"]
