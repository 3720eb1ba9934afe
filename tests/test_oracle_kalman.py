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
    for dx, dy in ((1, 1), (2, 2), (3, 2), (2, 1), (6, 5), (8, 8)):
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


def test_oracle_lg_pfilter_of_any_dimension_against_kalman():
    # the oracle's particle filter on LG models of the generic dimensions (the
    # reference defaults, data drawn by the oracle's simulator): the mean of
    # the replicate log-likelihoods lies within Monte Carlo error of the exact
    # Kalman value
    from oracle import filters as fl, lg as olg, threefry as tf
    for dx, dy, seed in ((3, 2, 0), (6, 5, 1)):
        spec = olg.LGSpec(dx, dy)
        th = olg.default_theta(dx, dy)
        theta = np.array([th[n] for n in spec.param_names], np.float32)
        model = olg.LGModel(spec)
        T = 10
        times = np.arange(1, T + 1.0)
        _, Y = fl.simulate(model, theta, fl.make_series(np.zeros((T, dy), np.float32), times, 0.0, nstep=1),
                           tf.split(tf.key(700 + seed), 1)[0])
        ys = np.ascontiguousarray(Y[:, 0, :], dtype=np.float32)
        A, C, LQ, LR, X0, L0 = (np.asarray(m, np.float64) for m in spec.unpack(theta))
        exact = kalman.kalman_loglik(ys, A, C, LQ @ LQ.T, LR @ LR.T, X0, L0 @ L0.T)
        ser = fl.make_series(ys, times, 0.0, nstep=1)
        keys = tf.split(tf.key(710 + seed), 12)
        ll = np.array([fl.pfilter(model, theta, ser, k, 1500)["logLik64"] for k in keys])
        se = ll.std(ddof=1) / np.sqrt(len(ll))
        assert abs(ll.mean() - exact) < 4 * se + 0.05, (dx, dy, ll.mean(), exact, se)
