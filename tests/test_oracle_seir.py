"""The SEIR oracle (oracle/seir.py): the binomial sampler against scipy's
binomial quantile and its exact edge cases, the normal quantile, the NegBin
measurement density and the counting invariants of the Euler-binomial step."""

import numpy as np
from scipy import special, stats

from oracle import seir, threefry as tf

F32 = np.float32


def test_ndtri_matches_scipy():
    rng = np.random.default_rng(0)
    u = np.concatenate([rng.uniform(1e-6, 1 - 1e-6, 50000), [0.5, 0.025, 0.975, 1e-7]]).astype(F32)
    got = seir.ndtri_f32(u).astype(np.float64)
    ref = special.ndtri(u.astype(np.float64))
    assert np.max(np.abs(got - ref) / np.maximum(1.0, np.abs(ref))) <= 2e-6
    assert seir.ndtri_f32(F32(0.5)) == 0


def test_binominv_matches_scipy_quantile():
    # the reference accepts a mismatch rate below 0.15 % against the exact
    # quantile (tests/test_random/test_sampler_accuracy.py:43-74)
    rng = np.random.default_rng(1)
    m = 40000
    n = np.floor(10 ** rng.uniform(0, 6, m)).astype(F32)
    p = rng.uniform(0, 1, m).astype(F32)
    u = rng.uniform(0, 1, m).astype(F32)
    got = seir.binominv_f32(u, n, p)
    ref = stats.binom.ppf(u.astype(np.float64), n.astype(np.float64), p.astype(np.float64))
    mism = np.mean(got != ref)
    assert mism < 0.0015, mism
    # where it differs it is off by one at most
    assert np.max(np.abs(got - ref)) <= 1


def test_binominv_edge_cases():
    u = np.array([0.3, 0.3, 0.3, 0.0, 1.0, 0.3, 0.3, 0.3], F32)
    n = np.array([0, 10, 10, 10, 10, -1, 10, 10], F32)
    p = np.array([0.5, 0.0, 1.0, 0.4, 0.4, 0.5, 1.5, -0.1], F32)
    r = seir.binominv_f32(u, n, p)
    assert r[0] == 0 and r[1] == 0 and r[2] == 10 and r[3] == 0 and r[4] == 10
    assert np.isnan(r[5]) and np.isnan(r[6]) and np.isnan(r[7])
    # small n*p uses the exact bottom-up CDF: equals the exact quantile
    uu = np.linspace(0.01, 0.99, 99).astype(F32)
    got = seir.binominv_f32(uu, F32(7), F32(0.2))
    assert np.array_equal(got, stats.binom.ppf(uu.astype(np.float64), 7, np.float64(F32(0.2))))


def test_negbin_dmeas_matches_scipy():
    model = seir.SEIRModel()
    th = seir.default_theta()
    J = 64
    rng = np.random.default_rng(2)
    X = np.zeros((J, 5), F32)
    X[:, 4] = rng.integers(1, 5000, J).astype(F32)
    for y in (0.0, 3.0, 250.0, 1800.0):
        got = model.dmeas(np.array([y], F32), X, th, 1.0).astype(np.float64)
        mu = X[:, 4].astype(np.float64) * np.float64(th[4])
        k = np.float64(th[5])
        ref = stats.nbinom.logpmf(y, k, k / (k + mu))
        assert np.max(np.abs(got - ref) / np.maximum(1, np.abs(ref))) <= 2e-6, y


def test_euler_binomial_step_conserves_and_accumulates():
    model = seir.SEIRModel()
    th = seir.default_theta()
    keys = tf.split(tf.key(3), 512)
    X = model.rinit(th, keys)
    N = float(th[6])
    assert np.all(X[:, :4].sum(1, dtype=np.float64) == N) and np.all(X[:, 4] == 0)
    H = np.zeros(512)
    t = 0.0
    for s in range(7):
        ks = tf.split_many(keys, 2)
        keys = ks[:, 0]
        Xn = model.rproc_step(X, th, ks[:, 1], t, 1.0 / 7)
        assert np.all(Xn >= 0)
        assert np.all(Xn[:, :4].sum(1, dtype=np.float64) == N)
        dEI = (Xn[:, 2] - X[:, 2]) + (Xn[:, 3] - X[:, 3])  # I + R gain = E->I transitions
        H += dEI
        assert np.all(Xn[:, 1] + Xn[:, 2] + Xn[:, 3] - (X[:, 1] + X[:, 2] + X[:, 3]) == X[:, 0] - Xn[:, 0])
        X = Xn
        t += 1.0 / 7
    assert np.array_equal(X[:, 4], H.astype(F32))


def test_first_substep_transition_means():
    # each transition count is Binomial(n, 1 - exp(-rate dt)); over 4096
    # particles the sample mean is within 5 standard errors of n p
    model = seir.SEIRModel()
    th = seir.default_theta().astype(np.float64)
    beta, amp, sig, gam, rho, k, N, S0f, E0, I0 = th
    keys = tf.split(tf.key(4), 4096)
    X = model.rinit(seir.default_theta(), keys)
    dt = 1.0 / 7
    Xn = model.rproc_step(X, seir.default_theta(), tf.split(tf.key(5), 4096), 0.0, dt)
    S0 = float(X[0, 0])
    dSE = X[:, 0] - Xn[:, 0]
    dIR = Xn[:, 3] - X[:, 3]
    dEI = Xn[:, 4]
    for d, n, rate in ((dSE, S0, beta * (1 + amp) * I0 / N), (dEI, E0, sig), (dIR, I0, gam)):
        p = 1 - np.exp(-rate * dt)
        se = np.sqrt(n * p * (1 - p) / d.size)
        assert abs(d.astype(np.float64).mean() - n * p) < 5 * se, (n, p, d.mean())


def test_negbin_inverse_summation_matches_scipy_ppf():
    # the device's NegBin sampler (upward CDF summation with rescaling) against
    # scipy's nbinom.ppf, including p^k below the float64 range (k = 200, mu = 1e4)
    rng = np.random.default_rng(12)
    for k, mu in ((10.0, 0.3), (10.0, 37.5), (10.0, 4200.0), (0.7, 15.0), (200.0, 1e4), (1e3, 2.0), (10.0, 1e-10)):
        u = np.concatenate([rng.uniform(0, 1, 3000).astype(np.float32).astype(np.float64), [1 - 2.0 ** -24, 2.0 ** -24]])
        got = seir.negbin_inverse_sum(u, k, mu)
        ref = stats.nbinom.ppf(u, k, k / (k + mu))
        assert np.array_equal(got, ref), (k, mu, np.flatnonzero(got != ref)[:5])
    assert seir.negbin_inverse_sum(np.array([0.0]), 10.0, 5.0)[0] == 0.0
