"""Accuracy of the deterministic elementary functions (oracle/detmath.py) against
float64 library references.  These functions define the engine's bits; here
they are pinned to the mathematics: a wrong constant, a dropped series term or
a mis-rounded reduction shows up as an error far above the stated bound."""

import math

import numpy as np
from scipy import special

from oracle import detmath as dm

F32 = np.float32


def ulp_err(got, exact):
    got = np.asarray(got, dtype=np.float64)
    exact = np.asarray(exact, dtype=np.float64)
    spacing = np.spacing(np.abs(exact).astype(F32)).astype(np.float64)
    return np.max(np.abs(got - exact) / spacing)


def test_log_accuracy_and_special_values():
    rng = np.random.default_rng(1)
    x = np.concatenate([np.exp(rng.uniform(-80, 80, 20000)), rng.uniform(0.5, 2, 20000),
                        [1.0, 2.0, 0.5, np.finfo(F32).tiny, np.finfo(F32).max]]).astype(F32)
    assert ulp_err(dm.d_log(x), np.log(x.astype(np.float64))) <= 4.0
    assert dm.d_log(F32(1)) == 0
    with np.errstate(all="ignore"):
        assert np.isneginf(dm.d_log(F32(0)))
        assert np.isnan(dm.d_log(F32(-1)))
        assert np.isposinf(dm.d_log(F32(np.inf)))


def test_log1p_expm1_accuracy():
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.uniform(-0.9, 3, 20000), 10.0 ** rng.uniform(-12, -1, 5000)]).astype(F32)
    assert ulp_err(dm.d_log1p(x), np.log1p(x.astype(np.float64))) <= 4.0
    y = np.concatenate([rng.uniform(-20, 20, 20000), 10.0 ** rng.uniform(-12, -1, 5000)]).astype(F32)
    assert ulp_err(dm.d_expm1(y), np.expm1(y.astype(np.float64))) <= 4.0


def test_exp_accuracy_and_range():
    rng = np.random.default_rng(3)
    x = rng.uniform(-87, 88, 40000).astype(F32)
    assert ulp_err(dm.d_exp(x), np.exp(x.astype(np.float64))) <= 2.0
    assert dm.d_exp(F32(0)) == 1
    assert dm.d_exp(F32(-200)) == 0           # flushes below the normal range
    assert np.isposinf(dm.d_exp(F32(100)))


def test_tanh_cos_pow():
    rng = np.random.default_rng(4)
    x = rng.uniform(-10, 10, 20000).astype(F32)
    t = dm.d_tanh(x).astype(np.float64)
    assert np.max(np.abs(t - np.tanh(x.astype(np.float64)))) <= 4 * np.finfo(F32).eps
    c = rng.uniform(-60, 60, 20000).astype(F32)
    assert np.max(np.abs(dm.d_cos(c).astype(np.float64) - np.cos(c.astype(np.float64)))) <= 4 * np.finfo(F32).eps
    b = rng.uniform(0.5, 1.0, 2000).astype(F32)
    e = rng.uniform(0, 60, 2000).astype(F32)
    exact = np.power(b.astype(np.float64), e.astype(np.float64))
    assert np.max(np.abs(dm.d_pow(b, e) / exact - 1)) <= 64 * np.finfo(F32).eps


def test_double_log_and_lgamma():
    rng = np.random.default_rng(5)
    x = np.exp(rng.uniform(-50, 50, 20000))
    assert np.max(np.abs(dm.dd_log(x) - np.log(x))) <= 4e-16 * np.maximum(1, np.abs(np.log(x))).max()
    z = np.concatenate([rng.uniform(0.05, 30, 20000), np.arange(1, 200, dtype=np.float64), 10 ** rng.uniform(0, 6, 2000)])
    got = dm.dd_lgamma(z)
    ref = special.gammaln(z)
    assert np.max(np.abs(got - ref) / np.maximum(1.0, np.abs(ref))) <= 1e-13
    # factorials are exact
    for n in range(1, 25):
        assert abs(float(dm.dd_lgamma(np.array([n + 1.0]))[0]) - math.lgamma(n + 1)) <= 1e-13 * max(1, math.lgamma(n + 1))


def test_double_exp():
    rng = np.random.default_rng(6)
    x = np.concatenate([rng.uniform(-745, 709.7, 40000), rng.uniform(-1, 1, 10000), np.arange(-50, 51.0)])
    got = dm.dd_exp(x)
    ref = np.exp(x)
    normal = ref > 2.2250738585072014e-308
    assert np.max(np.abs(got[normal] / ref[normal] - 1)) <= 4 * np.finfo(np.float64).eps
    assert np.max(np.abs(got[~normal] - ref[~normal])) <= 4 * 5e-324
    assert dm.dd_exp(np.array([0.0]))[0] == 1.0
    assert abs(dm.dd_exp(np.array([1.0]))[0] - math.e) <= 2 * np.spacing(math.e)
    edge = dm.dd_exp(np.array([800.0, -800.0, np.inf, -np.inf, np.nan]))
    assert edge[0] == np.inf and edge[1] == 0.0 and edge[2] == np.inf and edge[3] == 0.0 and np.isnan(edge[4])
