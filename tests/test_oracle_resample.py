"""Pins the oracle's systematic resamplers (oracle/resample.py) with exact
rational arithmetic, the exp definition, and systematic-resampling invariants."""

import math
from fractions import Fraction

import numpy as np
import pytest

from oracle import resample as rs

F32 = np.float32


def brute_ancestors(E, u):
    """a_j = #{i : C_i / total <= unif_j} in exact rationals, clamped to J-1."""
    J = len(E)
    C = np.cumsum([int(e) for e in E], dtype=object)
    total = int(C[-1])
    unifs = rs.comb(u, J)
    out = []
    for f in unifs:
        fr = Fraction(float(f))
        a = sum(1 for c in C if Fraction(int(c), total) <= fr)
        out.append(min(a, J - 1))
    return np.array(out)


@pytest.mark.parametrize("seed", range(6))
def test_fixed_resampler_matches_exact_rationals(seed):
    rng = np.random.default_rng(seed)
    J = int(rng.integers(1, 60))
    lw = (rng.normal(size=J) * rng.choice([0.1, 3.0, 30.0])).astype(F32)
    if J > 3:
        lw[rng.integers(0, J)] = -np.inf
    u = F32(rng.random() * (1 - 2 ** -23))
    E = rs.fixed_weights(lw)
    a = rs.resample_fixed(lw, u)
    assert np.array_equal(a, brute_ancestors(E, u))
    assert np.array_equal(a, rs.resample_fixed_fast(lw, u))


def test_exp_det_accuracy():
    x = np.concatenate([np.linspace(-40, 0, 100001), -np.logspace(-12, 1, 1000)])
    p, k = rs.exp_det(x)
    got = np.ldexp(p, k)
    ref = np.exp(x)
    assert np.max(np.abs(got - ref) / ref) < 4e-16
    assert rs.exp_det(np.array([0.0]))[0][0] == 1.0  # the max particle weighs exactly 2^sh


def test_fixed_weights_against_float_exp():
    rng = np.random.default_rng(1)
    lw = rng.normal(size=5000).astype(F32) * F32(5)
    E = rs.fixed_weights(lw).astype(np.float64)
    sh = rs.shift_for(lw.size)
    d = lw.astype(np.float64) - float(lw.max())
    approx = np.floor(np.exp(d) * 2.0 ** sh)
    assert np.max(np.abs(E - approx)) <= 2.0
    assert E.sum() < 2.0 ** 62


def test_uniform_and_collapsed_weights_give_identity():
    J = 1000
    for lw in (np.zeros(J, F32), np.full(J, -np.inf, F32)):
        a = rs.resample_fixed(lw, F32(0.37))
        assert np.array_equal(a, np.arange(J))


def test_nan_weights_get_no_offspring():
    lw = np.array([0.0, np.nan, 0.0, np.nan], dtype=F32)
    a = rs.resample_fixed(lw, F32(0.5))
    assert set(a.tolist()) <= {0, 2}


def test_systematic_offspring_counts():
    rng = np.random.default_rng(7)
    J = 4096
    lw = (rng.normal(size=J) * 2).astype(F32)
    a = rs.resample_fixed_fast(lw, F32(0.123))
    assert np.all(np.diff(a) >= 0) and a.min() >= 0 and a.max() < J
    w = np.exp(lw.astype(np.float64) - lw.max())
    w /= w.sum()
    n = np.bincount(a, minlength=J)
    assert np.all(np.abs(n - J * w) < 1.0 + 1e-6)


def test_agrees_with_reference_literal_on_generic_weights():
    rng = np.random.default_rng(3)
    agree = tot = 0
    for _ in range(50):
        J = 200
        lw = rng.normal(size=J).astype(F32)
        u = F32(rng.random())
        nw = (lw - F32(np.log(np.sum(np.exp(lw.astype(np.float64)))))).astype(F32)
        a1 = rs.resample_fixed_fast(lw, u)
        a2 = rs.resample_reference(nw, u)
        agree += int(np.sum(a1 == a2))
        tot += J
    assert agree / tot > 0.99


def test_comb_rounding_to_one_is_clamped():
    # (u + J - 1) / J rounds to 1.0 in float32 for u close to 1: index J -> J-1
    J = 3
    u = np.nextafter(F32(1), F32(0))
    assert rs.comb(u, J)[-1] <= 1.0
    lw = np.array([0.0, 0.0, -np.inf], dtype=F32)
    a = rs.resample_fixed(lw, u)
    assert a.max() <= J - 1
