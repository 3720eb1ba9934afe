"""Systematic resampling, two restatements -- TEST ORACLE ONLY.

``resample_reference`` is the reference's literal algorithm
(pypomp/core/algorithms/helpers.py:14-31): float32 ``unifs = (u + arange(J)) / J``,
float32 ``cumsum(exp(nw))``, ``searchsorted(csum / csum[-1], unifs, side='right')``;
the gather then clamps index J to J-1 (helpers.py:87, JAX's out-of-bounds gather
rule).  Its float32 prefix sum depends on summation order, so no two
implementations agree bit for bit at large J.

``resample_fixed`` is the build's order-independent definition, which the HIP
kernels implement and which this oracle restates:

* m = max(lw); e_i = exp_det(lw_i - m) in float64, where ``exp_det`` uses only
  correctly rounded IEEE multiplies and adds (no fused multiply-add), so every
  conforming implementation gets the same bits;
* E_i = floor(e_i * 2**SHIFT) as uint64, SHIFT = 62 - bitlen(J), so sum(E) < 2**62
  (a prefix fits in 62 bits next to a 2-bit state in one 64-bit word);
* C = inclusive prefix sum of E (integer addition is associative: the scan order
  cannot change it), total = C[J-1];
* unif_j = fl32(fl32(u + j) / J) exactly as the reference;
* T_j = floor(unif_j * total), computed exactly from unif_j's mantissa/exponent;
* ancestor a_j = #{i : C_i <= T_j} (searchsorted right), clamped to J-1.

It agrees with the reference's definition wherever the reference's float32
prefix sum is exact, and chooses the same comb.  Edge cases: NaN log-weights
get zero weight; if no log-weight is finite-or-+inf above -inf (all -inf / NaN),
every particle gets weight 1 (uniform); if m is +inf, the +inf particles share
the mass equally.  Pinned by brute-force exact rational arithmetic in
tests/test_oracle_resample.py.
"""

import math
from fractions import Fraction

import numpy as np

F32 = np.float32

# 1/n! for n = 0..13, correctly rounded to float64 (Horner from the top).
EXP_COEF = [float(Fraction(1, math.factorial(n))) for n in range(14)]
LOG2E = 1.4426950408889634
LN2_HI = 6.93147180369123816490e-01
LN2_LO = 1.90821492927058770002e-10


def shift_for(J):
    return 62 - int(J).bit_length()


def exp_det(x):
    """exp(x) for x <= 0 using only IEEE mul/add/rint/ldexp (float64, vectorised)."""
    x = np.asarray(x, dtype=np.float64)
    k = np.rint(x * LOG2E)
    r = x - k * LN2_HI
    r = r - k * LN2_LO
    p = np.full_like(r, EXP_COEF[13])
    for n in range(12, -1, -1):
        p = p * r + EXP_COEF[n]
    return p, k.astype(np.int64)


def fixed_weights(lw):
    """uint64 fixed-point weights E_i for log-weights lw (float32 array)."""
    lw = np.asarray(lw, dtype=F32)
    J = lw.shape[-1]
    sh = shift_for(J)
    with np.errstate(invalid="ignore"):
        finite_or_inf = lw[~np.isnan(lw)]
        m = finite_or_inf.max() if finite_or_inf.size else -np.inf
    if m == -np.inf:
        return np.ones(J, dtype=np.uint64)
    if m == np.inf:
        return (lw == np.inf).astype(np.uint64)
    with np.errstate(invalid="ignore", over="ignore"):
        d = lw.astype(np.float64) - np.float64(m)
        cutoff = -(sh + 2) * 0.6931471805599453
        live = (d >= cutoff) & ~np.isnan(d)
        dd = np.where(live, d, 0.0)
        p, k = exp_det(dd)
        val = np.ldexp(p, k + sh)
        E = np.where(live, np.floor(val), 0.0)
    return E.astype(np.uint64)


def comb(u, J):
    """Reference comb: unif_j = fl32(fl32(u + j) / J) as float32."""
    j = np.arange(J, dtype=np.int64).astype(F32)
    return ((F32(u) + j).astype(F32) / F32(J)).astype(F32)


def thresholds(unifs, total):
    """T_j = floor(unif_j * total) exactly (Python integers)."""
    out = np.empty(unifs.shape[0], dtype=object)
    total = int(total)
    for idx, f in enumerate(unifs):
        m, e = math.frexp(float(f))  # f = m * 2**e, m in [0.5, 1)
        mant = int(m * (1 << 24))    # exact: f has a 24-bit significand
        sh = e - 24
        prod = mant * total
        out[idx] = prod << sh if sh >= 0 else prod >> (-sh)
    return out


def resample_fixed(lw, u):
    """The build's systematic resampler: ancestor indices (int32), clamped."""
    E = fixed_weights(lw)
    J = E.shape[0]
    C = np.cumsum(E.astype(object))  # exact integers
    total = int(C[-1])
    T = thresholds(comb(u, J), total)
    Cint = np.array([int(c) for c in C], dtype=np.uint64)
    Tint = np.array([int(t) for t in T], dtype=np.uint64)
    a = np.searchsorted(Cint, Tint, side="right")
    return np.minimum(a, J - 1).astype(np.int32)


def resample_fixed_fast(lw, u):
    """Same result as ``resample_fixed`` with uint64 NumPy arithmetic (no overflow:
    sum(E) < 2**62 and unif_j <= 1)."""
    E = fixed_weights(lw)
    J = E.shape[0]
    C = np.cumsum(E, dtype=np.uint64)
    total = int(C[-1])
    unifs = comb(u, J)
    m, e = np.frexp(unifs.astype(np.float64))
    mant = (m * (1 << 24)).astype(np.uint64)
    sh = (e - 24).astype(np.int64)
    # mant * total < 2**24 * 2**62: split total into 32-bit halves
    t_hi = np.uint64(total >> 32)
    t_lo = np.uint64(total & 0xFFFFFFFF)
    p_hi = mant * t_hi            # < 2**55
    p_lo = mant * t_lo            # < 2**56
    # product = X * 2**32 + b with X = p_hi + (p_lo >> 32), b = p_lo & 0xffffffff;
    # unif <= 1 means the right shift s = 24 - e is at least 23.
    X = p_hi + (p_lo >> np.uint64(32))
    b = p_lo & np.uint64(0xFFFFFFFF)
    s = (-sh).astype(np.int64)
    big = s >= 32
    T = np.where(
        big,
        X >> np.where(big, s - 32, 0).astype(np.uint64),
        (X << np.where(big, 0, 32 - s).astype(np.uint64))
        + (b >> np.where(big, 0, s).astype(np.uint64)),
    ).astype(np.uint64)
    a = np.searchsorted(C, T, side="right")
    return np.minimum(a, J - 1).astype(np.int32)


def resample_reference(nw, u):
    """Reference-literal systematic resampling (float32 cumsum); clamped indices."""
    nw = np.asarray(nw, dtype=F32)
    J = nw.shape[-1]
    unifs = comb(u, J)
    csum = np.cumsum(np.exp(nw).astype(F32), dtype=F32)
    a = np.searchsorted((csum / csum[-1]).astype(F32), unifs, side="right")
    return np.minimum(a, J - 1).astype(np.int32)
