"""Deterministic elementary functions restated in NumPy -- TEST ORACLE ONLY.

The engine evaluates every transcendental function on its state / weight /
ancestry path as a fixed sequence of correctly rounded IEEE operations with no
fused multiply-adds (published constructions: Cody & Waite argument reduction,
atanh series for log, Kahan's log1p/expm1 identities, Stirling's series for
log-gamma).  These are the same sequences in float32 / float64 NumPy
arithmetic, so the oracle reproduces the engine's bits.  Pinned against
float64 references (np.log, np.exp, scipy.special.gammaln, ...) for accuracy
in tests/test_oracle_detmath.py; bit-equality with the device is checked by
the GPU parity tests.
"""

import numpy as np

F32 = np.float32
_inf = F32(np.inf)


def _f(a):
    return np.asarray(a, dtype=F32)


def d_log(x):
    x = _f(x).copy()
    out = np.empty_like(x)
    with np.errstate(all="ignore"):
        bad = ~(x > 0)
        inf = x == _inf
        e = np.zeros(x.shape, dtype=np.int64)
        sub = (x < F32(1.17549435e-38)) & ~bad
        x = np.where(sub, (x * F32(8388608.0)).astype(F32), x).astype(F32)
        e = np.where(sub, -23, 0)
        b = x.view(np.uint32)
        e = e + ((b >> np.uint32(23)) & np.uint32(0xFF)).astype(np.int64) - 127
        m = ((b & np.uint32(0x7FFFFF)) | np.uint32(0x3F800000)).view(np.float32)
        big = m > F32(1.41421356)
        m = np.where(big, (m * F32(0.5)).astype(F32), m).astype(F32)
        e = e + big.astype(np.int64)
        f = (m - F32(1.0)).astype(F32)
        s = (f / (F32(2.0) + f).astype(F32)).astype(F32)
        z = (s * s).astype(F32)
        p = np.full(x.shape, F32(0.111111112), dtype=F32)
        for c in (0.142857149, 0.2, 0.333333343, 1.0):
            p = ((p * z).astype(F32) + F32(c)).astype(F32)
        lm = ((F32(2.0) * s).astype(F32) * p).astype(F32)
        fe = e.astype(F32)
        r = ((fe * F32(0.693145751953125)).astype(F32)
             + (lm + (fe * F32(1.42860677e-06)).astype(F32)).astype(F32)).astype(F32)
        out = np.where(bad, np.where(x == 0, -_inf, F32(np.nan)), np.where(inf, _inf, r)).astype(F32)
    return out


def d_log1p(x):
    x = _f(x)
    with np.errstate(all="ignore"):
        u = (F32(1.0) + x).astype(F32)
        d = (u - F32(1.0)).astype(F32)
        r = (d_log(u) * (x / d).astype(F32)).astype(F32)
        r = np.where(u == F32(1.0), x, r)
        r = np.where(x == 0, x, r)
    return r.astype(F32)


def _pow2(k):
    return ((np.asarray(k, np.int64) + 127).astype(np.uint32) << np.uint32(23)).view(np.float32)


def d_exp(x):
    x = _f(x)
    with np.errstate(all="ignore"):
        k = np.rint((x * F32(1.44269502)).astype(F32)).astype(F32)
        r = ((x - (k * F32(0.693145751953125)).astype(F32)).astype(F32)
             - (k * F32(1.42860677e-06)).astype(F32)).astype(F32)
        p = np.full(x.shape, F32(1.98412701e-04), dtype=F32)
        for c in (1.38888892e-03, 8.33333377e-03, 4.16666679e-02, 0.166666672, 0.5, 1.0, 1.0):
            p = ((p * r).astype(F32) + F32(c)).astype(F32)
        ki = np.clip(k, -300, 300).astype(np.int64)
        lo = ki < -126
        hi = ki > 127
        mid = (p * _pow2(np.clip(ki, -126, 127))).astype(F32)
        low = ((p * _pow2(np.clip(ki + 64, -126, 127))).astype(F32) * F32(5.42101086e-20)).astype(F32)
        high = ((p * _pow2(np.clip(ki - 64, -126, 127))).astype(F32) * F32(1.84467441e+19)).astype(F32)
        r = np.where(lo, low, np.where(hi, high, mid))
        r = np.where(x < F32(-87.3365479), F32(0.0), r)
        r = np.where(x > F32(88.7228394), _inf, r)
        r = np.where(x != x, x, r)
    return r.astype(F32)


def d_expm1(x):
    x = _f(x)
    with np.errstate(all="ignore"):
        u = d_exp(x)
        um1 = (u - F32(1.0)).astype(F32)
        r = (um1 * (x / d_log(u)).astype(F32)).astype(F32)
        r = np.where(u == _inf, u, r)
        r = np.where(um1 == F32(-1.0), F32(-1.0), r)
        r = np.where(u == F32(1.0), x, r)
        r = np.where(x == 0, x, r)
    return r.astype(F32)


def d_tanh(x):
    x = _f(x)
    ax = np.where(x < 0, -x, x).astype(F32)
    t = d_expm1((F32(-2.0) * ax).astype(F32))
    with np.errstate(all="ignore"):
        r = (-t / (t + F32(2.0)).astype(F32)).astype(F32)
    return np.where(x < 0, -r, r).astype(F32)


def d_cos(x):
    x = _f(x)
    k = np.rint((x * F32(0.636619747)).astype(F32)).astype(F32)
    r = (x - (k * F32(1.5703125)).astype(F32)).astype(F32)
    r = (r - (k * F32(4.83751297e-04)).astype(F32)).astype(F32)
    r = (r - (k * F32(7.54978942e-08)).astype(F32)).astype(F32)
    z = (r * r).astype(F32)
    sp = np.full(x.shape, F32(-2.50521079e-08), dtype=F32)
    for c in (2.75573188e-06, -1.98412701e-04, 8.33333377e-03, -0.166666672):
        sp = ((sp * z).astype(F32) + F32(c)).astype(F32)
    sn = (r + ((r * z).astype(F32) * sp).astype(F32)).astype(F32)
    cp = np.full(x.shape, F32(2.08767559e-09), dtype=F32)
    for c in (-2.75573188e-07, 2.48015876e-05, -1.38888892e-03, 4.16666679e-02, -0.5):
        cp = ((cp * z).astype(F32) + F32(c)).astype(F32)
    cs = (F32(1.0) + (z * cp).astype(F32)).astype(F32)
    q = k.astype(np.int64) & 3
    return np.select([q == 0, q == 1, q == 2], [cs, -sn, -cs], sn).astype(F32)


def d_pow(b, e):
    return d_exp((_f(e) * d_log(b)).astype(F32))


def dd_log(x):
    x = np.asarray(x, dtype=np.float64).copy()
    with np.errstate(all="ignore"):
        bad = ~(x > 0)
        inf = x == np.inf
        sub = (x < 2.2250738585072014e-308) & ~bad
        x = np.where(sub, x * 4503599627370496.0, x)
        e = np.where(sub, -52, 0).astype(np.int64)
        b = x.view(np.uint64)
        e = e + ((b >> np.uint64(52)) & np.uint64(0x7FF)).astype(np.int64) - 1023
        m = ((b & np.uint64(0x000FFFFFFFFFFFFF)) | np.uint64(0x3FF0000000000000)).view(np.float64)
        big = m > 1.4142135623730951
        m = np.where(big, m * 0.5, m)
        e = e + big.astype(np.int64)
        f = m - 1.0
        s = f / (2.0 + f)
        z = s * s
        p = np.full(x.shape, 1.0 / 23.0)
        for den in (21.0, 19.0, 17.0, 15.0, 13.0, 11.0, 9.0, 7.0, 5.0, 3.0):
            p = p * z + 1.0 / den
        p = p * z + 1.0
        lm = (2.0 * s) * p
        de = e.astype(np.float64)
        r = de * 6.93147180369123816490e-01 + (lm + de * 1.90821492927058770002e-10)
        r = np.where(bad, np.where(x == 0, -np.inf, np.nan), np.where(inf, np.inf, r))
    return r


def dd_exp(x):
    """float64 exp: Cody-Waite reduction by a split ln 2, Taylor polynomial to
    r^13/13!, scaling by 2^n (two factors below 2^-1000)."""
    x = np.asarray(x, dtype=np.float64)
    with np.errstate(all="ignore"):
        n = np.rint(x * 1.4426950408889634)
        r = (x - n * 6.93147180369123816490e-01) - n * 1.90821492927058770002e-10
        p = np.full(x.shape, 1.0 / 6227020800.0)
        for den in (479001600.0, 39916800.0, 3628800.0, 362880.0, 40320.0, 5040.0, 720.0, 120.0, 24.0, 6.0):
            p = p * r + 1.0 / den
        p = p * r + 0.5
        p = p * r + 1.0
        p = p * r + 1.0
        k = np.nan_to_num(n, nan=0.0, posinf=0.0, neginf=0.0).astype(np.int64)
        low = k < -1000
        out = np.where(low, np.ldexp(np.ldexp(p, np.where(low, k + 600, 0)), np.where(low, -600, 0)),
                       np.ldexp(p, np.where(low, 0, k)))
        out = np.where(x > 709.782712893384, np.inf, out)
        out = np.where(x < -745.1332191019412, 0.0, out)
        out = np.where(x != x, x, out)
    return out


def dd_lgamma(x):
    """log Gamma for x > 0 (scalar loop per element, as the device)."""
    x = np.atleast_1d(np.asarray(x, dtype=np.float64))
    out = np.empty_like(x)
    for i, xi in enumerate(x.ravel()):
        if not (xi > 0):
            out.ravel()[i] = np.nan
            continue
        shift, prod = 0.0, 1.0
        while xi < 12.0:
            prod = prod * xi
            xi = xi + 1.0
            if prod > 1e280:
                shift = shift + float(dd_log(prod))
                prod = 1.0
        shift = shift + float(dd_log(prod))
        ix = 1.0 / xi
        ix2 = ix * ix
        s = 1.0 / 1188.0
        s = s * ix2 - 1.0 / 1680.0
        s = s * ix2 + 1.0 / 1260.0
        s = s * ix2 - 1.0 / 360.0
        s = s * ix2 + 1.0 / 12.0
        out.ravel()[i] = ((xi - 0.5) * float(dd_log(xi)) - xi + 0.91893853320467274178) + s * ix - shift
    return out
