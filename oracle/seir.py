"""SEIR model and the binomial sampler restated in NumPy float32 -- TEST ORACLE ONLY.

The model is defined here (the reference has no SEIR); it follows the
reference's SIR model pattern (pypomp/models/sir.py:171-249) and draws every
binomial with the reference's sampler, ``fast_binomial`` / ``binominv``
(pypomp/random/binom.py:130-522: Giles & Beentjes (2024) order-2 normal
asymptotic inverse with a <= 5-term bottom-up exact stage).  The normal
quantile inside it is sqrt(2) * erfinv(2u - 1) with Giles' single-precision
erfinv evaluated at w = -log(4u(1-u)) (the reference calls
jax.scipy.stats.norm.ppf there).

Pins: ``binominv_f32`` against scipy.stats.binom.ppf (mismatch rate bound as
the reference's tests/test_random/test_sampler_accuracy.py:43-74) and exact
edge cases (tests/test_oracle_seir.py); ``ndtri_f32`` against
scipy.special.ndtri.  The SEIR step itself (an integer Markov chain defined
here) is pinned by conservation (S+E+I+R = N), the accumulator reset, and the
binomial moments; its GPU parity is bit-exact on counts for a fixed key.

State words: S, E, I, R, H (H = E->I transitions since the last observation).
theta = (beta, amp, sigma, gamma, rho, k, N, S0f, E0, I0), natural scale.
"""

import numpy as np
from scipy import stats

from . import filters as fl
from . import threefry as tf
from .detmath import d_cos, d_exp, d_expm1, d_log, d_log1p, dd_lgamma, dd_log

F32 = np.float32
_C = tf._ERFINV_CENTRAL
_T = tf._ERFINV_TAIL

PARAM_NAMES = ["beta", "amp", "sigma", "gamma", "rho", "k", "N", "S0f", "E0", "I0"]
DEFAULTS = {"beta": 1.5, "amp": 0.2, "sigma": 1.0, "gamma": 0.7, "rho": 0.4, "k": 10.0,
            "N": 1e6, "S0f": 0.95, "E0": 500.0, "I0": 500.0}
_LOG = [0, 2, 3, 5, 6, 8, 9]
_LOGIT = [1, 4, 7]


def ndtri_f32(u):
    u = np.asarray(u, dtype=F32)
    with np.errstate(divide="ignore", invalid="ignore"):
        w = (-d_log((F32(4.0) * u * (F32(1.0) - u)).astype(F32))).astype(F32)
        x = (F32(2.0) * u - F32(1.0)).astype(F32)
        lt = w < F32(5.0)
        wc = np.where(lt, w - F32(2.5), np.sqrt(w) - F32(3.0)).astype(F32)
        p = np.where(lt, _C[0], _T[0]).astype(F32)
        for i in range(1, 9):
            p = (np.where(lt, _C[i], _T[i]).astype(F32) + p * wc).astype(F32)
    return (F32(1.41421354) * (p * x).astype(F32)).astype(F32)


def binominv_f32(u, n, p):
    """binominv (binom.py:236-393) for float32 arrays u, n, p; order 2, exact_max 5."""
    u = np.asarray(u, dtype=F32)
    n = np.asarray(n, dtype=F32)
    p = np.asarray(p, dtype=F32)
    u, n, p = np.broadcast_arrays(u, n, p)
    tiny, eps = np.finfo(F32).tiny, np.finfo(F32).eps
    with np.errstate(all="ignore"):
        invalid = (n < 0) | (p < 0) | (p > 1) | (u < 0) | (u > 1)
        flip = p > F32(0.5)
        ps = np.where(flip, F32(1) - p, p).astype(F32)
        uf = np.where(flip, F32(1) - u, u).astype(F32)
        ns = np.where(n < 0, F32(1), n).astype(F32)
        ps = np.clip(ps, F32(0), F32(1)).astype(F32)
        us = np.clip(uf, eps, F32(1) - eps).astype(F32)
        q = (F32(1) - ps).astype(F32)
        w = ndtri_f32(us)
        w2 = (w * w).astype(F32)
        np_ = (ns * ps).astype(F32)
        npq = (np_ * q).astype(F32)
        sq = np.sqrt(np.maximum(npq, tiny)).astype(F32)
        pq = (ps * q).astype(F32)
        q0 = (np_ + sq * w + (F32(2) + F32(2) * ps + (q - ps) * w2) / F32(6)).astype(F32)
        w3 = (w2 * w).astype(F32)
        q1 = (q0 + ((F32(-2) + F32(14) * pq) * w + (F32(-1) - F32(2) * pq) * w3)
              / np.maximum(F32(72) * sq, tiny)).astype(F32)
        w4 = (w3 * w).astype(F32)
        q2 = (q1 + ((ps - q) * (F32(2) + pq) * (F32(16) - F32(7) * w2 - F32(3) * w4))
              / np.maximum(F32(1620) * (sq * sq), tiny)).astype(F32)
        k = np.minimum(np.maximum(np.floor(q2), F32(0)), ns).astype(F32)
        k = np.minimum(k, (np_ + F32(6) * sq + F32(5)).astype(F32)).astype(F32)
        # bottom-up exact stage (binom.py:396-451)
        ue = np.clip(uf, F32(0), F32(1)).astype(F32)
        qc = np.clip(F32(1) - ps, tiny, F32(1)).astype(F32)
        ratio = (ps / qc).astype(F32)
        logq = d_log1p(-ps)
        pmf = np.where(ns == 0, F32(1), d_exp((ns * logq).astype(F32))).astype(F32)
        cdf = pmf.copy()
        found = cdf >= ue
        res = np.where(found, F32(0), k).astype(F32)
        for i in range(1, 5):
            num = np.maximum(ns - F32(i - 1), F32(0)).astype(F32)
            pmf = ((pmf * (num / F32(i)).astype(F32)).astype(F32) * ratio).astype(F32)
            cdf = (cdf + pmf).astype(F32)
            stalled = (cdf > F32(1) - eps * F32(10)) & (pmf < eps)
            fnow = (cdf >= ue) | stalled
            res = np.where(~found & fnow, F32(i), res).astype(F32)
            found = found | fnow
        kb = np.minimum(np.maximum(res, F32(0)), ns).astype(F32)
        use_bu = (k < F32(10)) | (np_ < F32(4))
        k = np.where(use_bu, kb, k).astype(F32)
        r = np.where(flip, ns - k, k).astype(F32)
        r = np.where(n == 0, F32(0), r)
        r = np.where(u == 0, F32(0), r)
        r = np.where(u == 1, ns, r)
        r = np.where(p == 0, F32(0), r)
        r = np.where(p == 1, ns, r)
        r = np.clip(r, F32(0), ns)
        r = np.where(invalid, F32(np.nan), r)
    return r.astype(F32)


def fast_binomial(keys, n, p):
    """fast_binomial with one scalar uniform per key (binom.py:211-233)."""
    u = tf.uniform_from_bits(tf.random_bits_many(keys, 1)[:, 0])
    u = np.clip(u, np.finfo(F32).tiny, np.nextafter(F32(1), F32(0))).astype(F32)
    r = binominv_f32(u, np.asarray(n, dtype=F32), p)
    return np.where(np.isnan(r), F32(-1), r).astype(F32)


class SEIRModel:
    S = 5
    accumvars = (4,)
    param_names = PARAM_NAMES

    def from_est(self, th):
        th = np.array(th, dtype=F32, copy=True)
        th[..., _LOG] = d_exp(th[..., _LOG])
        th[..., _LOGIT] = F32(1) / (F32(1) + d_exp(-th[..., _LOGIT]))
        return th.astype(F32)

    def to_est(self, th):
        th = np.array(th, dtype=F32, copy=True)
        th[..., _LOG] = np.log(th[..., _LOG])
        pp = th[..., _LOGIT]
        th[..., _LOGIT] = np.log(pp / (F32(1) - pp))
        return th.astype(F32)

    @staticmethod
    def _cols(theta, J):
        th = np.asarray(theta, dtype=F32)
        if th.ndim == 1:
            th = np.broadcast_to(th, (J, th.shape[0]))
        return [th[:, i] for i in range(th.shape[1])]

    def rinit(self, theta, keys, t0=0.0):
        J = keys.shape[0]
        beta, amp, sig, gam, rho, k, N, S0f, E0, I0 = self._cols(theta, J)
        S = np.rint((S0f * N).astype(F32)).astype(F32)
        E = np.rint(E0).astype(F32)
        I = np.rint(I0).astype(F32)
        R = (((N - S).astype(F32) - E).astype(F32) - I).astype(F32)
        return np.stack([S, E, I, R, np.zeros(J, F32)], axis=1).astype(F32)

    def rproc_step(self, X, theta, keys, t, dt):
        J = X.shape[0]
        beta, amp, sig, gam, rho, k, N, S0f, E0, I0 = self._cols(theta, J)
        S, E, I, R, H = (X[:, i] for i in range(5))
        invN = (F32(1) / N).astype(F32)
        bt = (beta * (F32(1) + (amp * d_cos((F32(6.28318531) * F32(t) / F32(52)).astype(F32))).astype(F32))).astype(F32)
        foi = (bt * I * invN).astype(F32)
        dt = F32(dt)
        pSE = (-d_expm1(-(foi * dt).astype(F32))).astype(F32)
        pEI = (-d_expm1(-(sig * dt).astype(F32))).astype(F32)
        pIR = (-d_expm1(-(gam * dt).astype(F32))).astype(F32)
        k3 = tf.split_many(keys, 3)
        dSE = fast_binomial(k3[:, 0], S, pSE)
        dEI = fast_binomial(k3[:, 1], E, pEI)
        dIR = fast_binomial(k3[:, 2], I, pIR)
        S = S - dSE
        E = E + dSE - dEI
        I = I + dEI - dIR
        R = R + dIR
        H = H + dEI
        return np.stack([S, E, I, R, H], axis=1).astype(F32)

    def dmeas(self, y, X, theta, t):
        J = X.shape[0]
        beta, amp, sig, gam, rho, k, N, S0f, E0, I0 = self._cols(theta, J)
        mu = np.maximum((X[:, 4] * rho).astype(F32), F32(1e-10)).astype(F32).astype(np.float64)
        yy = float(np.rint(np.asarray(y, dtype=F32)[0]))
        k64 = k.astype(np.float64)
        ku, kinv = np.unique(k64, return_inverse=True)
        lgk = dd_lgamma(ku)[kinv]
        lgyk = dd_lgamma(ku + yy)[kinv]
        lgy1 = dd_lgamma(np.array([yy + 1.0]))[0]
        v = (((lgyk - lgk) - lgy1) + k64 * dd_log(k64 / (k64 + mu))) + yy * dd_log(mu / (k64 + mu))
        return v.astype(F32)

    def rmeas(self, X, theta, keys, t):
        """NegBin(mean rho H, size k) by inversion of one uniform u: the smallest
        y >= 0 with F(y) >= u (scipy's ppf), and 0 at u = 0."""
        J = X.shape[0]
        beta, amp, sig, gam, rho, k, N, S0f, E0, I0 = self._cols(theta, J)
        mu = np.maximum((X[:, 4] * rho).astype(np.float64), 1e-10)
        u = tf.uniform_from_bits(tf.random_bits_many(keys, 1)[:, 0]).astype(np.float64)
        kk = k.astype(np.float64)
        y = np.where(u > 0, stats.nbinom.ppf(u, kk, kk / (kk + mu)), 0.0)
        return y.astype(F32)[:, None]


def negbin_inverse_sum(u, k, mu):
    """Smallest y >= 0 with F(y) >= u for NegBin(size k, mean mu) by upward
    summation in float64 (vectorised over draws): pmf(0) = p^k, pmf(y+1) =
    pmf(y) (y + k)/(y + 1) q with p = k/(k+mu), q = mu/(k+mu); F and pmf are
    carried in units of pmf(0) 2^(600 r) (r rescalings), so p^k may underflow.
    The device sampler uses this construction; it agrees with scipy's
    ``nbinom.ppf`` except where u lies within rounding of a CDF step."""
    from .detmath import dd_exp, dd_log
    u, k, mu = (np.asarray(a, np.float64) for a in np.broadcast_arrays(u, k, mu))
    p, q = k / (k + mu), mu / (k + mu)
    with np.errstate(divide="ignore"):
        lt = dd_log(u) - k * dd_log(p)
    term = np.ones_like(u)
    cdf = np.ones_like(u)
    y = np.zeros_like(u)
    big, lbig = 2.0 ** 600, 600 * 0.6931471805599453
    target = np.where(lt > 700.0, np.inf, dd_exp(np.minimum(lt, 700.0)))
    active = (u > 0) & (cdf < target)
    while np.any(active):
        ratio = (y + k) / (y + 1.0)
        term = np.where(active, (term * ratio) * q, term)
        cdf = np.where(active, cdf + term, cdf)
        y = np.where(active, y + 1.0, y)
        resc = active & (lt > 700.0) & (cdf > big)
        term = np.where(resc, term * 2.0 ** -600, term)
        cdf = np.where(resc, cdf * 2.0 ** -600, cdf)
        lt = np.where(resc, lt - lbig, lt)
        target = np.where(resc & (lt <= 700.0), dd_exp(np.minimum(lt, 700.0)), target)
        active = active & (cdf < target) & ~((term == 0) & (y > mu))
    return np.where(u > 0, y, 0.0)


def default_theta():
    return np.array([DEFAULTS[n] for n in PARAM_NAMES], dtype=F32)


def simulate_data(seed, T, theta=None, nstep=7):
    model = SEIRModel()
    theta = default_theta() if theta is None else np.asarray(theta, dtype=F32)
    ser = fl.make_series(np.zeros((T, 1)), np.arange(1, T + 1, dtype=float), 0.0, nstep=nstep)
    k = tf.split(fl.pomp_new_key(tf.key(seed)), 1)[0]
    _, Y = fl.simulate(model, theta, ser, k)
    return Y[:, 0, :].astype(F32), theta


def simulate_panel(seed, U, Tmin=52, Tmax=520):
    """Per-unit beta ~ LogNormal(log 1.5, 0.2) and T ~ U{Tmin..Tmax}, from key(seed):
    unit key ku = split(key(seed), U)[u]; beta from normal(split(ku)[0]), T from
    uniform(split(ku)[1]); data simulated from the key split(ku, 3)[2]."""
    keys = tf.split(tf.key(seed), U)
    betas, Ts, ys = [], [], []
    for u in range(U):
        ku = tf.split(keys[u], 3)
        z = tf.normal(ku[0], 1)[0]
        beta = F32(np.exp(np.log(1.5) + 0.2 * float(z)))
        Tu = Tmin + int(np.floor(float(tf.uniform(ku[1])) * (Tmax - Tmin + 1)))
        theta = default_theta()
        theta[0] = beta
        model = SEIRModel()
        ser = fl.make_series(np.zeros((Tu, 1)), np.arange(1, Tu + 1, dtype=float), 0.0, nstep=7)
        kk = tf.split(fl.pomp_new_key(ku[2]), 1)[0]
        _, Y = fl.simulate(model, theta, ser, kk)
        betas.append(beta)
        Ts.append(Tu)
        ys.append(Y[:, 0, 0])
    off = np.concatenate([[0], np.cumsum(Ts)]).astype(np.int64)
    return {"beta": np.array(betas, F32), "T": np.array(Ts, np.int64), "off": off,
            "ys": np.concatenate(ys).astype(F32)}
