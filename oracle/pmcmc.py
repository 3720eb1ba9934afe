"""PMMH restated on the host -- TEST ORACLE ONLY.

``_pmcmc_internal`` / ``_pmcmc_step`` (core/algorithms/pmcmc.py:30-163) and
``F.pmcmc`` (functional/pmcmc.py:15-112) with the three proposals of
pypomp/proposals.py, over any log-likelihood function; ``pmcmc`` plugs in the
oracle's bootstrap filter.  Chain arithmetic in float32 as JAX's default.

Pinned: the Metropolis-Hastings kernel by the posterior of a one-parameter
linear-Gaussian model under the exact Kalman likelihood (grid integration);
the adaptive proposal's running covariance by a batch covariance of the same
states; the change-of-variables term by finite differences of float64
transforms (tests/test_oracle_pmcmc.py).
"""

import numpy as np

from . import filters as fl
from . import threefry as tf

F32 = np.float32

# elementwise transform of each parameter from the estimation scale
ID, EXP, TANH, LOGISTIC = 0, 1, 2, 3


def transform_kinds(model):
    """Per-parameter transform kind of an oracle model's from_est."""
    from .lg import LGModel
    from .seir import SEIRModel
    from .sv import SVModel
    if isinstance(model, fl.IdentityTrans):
        return [ID] * len(model.param_names)
    if isinstance(model, LGModel):
        return [EXP if (n[0] in "QR" and len(n) == 3 and n[1] == n[2]) else ID for n in model.spec.param_names]
    if isinstance(model, SVModel):
        return [ID, TANH, EXP, TANH, ID, EXP]
    if isinstance(model, SEIRModel):
        return [EXP, LOGISTIC, EXP, EXP, LOGISTIC, EXP, EXP, LOGISTIC, EXP, EXP]
    raise TypeError(type(model))


def log_jac(kinds, th_est):
    """log |det d from_est / d theta| for elementwise transforms (float64)."""
    out = 0.0
    for k, x in zip(kinds, np.asarray(th_est, np.float64)):
        if k == EXP:
            out += x
        elif k == TANH:
            out += np.log1p(-np.tanh(x) ** 2) if abs(x) < 10 else np.log(4.0) - 2 * abs(x)
        elif k == LOGISTIC:
            s = 1.0 / (1.0 + np.exp(-x))
            out += np.log(s * (1.0 - s))
    return out


def diag_rw(sd):
    sd = np.asarray(sd, np.float64)

    def step(state, th, z, n, acc):
        return (th + (z * sd.astype(F32)).astype(F32)).astype(F32), state
    return (lambda th: None), step


def full_rw(chol):
    """MVNRWFull with the (canonicalized) Cholesky factor ``chol``."""
    L = np.asarray(chol, np.float64).astype(F32)

    def step(state, th, z, n, acc):
        d = len(th)
        out = np.empty(d, F32)
        for i in range(d):
            a = F32(0)
            for j in range(d):
                a = F32(a + F32(L[i, j] * z[j]))
            out[i] = a
        return (th + out).astype(F32), state
    return (lambda th: None), step


def adaptive_rw(init_var, scale_start=200, scale_cooling=0.999, shape_start=200, target=0.234, max_scaling=50.0):
    """MVNRWAdaptive (proposals.py): Roberts & Rosenthal (2009) two-phase
    adaptation with a Welford running mean and covariance."""
    V = np.asarray(init_var, np.float64)
    active = (np.any(V != 0, axis=0) | np.any(V != 0, axis=1)).astype(F32)

    def init(th):
        d = len(th)
        return [F32(1), np.zeros(d, F32), np.zeros((d, d), F32), False]

    def step(state, th, z, n, acc):
        scaling, mean, cov_emp, seeded = state
        d = len(th)
        nf, af = F32(max(n, 1)), F32(acc)
        mean0 = mean if seeded else th
        phase1 = n >= scale_start and acc < shape_start
        if phase1:
            cool = np.power(F32(scale_cooling), F32(max(n - scale_start, 0)), dtype=F32)
            scaling = F32(min(F32(scaling * np.exp(F32(cool * F32(F32(af / nf) - F32(target))), dtype=F32)),
                              F32(max_scaling)))
        if acc >= shape_start:
            C = (F32(2.38 ** 2 / d) * cov_emp).astype(F32)
        elif phase1:
            C = (F32(scaling * scaling) * V.astype(F32)).astype(F32)
        else:
            C = V.astype(F32)
        new_mean = (mean0 + ((th - mean0) / nf).astype(F32)).astype(F32)
        upd = np.outer((th - mean0).astype(F32), (th - new_mean).astype(F32)).astype(F32)
        new_cov = ((((nf - F32(1)) * cov_emp).astype(F32) + upd).astype(F32) / nf).astype(F32)
        L = np.linalg.cholesky((C + F32(1e-10) * np.eye(d, dtype=F32)).astype(F32)).astype(F32)
        inc = np.empty(d, F32)
        for i in range(d):
            a = F32(0)
            for j in range(d):
                a = F32(a + F32(L[i, j] * z[j]))
            inc[i] = a
        return (th + (inc * active).astype(F32)).astype(F32), [scaling, new_mean, new_cov, True]
    return init, step


def mh_chain(loglik, log_prior, proposal, theta_est0, key, M, d):
    """_pmcmc_internal for one chain over ``loglik(theta_est, key) -> float32``
    and ``log_prior(theta_est) -> float32``.  Returns (ll (M+1), theta_est
    (M+1, d), accepts)."""
    init, step = proposal
    th = np.asarray(theta_est0, F32)
    k2 = tf.split(key)
    key, k_init = k2[0], k2[1]
    ll = F32(loglik(th, k_init))
    lp = F32(log_prior(th))
    state = init(th)
    acc = 0
    lls, ths = [ll], [th.copy()]
    for n in range(1, M + 1):
        ks = tf.split(key, 4)
        key, kp, kf, ka = ks[0], ks[1], ks[2], ks[3]
        z = tf.normal_many(kp[None], d)[0]
        prop, state = step(state, th, z, n, acc)
        lp_p = F32(log_prior(prop))
        ll_p = F32(loglik(prop, kf))
        u = tf.uniform_from_bits(tf.random_bits_many(ka[None], 1)[:, 0])[0]
        with np.errstate(invalid="ignore", divide="ignore"):
            la = F32(F32(F32(ll_p + lp_p) - ll) - lp)
            ok = bool(np.isfinite(la) and np.isfinite(lp_p) and np.log(u) < la)
        if ok:
            th, ll, lp = prop, ll_p, lp_p
            acc += 1
        lls.append(ll)
        ths.append(th.copy())
    return np.array(lls, F32), np.stack(ths), acc


def pmcmc(model, series, theta_nat0, proposal, M, J, key, thresh=0.0, dprior=None):
    """F.pmcmc for one chain with the oracle's bootstrap filter (should_trans).
    ``dprior(theta_nat) -> float`` on the natural scale, flat when None."""
    kinds = transform_kinds(model)
    names = list(model.param_names)

    def prior_nat(th_nat):
        return F32(0.0) if dprior is None else F32(dprior(dict(zip(names, th_nat))))

    def prior_est(th):
        return F32(prior_nat(model.from_est(th)) + F32(log_jac(kinds, th)))

    def loglik(th, k):
        return F32(fl.pfilter(model, th, series, k, J, thresh, should_trans=True)["logLik"])

    th0 = np.asarray(model.to_est(np.asarray(theta_nat0, F32)), F32)
    ll, est, acc = mh_chain(loglik, prior_est, proposal, th0, key, M, len(th0))
    nat = np.stack([model.from_est(e) for e in est]).astype(F32)
    lp = np.array([prior_nat(t) for t in nat], F32)
    return ll, lp, nat, acc
