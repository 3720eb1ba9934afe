"""Filtering algorithms restated in NumPy -- TEST ORACLE ONLY.

Each function follows a reference function step for step (file:line cited in
its docstring), including the key tree, the float32 arithmetic and the special
cases.  ``resampler`` selects the systematic resampler: ``"fixed"`` (the
build's order-independent definition, which the HIP kernels implement) or
``"reference"`` (the literal float32-cumsum algorithm).

A model object provides ``S`` (state words), ``accumvars`` (indices reset to 0
at the start of each observation interval), ``rinit(theta, keys, t0)``,
``rproc_step(X, theta, keys, t, dt)``, ``dmeas(y, X, theta, t)`` and
``rmeas(X, theta, keys, t)``; theta is (d,) or per particle (J, d), on the
natural scale.  ``from_est``/``to_est`` map parameter arrays between scales.
A model with ``uses_covars`` also takes ``cov``, the covariate row of the
call (the reference's covars_extended[t_idx]); a model with ``vectorized``
gets the Euler sub-step key itself instead of per-particle keys
(core/model_mechanics.py:553-568).
"""

from dataclasses import dataclass

import numpy as np

from . import threefry as tf
from .detmath import d_log, d_pow, dd_log
from .resample import fixed_weights, resample_fixed_fast, resample_reference, shift_for

F32 = np.float32


@dataclass
class Series:
    """Observation series and Euler grid (reference: core/algorithms/helpers.py:191-299)."""
    ys: np.ndarray          # (T, dy) float32
    times: np.ndarray       # (T,) float32
    t0: float
    nstep: np.ndarray       # (T,) int
    dt_ext: np.ndarray      # (sum nstep,) float32
    covars: np.ndarray = None  # (sum nstep + 1, dc) float32 on the grid, or None

    @property
    def T(self):
        return self.ys.shape[0]


def interp_covars(t, ctimes, covars, order="linear"):
    """``_interp_covars`` (core/algorithms/helpers.py:218-262), float64:
    "linear" takes the bracketing covariate times (the first or last pair
    beyond the table, extrapolating) and a + (b - a) frac; "constant" the
    right-continuous step function, the first value before the table."""
    if order not in ("linear", "constant"):
        raise ValueError(f"Unsupported interpolation order: '{order}'. Must be 'linear' or 'constant'.")
    t = np.atleast_1d(np.asarray(t, dtype=np.float64))
    ct = np.asarray(ctimes, dtype=np.float64)
    cv = np.asarray(covars, dtype=np.float64)
    cv2 = cv.reshape(cv.shape[0], -1)
    out = np.empty((t.shape[0], cv2.shape[1]))
    n = ct.shape[0]
    for q, tq in enumerate(t):
        if order == "linear":
            up = int(np.searchsorted(ct, tq, side="left"))
            up = min(max(up, 1), n - 1)
            lo = up - 1
            frac = (tq - ct[lo]) / (ct[up] - ct[lo])
            out[q] = cv2[lo] + (cv2[up] - cv2[lo]) * frac
        else:
            idx = min(max(int(np.searchsorted(ct, tq, side="right")), 1), n) - 1
            out[q] = cv2[idx]
    return out if cv.ndim > 1 else out[:, 0]


def make_series(ys, times, t0, nstep=None, dt=None, ctimes=None, covars=None, order="linear"):
    """Fixed ``nstep`` or fixed ``dt`` Euler grid, as ``_calc_steps`` (helpers.py:158-215):
    the loop rebinds ``dt`` to each interval's result, so the next interval
    starts from the previous interval's step, not the caller's."""
    times = np.asarray(times, dtype=np.float64)
    times0 = np.concatenate([[float(t0)], times])
    ns, dts = [], []
    tol = np.sqrt(np.finfo(float).eps)
    h = None if dt is None else float(dt)
    for i in range(len(times)):
        t1, t2 = float(times0[i]), float(times0[i + 1])
        if nstep is not None:
            n, h = int(nstep), (t2 - t1) / int(nstep)
        elif t1 >= t2:
            n, h = 0, 0.0
        elif t1 + h >= t2:
            n, h = 1, t2 - t1
        else:
            n = int(np.ceil((t2 - t1) / h / (1 + tol)))
            h = (t2 - t1) / n
        ns.append(n)
        dts.append(h)
    ns = np.array(ns, dtype=np.int64)
    dt_ext = np.repeat(np.array(dts), ns).astype(F32)
    cv = None
    if covars is not None:
        # grid of _calc_interp_covars (helpers.py:265-276): each interval's
        # sub-step start times, then the last observation time
        grid = [float(times0[i]) + k * dts[i] for i in range(len(ns)) for k in range(int(ns[i]))]
        grid.append(float(times0[-1]))
        cvals = np.asarray(covars, dtype=np.float64)
        cv = interp_covars(np.array(grid), ctimes, cvals.reshape(cvals.shape[0], -1), order).astype(F32)
    return Series(ys=np.asarray(ys, dtype=F32).reshape(len(times), -1),
                  times=times.astype(F32), t0=float(t0), nstep=ns, dt_ext=dt_ext, covars=cv)


def _cov_kw(model, series, i):
    """The covariate row of grid index i, for a model that reads covariates."""
    if not getattr(model, "uses_covars", False):
        return {}
    return {"cov": series.covars[i]}


def rinit_at(model, th, pkeys, series):
    """``rinit`` with the covariates of grid row 0 (pfilter.py:28-35)."""
    return model.rinit(th, pkeys, series.t0, **_cov_kw(model, series, 0))


def dmeas_at(model, y, X, th, t, series, t_idx):
    """``dmeas`` with the covariates of grid row t_idx (pfilter.py:124-133)."""
    return model.dmeas(y, X, th, t, **_cov_kw(model, series, t_idx))


def rprocess_interp(model, X, theta, key, series, t, t_idx, nstep):
    """Euler sub-step loop (core/model_mechanics.py:908-947 and step_pf :584-588):
    reset accumvars; per sub-step ``(next, sk) = split(key)``, particle keys
    ``split(sk, J)``, one model step, ``t += dt``."""
    X = X.copy()
    for a in model.accumvars:
        X[:, a] = 0
    J = X.shape[0]
    t = F32(t)
    for _ in range(int(nstep)):
        k2 = tf.split(key)
        nxt, sk = k2[0], k2[1]
        # @vectorized: the sub-step key itself (step_vec, model_mechanics.py:553-568)
        pkeys = sk if getattr(model, "vectorized", False) else tf.split(sk, J)
        dt = series.dt_ext[t_idx]
        X = model.rproc_step(X, theta, pkeys, t, dt, **_cov_kw(model, series, t_idx))
        t = F32(t + dt)
        t_idx += 1
        key = nxt
    return X, t_idx


def step_loglik(m, total, J):
    """The engine's log-likelihood term of one observation (float64):
    m + log(total 2^-sh), total = exact integer sum of the resampler's
    fixed-point weights (the device's step_loglik).  m = -inf gives -inf; +inf or
    NaN give NaN."""
    m = F32(m)
    if m == -np.inf:
        return -np.inf
    if not (m < np.inf):
        return np.nan
    return float(np.float64(m) + dd_log(np.ldexp(np.float64(int(total)), -shift_for(J))))


def normalize_weights(w, engine=True):
    """``_normalize_weights`` (core/algorithms/helpers.py:34-57).  ``engine``:
    the log-likelihood term is ``step_loglik`` of the fixed-point weights
    (float64), the normalised weights ``w - fl32(term)``; otherwise the
    reference's float32 arithmetic literally."""
    w = np.asarray(w, dtype=F32)
    J = w.shape[-1]
    mw = np.max(w)
    if mw == -np.inf:
        return np.full(J, -d_log(F32(J)), dtype=F32), (-np.inf if engine else F32(-np.inf))
    if engine:
        total = int(np.sum(fixed_weights(w), dtype=np.uint64))
        ll = step_loglik(mw, total, J)
        with np.errstate(invalid="ignore"):
            nw = (w - F32(ll)).astype(F32)
        return nw, ll
    with np.errstate(invalid="ignore", over="ignore"):
        ll = F32(mw + np.log(np.nansum(np.exp((w - mw).astype(F32)), dtype=F32)))
        nw = (w - ll).astype(F32)
    return nw, ll


def _resample(resampler, nw, w_unnorm, key):
    u = tf.uniform(key)
    if resampler == "reference":
        return resample_reference(nw, u)
    return resample_fixed_fast(w_unnorm, u)


def pfilter(model, theta, series, key, J, thresh=0.0, resampler="fixed",
            diagnostics=False, should_trans=False, return_trace=False):
    """One bootstrap particle filter: ``_pfilter_internal`` + ``_pfilter_step``
    (core/algorithms/pfilter.py:15-182).  Returns ``logLik`` and, with
    ``diagnostics``, CLL/ESS/filter and prediction means; ``return_trace`` adds
    per-step ancestors and weights.  resampler "fixed" is the engine's
    definition (fixed-point resampler, float64 sum over t of ``step_loglik``,
    also as ``logLik64``); "reference" the reference's float32 arithmetic."""
    theta = np.asarray(theta, dtype=F32)
    th = model.from_est(theta) if should_trans else theta
    K = tf.split(key, J + 1)
    key, pkeys = K[0], K[1:]
    X = rinit_at(model, th, pkeys, series)
    nw = np.full(J, d_log(F32(1.0) / F32(J)), dtype=F32)
    engine = resampler != "reference"
    loglik = 0.0 if engine else F32(0.0)
    t, t_idx = F32(series.t0), 0
    out = {"CLL": [], "ESS": [], "filter_mean": [], "prediction_mean": [],
           "ancestors": [], "weights": [], "states": []}
    logthresh = d_log(F32(thresh)) if thresh > 0 else -np.inf
    for i in range(series.T):
        k2 = tf.split(key)
        key, sub = k2[0], k2[1]
        X, t_idx = rprocess_interp(model, X, th, sub, series, t, t_idx, series.nstep[i])
        t = series.times[i]
        meas = dmeas_at(model, series.ys[i], X, th, t, series, t_idx).astype(F32)
        w = (nw + meas).astype(F32)
        nw, ll_t = normalize_weights(w, engine)
        loglik = loglik + ll_t if engine else F32(loglik + ll_t)
        if diagnostics:
            out["CLL"].append(F32(ll_t))
            out["ESS"].append(F32(1.0) / np.sum(np.exp(F32(2.0) * nw), dtype=F32))
            out["filter_mean"].append((X * np.exp(nw)[:, None]).sum(0, dtype=F32))
            out["prediction_mean"].append(X.mean(0, dtype=F32))
        if return_trace:
            out["weights"].append(w.copy())
            out["states"].append(X.copy())
        with np.errstate(invalid="ignore"):
            do_res = (np.max(nw) - np.min(nw)) > logthresh
        k2 = tf.split(key)
        key, sub = k2[0], k2[1]
        if do_res:
            a = _resample(resampler, nw, w, sub)
            X = X[a]
            nw = np.full(J, F32(0.0) - d_log(F32(J)), dtype=F32)
        else:
            a = np.arange(J, dtype=np.int32)
        if return_trace:
            out["ancestors"].append(a)
    out["logLik"] = F32(loglik)
    if engine:
        out["logLik64"] = loglik
    return out


def mop(model, theta_est, series, key, J, alpha, resampler="fixed", return_trace=False):
    """``_mop_internal`` + ``_mop_step`` (core/algorithms/mop.py:19-135): theta on the
    estimation scale; returns the MOP log-likelihood (the reference returns its
    negative).  Value-wise the weights stay constant across particles, so each
    term LSE(wP + m) - LSE(wP) is max(m) + log sum exp(m - max) - log J: with the
    "fixed" resampler the engine's float64 ``step_loglik`` minus fl32(log J),
    summed in float64; with "reference" the reference's float32 LSEs."""
    th = model.from_est(np.asarray(theta_est, dtype=F32))
    K = tf.split(key, J + 1)
    key, pkeys = K[0], K[1:]
    X = rinit_at(model, th, pkeys, series)
    wF = np.full(J, d_log(F32(1.0) / F32(J)), dtype=F32)
    engine = resampler != "reference"
    loglik = 0.0 if engine else F32(0.0)
    logJ = float(d_log(F32(J)))
    t, t_idx = F32(series.t0), 0
    trace = {"ancestors": [], "meas": [], "states": []}
    for i in range(series.T):
        wP = (F32(alpha) * wF).astype(F32)
        k2 = tf.split(key)
        key, sub = k2[0], k2[1]
        X, t_idx = rprocess_interp(model, X, th, sub, series, t, t_idx, series.nstep[i])
        t = series.times[i]
        meas = dmeas_at(model, series.ys[i], X, th, t, series, t_idx).astype(F32)
        if engine:
            mm = np.max(meas)
            loglik = loglik + (step_loglik(mm, int(np.sum(fixed_weights(meas), dtype=np.uint64)), J) - logJ)
        else:
            loglik = F32(loglik + lse(wP + meas) - lse(wP))
        nw, _ = normalize_weights(meas, engine)
        k2 = tf.split(key)
        key, sub = k2[0], k2[1]
        a = _resample(resampler, nw, meas, sub)
        if return_trace:
            trace["ancestors"].append(a)
            trace["meas"].append(meas)
            trace["states"].append(X.copy())
        X = X[a]
        wF = wP[a]
    if return_trace:
        return loglik, trace
    return loglik


def lse(a):
    """``jax.scipy.special.logsumexp`` in float32."""
    a = np.asarray(a, dtype=F32)
    amax = np.max(a)
    if not np.isfinite(amax):
        amax = F32(0.0)
    with np.errstate(divide="ignore"):
        return F32(np.log(np.sum(np.exp((a - amax).astype(F32)), dtype=F32)) + amax)


@dataclass
class RWSigma:
    """Random-walk sigmas and the cooling schedules (core/rw_sigma.py:184-225):
    geometric (base a), cosine (final factor c over M iterations), hyperbolic
    (rate s) or custom (fn(nt, m, ntimes))."""
    sigmas: np.ndarray      # (d,) canonical order
    init_mask: np.ndarray   # (d,) 1.0 for initial-value parameters
    a: float = 0.5
    cooling: str = "geometric"
    c: float = None
    M: int = None
    s: float = None
    fn: object = None

    @property
    def sigmas_array(self):
        return (self.sigmas.astype(F32) * (F32(1.0) - self.init_mask.astype(F32))).astype(F32)

    @property
    def sigmas_init_array(self):
        return (self.sigmas.astype(F32) * self.init_mask.astype(F32)).astype(F32)

    def cooling_factor(self, nt, m, ntimes):
        """frac = nt / ntimes + m in float32, then the schedule (rw_sigma.py:209-223)."""
        frac = F32(F32(nt) / F32(ntimes) + F32(m))
        if self.cooling == "cosine":
            progress = F32(frac / F32(self.M))
            half = F32((1.0 - self.c) * 0.5)   # a Python-float product, rounded once
            return F32(F32(self.c) + F32(half * F32(F32(1) + np.cos(F32(F32(np.pi) * progress), dtype=F32))))
        if self.cooling == "hyperbolic":
            return F32(F32(1) / F32(F32(1) + F32(F32(self.s) * frac)))
        if self.cooling == "custom":
            return F32(self.fn(nt, m, ntimes))
        factor = F32(self.a ** (1.0 / 50.0))
        return F32(d_pow(factor, frac))


def perfilter(model, m, thetas, series, key, J, rw, thresh=0.0, resampler="fixed", return_ancestry=False):
    """``_perfilter_internal`` + ``_perfilter_scan_body`` (core/algorithms/mif.py:88-258).
    ``thetas`` (J, d) on the estimation scale.  Returns (thetas, loglik), plus the
    composed resampling ancestry (mif.py:237-245) when ``return_ancestry``."""
    thetas = np.asarray(thetas, dtype=F32).copy()
    d = thetas.shape[1]
    T = series.T
    nw = np.full(J, -d_log(F32(J)), dtype=F32)
    engine = resampler != "reference"
    loglik = 0.0 if engine else F32(0.0)
    cf = [rw.cooling_factor(i, m, T) for i in range(T)]
    k2 = tf.split(key)
    key, sub = k2[0], k2[1]
    sic = (rw.cooling_factor(0, m, T) * rw.sigmas_init_array).astype(F32)
    thetas = (thetas + sic * tf.normal(sub, J * d).reshape(J, d)).astype(F32)
    K = tf.split(key, J + 1)
    key, pkeys = K[0], K[1:]
    X = rinit_at(model, model.from_est(thetas), pkeys, series)
    step_keys = tf.split(key, T + 1)[1:]
    t, t_idx = F32(series.t0), 0
    logthresh = d_log(F32(thresh)) if thresh > 0 else -np.inf
    nan_sub = d_log(F32(1e-18))
    sig = rw.sigmas_array
    anc = np.arange(J)
    for i in range(T):
        k3 = tf.split(step_keys[i], 3)
        kpert, kproc, kres = k3[0], k3[1], k3[2]
        thetas = (thetas + (cf[i] * sig).astype(F32) * tf.normal(kpert, J * d).reshape(J, d)).astype(F32)
        th_nat = model.from_est(thetas)
        X, t_idx = rprocess_interp(model, X, th_nat, kproc, series, t, t_idx, series.nstep[i])
        t = series.times[i]
        meas = dmeas_at(model, series.ys[i], X, th_nat, t, series, t_idx).astype(F32)
        meas = np.nan_to_num(meas, nan=nan_sub).astype(F32)
        w = (nw + meas).astype(F32)
        nw, ll_t = normalize_weights(w, engine)
        loglik = loglik + ll_t if engine else F32(loglik + ll_t)
        with np.errstate(invalid="ignore"):
            do_res = (np.max(nw) - np.min(nw)) > logthresh
        if do_res:
            a = _resample(resampler, nw, w, kres)
            X = X[a]
            thetas = thetas[a]
            anc = anc[a]
            nw = np.full(J, F32(0.0) - d_log(F32(J)), dtype=F32)
    return (thetas, loglik, anc) if return_ancestry else (thetas, loglik)


def mif(model, thetas_nat, series, key, J, M, rw, thresh=0.0, n_monitors=0,
        resampler="fixed"):
    """``_mif_internal`` / ``_mif_scan_body`` (core/algorithms/mif.py:16-85) plus the
    scale handling of ``F.mif`` (functional/mif.py:85-122).  ``thetas_nat`` (J, d)
    natural scale.  Returns (logliks (M,), traces (M+1, d) natural, final swarm)."""
    theta = model.to_est(np.asarray(thetas_nat, dtype=F32))
    mkeys = tf.split(key, M + 1)[1:]
    traces = [theta.mean(0, dtype=F32)]
    lls = []
    for m in range(M):
        new, ll_per = perfilter(model, m, theta, series, mkeys[m], J, rw, thresh, resampler)
        if n_monitors >= 1:
            mean = theta.mean(0, dtype=F32)
            mon = tf.split(mkeys[m], n_monitors)
            vals = [pfilter(model, mean, series, mon[k], J, thresh, resampler,
                            should_trans=True)["logLik"] for k in range(n_monitors)]
            ll = F32(np.mean(np.array(vals, dtype=F32), dtype=F32))
        else:
            ll = ll_per
        lls.append(ll)
        traces.append(new.mean(0, dtype=F32))
        theta = new
    traces = np.stack(traces).astype(F32)
    return (np.array(lls, dtype=F32), model.from_est(traces), model.from_est(theta))


def panel_mif(model, shared_nat, unit_nat, series_list, key, J, M, rw, perm, thresh=0.0, block=True,
              resampler="fixed", par_trans=None):
    """``_panel_mif_internal`` / ``_panel_mif_iter_body`` / ``_panel_mif_unit_scan_fn``
    (core/algorithms/panel_mif.py:17-247) for one replicate, no monitors, with the
    scale handling of ``F.panel_mif`` (functional/mif.py:212-294).

    shared_nat (J, ns), unit_nat (J, U, nu) natural scale; ``perm`` gives, for
    each model-order parameter, its index in the panel order (shared names,
    then unit-specific names); ``rw`` is in model order.  ``par_trans`` (default:
    the model) supplies the panel-level to_est/from_est (PanelPompStruct.par_trans,
    panel/panel.py:146); the filter itself always applies the model's own
    transforms, which the reference binds into the mechanics when the unit Pomp
    is built (core/pomp.py:271-313).  Returns
    (shared_traces (M+1, ns+1), unit_traces (M+1, U, nu+1), final shared, final
    unit), natural scale, log-likelihood in column 0 (NaN at m = 0)."""
    shared_nat = np.asarray(shared_nat, dtype=F32)
    unit_nat = np.asarray(unit_nat, dtype=F32)
    ns, U, nu = shared_nat.shape[1], unit_nat.shape[1], unit_nat.shape[2]
    perm = np.asarray(perm)
    inv = np.argsort(perm)
    pt = model if par_trans is None else par_trans

    def to_model(shared, unit):  # (..., ns), (..., U, nu) -> (..., U, d) model order
        lead = unit.shape[:-2]
        full = np.concatenate([np.broadcast_to(shared[..., None, :], lead + (U, ns)), unit], axis=-1)
        return full[..., perm]

    def to_panel(full):          # (..., U, d) model order -> shared (..., ns), unit (..., U, nu)
        p = full[..., inv]
        return p[..., 0, :ns], p[..., ns:]

    sh, un = to_panel(pt.to_est(to_model(shared_nat, unit_nat)))
    sh, un = sh.astype(F32).copy(), un.astype(F32).copy()
    sh_tr = [np.concatenate([[np.nan], sh.mean(0, dtype=F32)]).astype(F32)]
    un_tr = [np.concatenate([np.full((U, 1), np.nan, F32), un.mean(0, dtype=F32)], axis=1)]
    mkeys = tf.split(key, M + 1)[1:]
    for m in range(M):
        ukeys = tf.split(mkeys[m], U)
        rows, total = [], F32(0.0)
        for u in range(U):
            th = np.concatenate([sh, un[:, u, :]], axis=1)[:, perm]
            new, ll, anc = perfilter(model, m, th, series_list[u], ukeys[u], J, rw, thresh, resampler,
                                     return_ancestry=True)
            upd = new[:, inv]
            if ns:
                sh = upd[:, :ns].copy()
            if not block:
                un = un[anc]
            un = un.copy()
            un[:, u, :] = upd[:, ns:]
            rows.append(np.concatenate([[F32(ll)], upd[:, ns:].mean(0, dtype=F32)]))
            total = F32(total + F32(ll))
        un_tr.append(np.stack(rows).astype(F32))
        sh_tr.append(np.concatenate([[total], sh.mean(0, dtype=F32)]).astype(F32))
    sh_tr, un_tr = np.stack(sh_tr), np.stack(un_tr)
    s_nat, u_nat = to_panel(pt.from_est(to_model(sh_tr[:, 1:], un_tr[..., 1:])))
    sh_tr[:, 1:], un_tr[..., 1:] = s_nat, u_nat
    fs, fu = to_panel(pt.from_est(to_model(sh, un)))
    return sh_tr, un_tr, fs, fu


class IdentityTrans:
    """A model whose estimation scale is its natural scale (the reference's
    ``ParTrans()`` with no transforms, core/par_trans.py:70-111)."""

    def __init__(self, model):
        self._m = model

    def __getattr__(self, name):
        return getattr(self._m, name)

    def to_est(self, theta):
        return np.array(theta, dtype=F32, copy=True)

    def from_est(self, theta):
        return np.array(theta, dtype=F32, copy=True)


def simulate(model, theta, series, key, nsim=1):
    """``_simulate_internal`` / ``_simulate_helper`` (core/algorithms/simulate.py:15-142).
    Returns X (T+1, nsim, S) and Y (T, nsim, dy)."""
    theta = np.asarray(theta, dtype=F32)
    K = tf.split(key, nsim + 1)
    key, skeys = K[0], K[1:]
    X = rinit_at(model, theta, skeys, series)
    Xs, Ys = [X.copy()], []
    t, t_idx = F32(series.t0), 0
    times0 = np.concatenate([[series.t0], series.times]).astype(F32)
    for i in range(series.T):
        k2 = tf.split(key)
        key, sub = k2[0], k2[1]
        X, t_idx = rprocess_interp(model, X, theta, sub, series, t, t_idx, series.nstep[i])
        t = times0[i]
        K = tf.split(key, nsim + 1)
        key, mkeys = K[0], K[1:]
        Y = model.rmeas(X, theta, mkeys, t, **_cov_kw(model, series, t_idx))
        Xs.append(X.copy())
        Ys.append(Y)
    return np.stack(Xs), np.stack(Ys)


def pomp_new_key(key):
    """``Pomp._update_fresh_key`` (core/pomp.py:391-408): returns the key used by the call."""
    return tf.split(key)[1]


def logmeanexp(x):
    """``logmeanexp`` over replicates (pypomp/maths.py:24-84), float64."""
    x = np.asarray(x, dtype=np.float64)
    mx = np.max(x)
    return float(mx + np.log(np.mean(np.exp(x - mx))))
