"""Array-level API mirroring ``pypomp.functional`` for the particle-filter path.

``pfilter`` (functional/pfilter.py:11-90), ``mop`` (functional/mop.py:8-62),
``mif`` (functional/mif.py:15-122) and ``panel_mif`` for panels whose
parameters are all unit-specific (functional/mif.py:125-294) -- same argument
meaning, shapes and scales, with NumPy in and out.  Everything runs in
libmipomp.so on the GPU; ``form`` picks the batched (one persistent launch)
or the plain (one filter at a time) execution.
"""

from dataclasses import dataclass, field

import numpy as np

from . import _native
from . import random as prandom
from .series import DeviceSeries, Series

F32 = np.float32


@dataclass
class PompStruct:
    """Model + data of one Pomp (or the units of a panel), cf. functional/structs.py:10-81."""
    model: object
    units: list
    _dev: dict = field(default_factory=dict, repr=False)

    @property
    def param_names(self):
        return list(self.model.param_names)

    def device_series(self):
        import torch
        dev = _native.require_device()
        k = (dev.index,)
        if k not in self._dev:
            self._dev[k] = DeviceSeries(self.units, dev)
        return self._dev[k]


# The plain form (one filter at a time over the whole GPU) pays two launches
# per observation; the batched form runs one workgroup per filter.  A filter
# of at least this many particles fills the GPU on its own, so when fewer
# filters than CUs are asked for, the plain form runs them faster.
PLAIN_MIN_PARTICLES = 65536


def choose_form(R, J, multi_unit=False):
    """The execution form for R filters of J particles: "plain" when the
    filters are single-series, fewer than the GPU's compute units and each has
    at least PLAIN_MIN_PARTICLES particles; otherwise "batched"."""
    import torch
    if multi_unit or J < PLAIN_MIN_PARTICLES:
        return "batched"
    cus = torch.cuda.get_device_properties(_native.require_device()).multi_processor_count
    return "plain" if R < cus else "batched"


def _form_id(form):
    if form in ("batched", _native.MP_FORM_BATCHED):
        return _native.MP_FORM_BATCHED
    if form in ("plain", _native.MP_FORM_PLAIN):
        return _native.MP_FORM_PLAIN
    raise ValueError(f"unknown form {form!r}")


def _keys_tensor(keys, dev):
    import torch
    k = np.ascontiguousarray(np.asarray(keys, dtype=np.uint32).reshape(-1, 2))
    return torch.from_numpy(k.view(np.int32)).to(dev)


def _unit_tensor(unit, R, dev, U):
    """(R,) int32 device tensor of panel unit indices in [0, U), or None."""
    import torch
    if unit is None:
        if U != 1:
            raise ValueError(f"a struct of {U} units needs the unit of every row")
        return None
    u = np.array(np.broadcast_to(np.asarray(unit, dtype=np.int32), (R,)))
    if u.size and (u.min() < 0 or u.max() >= U):
        raise ValueError(f"unit indices must lie in [0, {U})")
    return torch.from_numpy(u).to(dev)


def _workspace(algo, md, ser, R, J, form, dev):
    import torch
    n = _native.lib().mp_filter_workspace_size(algo, md, ser.desc, R, J, form)
    if n == 0:
        raise _native.NativeError("workspace size query failed")
    return torch.empty(int(n), dtype=torch.uint8, device=dev), int(n)


def pfilter(struct, thetas_array, J, keys, thresh=0.0, CLL=False, ESS=False, filter_mean=False,
            prediction_mean=False, form="auto", should_trans=False, unit=None):
    """Bootstrap particle filter.  thetas (n, d) natural scale, keys (n, reps, 2);
    ``unit`` (n,) selects each parameter set's unit of a multi-unit struct.
    ``form``: "batched", "plain" or "auto" (``choose_form``); both forms give
    the same results bit for bit.  Returns a dict with ``logLik`` (n, reps) and
    the requested diagnostics."""
    import torch
    dev = _native.require_device()
    thetas = np.asarray(thetas_array, dtype=F32)
    if thetas.ndim == 1:
        thetas = thetas[None]
    keys = np.asarray(keys, dtype=np.uint32)
    if keys.ndim == 2:
        keys = keys[:, None, :]
    n, reps = keys.shape[0], keys.shape[1]
    if thetas.shape[0] != n:
        raise ValueError("thetas and keys disagree on the number of parameter sets")
    if J < 1:
        raise ValueError("J should be greater than 0.")
    thresh = float(max(0.0, thresh))
    R = n * reps
    if form == "auto":
        form = choose_form(R, J, multi_unit=unit is not None or len(struct.units) > 1)
    ser = struct.device_series()
    md = struct.model.desc()
    th = torch.from_numpy(np.repeat(thetas, reps, axis=0)).to(dev)
    kt = _keys_tensor(keys.reshape(R, 2), dev)
    ut = None
    if unit is not None:
        if form not in ("batched", _native.MP_FORM_BATCHED):
            raise ValueError("multi-unit filters run in the batched form")
        ut = torch.from_numpy(np.repeat(np.asarray(unit, dtype=np.int32).reshape(n), reps)).to(dev)
    T, S = ser.Tmax, struct.model.S
    ll = torch.empty(R, dtype=torch.float64, device=dev)
    cll = torch.empty((R, T), dtype=torch.float32, device=dev) if CLL else None
    ess = torch.empty((R, T), dtype=torch.float32, device=dev) if ESS else None
    fm = torch.empty((R, T, S), dtype=torch.float32, device=dev) if filter_mean else None
    pm = torch.empty((R, T, S), dtype=torch.float32, device=dev) if prediction_mean else None
    flags = ((_native.MP_CLL if CLL else 0) | (_native.MP_ESS if ESS else 0)
             | (_native.MP_FILTER_MEAN if filter_mean else 0)
             | (_native.MP_PREDICTION_MEAN if prediction_mean else 0)
             | (_native.MP_SHOULD_TRANS if should_trans else 0))
    fid = _form_id(form)
    ws, nws = _workspace(_native.MP_ALGO_PF, md, ser, R, J, fid, dev)
    _native.check(_native.lib().mp_pfilter(
        md, ser.desc, _native.ptr(th), struct.model.d, _native.ptr(kt), _native.ptr(ut), R, J, float(thresh), flags,
        fid, _native.ptr(ll), _native.ptr(cll), _native.ptr(ess), _native.ptr(fm), _native.ptr(pm),
        _native.ptr(ws), nws, _native.stream_ptr()))
    torch.cuda.current_stream().synchronize()
    out = {"logLik": ll.cpu().numpy().astype(F32).reshape(n, reps),
           "logLik64": ll.cpu().numpy().reshape(n, reps)}
    if CLL:
        out["CLL"] = cll.cpu().numpy().reshape(n, reps, T)
    if ESS:
        out["ESS"] = ess.cpu().numpy().reshape(n, reps, T)
    if filter_mean:
        out["filter_mean"] = fm.cpu().numpy().reshape(n, reps, T, S)
    if prediction_mean:
        out["prediction_mean"] = pm.cpu().numpy().reshape(n, reps, T, S)
    return out


def simulate(struct, thetas_array, nsim, keys, times=None, should_trans=False):
    """Simulate states and observations (functional/simulate.py:11-78;
    core/algorithms/simulate.py:15-142).  thetas (n, d) natural scale, keys
    (n, 2); ``times`` (T,) replaces the observation times (the Euler grid stays
    the struct's).  Returns X (n, nsim, T+1, S) and Y (n, nsim, T, dy), float32."""
    import torch
    dev = _native.require_device()
    thetas = np.ascontiguousarray(np.asarray(thetas_array, dtype=F32))
    if thetas.ndim == 1:
        thetas = thetas[None]
    keys = np.asarray(keys, dtype=np.uint32).reshape(-1, 2)
    n = keys.shape[0]
    if thetas.shape[0] != n:
        raise ValueError("thetas and keys disagree on the number of parameter sets")
    if len(struct.units) != 1:
        raise ValueError("simulate runs single-series models")
    if int(nsim) < 1:
        raise ValueError("nsim must be at least 1")
    ser = struct.device_series()
    md = struct.model.desc()
    T, S, dy = ser.Tmax, struct.model.S, struct.units[0].ys.shape[1]
    tt = None
    if times is not None:
        times = np.asarray(times, dtype=F32).reshape(-1)
        if times.shape[0] != T:
            raise ValueError(f"times must have the series' length {T} (the Euler grid is the struct's)")
        tt = torch.from_numpy(times).to(dev)
    th = torch.from_numpy(thetas).to(dev)
    kt = _keys_tensor(keys, dev)
    X = torch.empty((n, T + 1, S, int(nsim)), dtype=torch.float32, device=dev)
    Y = torch.empty((n, T, dy, int(nsim)), dtype=torch.float32, device=dev)
    _native.check(_native.lib().mp_simulate(md, ser.desc, _native.ptr(th), struct.model.d, _native.ptr(kt),
                                            _native.ptr(tt), n, int(nsim),
                                            _native.MP_SHOULD_TRANS if should_trans else 0,
                                            _native.ptr(X), _native.ptr(Y), _native.stream_ptr()))
    Xo = X.permute(0, 3, 1, 2).contiguous().cpu().numpy()
    Yo = Y.permute(0, 3, 1, 2).contiguous().cpu().numpy()
    return Xo, Yo


def _est_prior(model, dprior):
    """(estimation-scale log prior with the change-of-variables term, natural
    log prior) for a user prior ``dprior(theta: dict) -> float`` on the natural
    scale, or a flat prior (core/model_mechanics.py:799-867;
    functional/structs.py:271-295)."""
    names = list(model.param_names)

    def natural(th_nat):
        if dprior is None:
            return F32(0.0)
        return F32(dprior({n: F32(v) for n, v in zip(names, th_nat)}))

    def est(th_est):
        return F32(natural(model.from_est(th_est)) + model.log_jac_from_est(th_est))

    return est, natural


def pmcmc(struct, thetas_array, proposal, M, J, keys, thresh=0.0, dprior=None):
    """Particle marginal Metropolis-Hastings (functional/pmcmc.py:15-112;
    core/algorithms/pmcmc.py:30-163): one chain per row of thetas (n, d,
    natural scale), keys (n, 2).  The chain runs on the estimation scale; each
    step draws (key, prop, pf, accept) = split(key, 4), proposes, evaluates the
    prior and a bootstrap particle filter at the proposal and accepts when
    log(u) < ll' + lp' - ll - lp with both finite.  The n filters of a step run
    in one launch; the proposal normals and accept uniforms of all steps are
    drawn on the device up front.  Returns (loglik (n, M+1), log_prior (n,
    M+1) natural scale, theta (n, M+1, d) natural scale, accepts (n,))."""
    import torch
    dev = _native.require_device()
    model = struct.model
    if J < 1:
        raise ValueError("J should be greater than 0.")
    th_nat0 = np.ascontiguousarray(np.asarray(thetas_array, dtype=F32))
    if th_nat0.ndim == 1:
        th_nat0 = th_nat0[None]
    keys = np.asarray(keys, dtype=np.uint32).reshape(-1, 2)
    n, d = th_nat0.shape
    if keys.shape[0] != n:
        raise ValueError("thetas and keys disagree on the number of chains")
    M = int(M)
    proposal = proposal.canonicalize(model.param_names)
    lp_est_fn, lp_nat_fn = _est_prior(model, dprior)
    # key tree of every chain: (key, init_pf) = split(key); per step split(key, 4)
    init_pf = np.zeros((n, 2), np.uint32)
    step_keys = np.zeros((M, 3, n, 2), np.uint32)  # prop, pf, accept
    for c in range(n):
        k2 = prandom.split(keys[c], 2)
        k, init_pf[c] = k2[0], k2[1]
        for m in range(M):
            ks = prandom.split(k, 4)
            k = ks[0]
            step_keys[m, :, c] = ks[1:]
    if M > 0:
        z = prandom.normal(step_keys[:, 0].reshape(-1, 2), d).cpu().numpy().reshape(M, n, d)
        u = prandom.uniform(step_keys[:, 2].reshape(-1, 2), 1).cpu().numpy().reshape(M, n)
    ser = struct.device_series()
    md = model.desc()
    ws, nws = _workspace(_native.MP_ALGO_PF, md, ser, n, J, _native.MP_FORM_BATCHED, dev)
    th_dev = torch.empty((n, d), dtype=torch.float32, device=dev)
    pf_keys = torch.from_numpy(np.concatenate([init_pf[None], step_keys[:, 1]]).view(np.int32)).to(dev)
    ll_dev = torch.empty(n, dtype=torch.float64, device=dev)
    thresh = float(max(0.0, thresh))

    def loglik(theta_est, m):
        th_dev.copy_(torch.from_numpy(np.ascontiguousarray(theta_est, dtype=F32)))
        _native.check(_native.lib().mp_pfilter(
            md, ser.desc, _native.ptr(th_dev), d, _native.ptr(pf_keys[m]), None, n, J, thresh,
            _native.MP_SHOULD_TRANS, _native.MP_FORM_BATCHED, _native.ptr(ll_dev), None, None, None, None,
            _native.ptr(ws), nws, _native.stream_ptr()))
        return ll_dev.cpu().numpy().astype(F32)

    cur = model.to_est(th_nat0).astype(F32)
    ll_cur = loglik(cur, 0)
    lp_cur = np.array([lp_est_fn(cur[c]) for c in range(n)], dtype=F32)
    states = [proposal.init_state(cur[c]) for c in range(n)]
    accepts = np.zeros(n, np.int32)
    ll_tr = np.empty((n, M + 1), F32)
    est_tr = np.empty((n, M + 1, d), F32)
    ll_tr[:, 0], est_tr[:, 0] = ll_cur, cur
    for m in range(M):
        prop = np.empty_like(cur)
        for c in range(n):
            prop[c], states[c] = proposal.step(states[c], cur[c], z[m, c], m + 1, int(accepts[c]))
        lp_prop = np.array([lp_est_fn(prop[c]) for c in range(n)], dtype=F32)
        ll_prop = loglik(prop, m + 1)
        with np.errstate(invalid="ignore", divide="ignore"):
            log_alpha = (((ll_prop + lp_prop).astype(F32) - ll_cur).astype(F32) - lp_cur).astype(F32)
            acc = np.isfinite(log_alpha) & np.isfinite(lp_prop) & (np.log(u[m]) < log_alpha)
        cur = np.where(acc[:, None], prop, cur)
        ll_cur = np.where(acc, ll_prop, ll_cur)
        lp_cur = np.where(acc, lp_prop, lp_cur)
        accepts += acc.astype(np.int32)
        ll_tr[:, m + 1], est_tr[:, m + 1] = ll_cur, cur
    nat_tr = model.from_est(est_tr.reshape(-1, d)).reshape(n, M + 1, d)
    lp_tr = np.array([[lp_nat_fn(nat_tr[c, m]) for m in range(M + 1)] for c in range(n)], dtype=F32)
    return ll_tr, lp_tr, nat_tr, accepts


def mop(struct, thetas_array, J, alpha, keys, form="batched", return_float64=False, unit=None):
    """MOP(alpha) objective: thetas (n, d) on the estimation scale, keys (n, 2),
    unit (n,) the panel unit each row filters (None: a one-unit struct).
    Returns the NEGATIVE log-likelihood estimates (n,), as the reference."""
    import torch
    dev = _native.require_device()
    thetas = np.asarray(thetas_array, dtype=F32)
    if thetas.ndim == 1:
        thetas = thetas[None]
    keys = np.asarray(keys, dtype=np.uint32).reshape(-1, 2)
    R = keys.shape[0]
    if thetas.shape[0] != R:
        raise ValueError("thetas and keys disagree on the number of rows")
    ser = struct.device_series()
    md = struct.model.desc()
    th = torch.from_numpy(np.ascontiguousarray(thetas)).to(dev)
    kt = _keys_tensor(keys, dev)
    ll = torch.empty(R, dtype=torch.float64, device=dev)
    fid = _form_id(form)
    ws, nws = _workspace(_native.MP_ALGO_MOP, md, ser, 1 if fid == _native.MP_FORM_PLAIN else R, J,
                         _native.MP_FORM_BATCHED, dev)
    ut = _unit_tensor(unit, R, dev, len(struct.units))
    _native.check(_native.lib().mp_mop(md, ser.desc, _native.ptr(th), struct.model.d, _native.ptr(kt),
                                       None if ut is None else _native.ptr(ut), R, J, float(alpha), fid,
                                       _native.ptr(ll), _native.ptr(ws), nws, _native.stream_ptr()))
    torch.cuda.current_stream().synchronize()
    v = -ll.cpu().numpy()
    return v if return_float64 else v.astype(F32)


class _PerRunner:
    """IF2 iterations on device-resident swarms: one ``mp_perfilter`` launch per
    iteration over R swarms (replicates x units), the (R, J, d) parameter
    swarm staying in HBM between iterations."""

    def __init__(self, struct, R, J, rw, thresh=0.0, unit=None, ancestry=False):
        import torch
        self.dev = dev = _native.require_device()
        self.ser = struct.device_series()
        self.md = struct.model.desc()
        self.R, self.J, self.d = int(R), int(J), struct.model.d
        t = lambda a: torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=F32))).to(dev)  # noqa: E731
        self.sig, self.sig0 = t(rw.sigmas_array), t(rw.sigmas_init_array)
        self.factor = float(rw.factor)
        self.rw = rw
        self.thresh = float(max(0.0, thresh))
        self.unit = None if unit is None else torch.from_numpy(np.ascontiguousarray(unit, dtype=np.int32)).to(dev)
        self.ws, self.nws = _workspace(_native.MP_ALGO_PER, self.md, self.ser, self.R, self.J,
                                       _native.MP_FORM_BATCHED, dev)
        self.anc = torch.empty((self.R, self.J), dtype=torch.int32, device=dev) if ancestry else None

    def cooling(self, ms):
        """Device tables (len(ms), sum T_u) of the schedule's factors for the
        iterations ms, or None for the geometric schedule."""
        import torch
        if self.rw.geometric:
            return None
        tab = np.stack([self.rw.cooling_table(m, self.ser.T_units) for m in ms])
        return torch.from_numpy(np.ascontiguousarray(tab)).to(self.dev)

    def run(self, tin, tout, keys, m, tmean, ll, cool=None):
        """One iteration: tin -> tout (R, J, d), keys (R, 2) int32 view, means
        (R, d), loglik (R); ``cool`` this iteration's factor table (device) for
        a non-geometric schedule (made here when not given)."""
        if cool is None and not self.rw.geometric:
            cool = self.cooling([m])[0]
        _native.check(_native.lib().mp_perfilter(
            self.md, self.ser.desc, _native.ptr(tin), _native.ptr(keys), _native.ptr(self.unit), self.R, self.J,
            self.thresh, _native.ptr(self.sig), _native.ptr(self.sig0), self.factor, _native.ptr(cool), int(m),
            _native.ptr(tout), _native.ptr(tmean), _native.ptr(ll), _native.ptr(self.anc), _native.ptr(self.ws),
            self.nws, _native.stream_ptr()))

    def iterate(self, theta0, mkeys, graph=False):
        """M iterations from theta0 (R, J, d) est. scale (device tensor), keys
        mkeys (M, R, 2).  Returns (final swarm, means (M, R, d), logliks (M, R)),
        device tensors.  graph=True records the M launches as one HIP graph
        and replays it once."""
        import torch
        M = int(mkeys.shape[0])
        kt = torch.from_numpy(np.ascontiguousarray(np.asarray(mkeys, dtype=np.uint32)).view(np.int32)).to(self.dev)
        bufs = [theta0.contiguous(), torch.empty_like(theta0)]
        means = torch.empty((M, self.R, self.d), dtype=torch.float32, device=self.dev)
        lls = torch.empty((M, self.R), dtype=torch.float64, device=self.dev)
        cools = self.cooling(range(M))

        def body():
            for m in range(M):
                self.run(bufs[m % 2], bufs[(m + 1) % 2], kt[m], m, means[m], lls[m],
                         None if cools is None else cools[m])

        if graph and M > 0:
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                body()
            # the graph holds raw addresses: keep every buffer it reads or
            # writes alive for as long as the graph may be replayed
            self.graph, self.graph_buffers = g, (kt, bufs, means, lls, cools)
            g.replay()
        else:
            body()
        return bufs[M % 2], means, lls


def perfilter(struct, thetas_est, keys, J, rw_sd, m, thresh=0.0, unit=None, ancestry=False):
    """One IF2 iteration for R swarms: thetas_est (R, J, d), keys (R, 2).
    Returns (thetas_out (R, J, d), theta_mean (R, d), loglik (R,)), all est. scale,
    plus the final ancestry (R, J) when ``ancestry`` (mif.py:237-245)."""
    import torch
    thetas = np.ascontiguousarray(np.asarray(thetas_est, dtype=F32))
    R = thetas.shape[0]
    run = _PerRunner(struct, R, J, rw_sd, thresh, unit, ancestry)
    tin = torch.from_numpy(thetas).to(run.dev)
    tout = torch.empty_like(tin)
    tmean = torch.empty((R, run.d), dtype=torch.float32, device=run.dev)
    ll = torch.empty(R, dtype=torch.float64, device=run.dev)
    run.run(tin, tout, _keys_tensor(keys, run.dev), m, tmean, ll)
    torch.cuda.current_stream().synchronize()
    out = (tout.cpu().numpy(), tmean.cpu().numpy(), ll.cpu().numpy())
    return out + (run.anc.cpu().numpy(),) if ancestry else out


def mif(struct, thetas_array, rw_sd, M, J, keys, thresh=0.0, n_monitors=0, unit=None, graph=False):
    """IF2 (functional/mif.py:15-122; core/algorithms/mif.py:16-85): thetas (n, J, d)
    natural scale, keys (n, 2).  Returns (logliks (n, M), traces (n, M+1, d)
    natural, final swarm (n, J, d) natural).  The swarms stay on the device for
    all M iterations; ``graph`` records them as one HIP graph (n_monitors = 0)."""
    import torch
    model = struct.model
    thetas = np.asarray(thetas_array, dtype=F32)
    keys = np.asarray(keys, dtype=np.uint32).reshape(-1, 2)
    n = keys.shape[0]
    if thetas.shape[0] != n:
        raise ValueError("thetas and keys disagree on the number of rows")
    rw = rw_sd.canonicalize(model.param_names)
    theta = model.to_est(thetas)
    mkeys = np.stack([prandom.split(k, M + 1)[1:] for k in keys], axis=1)  # (M, n, 2)
    run = _PerRunner(struct, n, J, rw, thresh, unit)
    t0 = torch.from_numpy(np.ascontiguousarray(theta)).to(run.dev)
    if n_monitors >= 1:
        lls = np.zeros((n, M), dtype=F32)
        means = np.zeros((M, n, model.d), dtype=F32)
        cur, nxt = t0, torch.empty_like(t0)
        tm = torch.empty((n, model.d), dtype=torch.float32, device=run.dev)
        llt = torch.empty(n, dtype=torch.float64, device=run.dev)
        kt = _keys_tensor(mkeys.reshape(-1, 2), run.dev).view(M, n, 2)
        for m in range(M):
            mean_now = cur.cpu().numpy().mean(axis=1, dtype=F32)
            mon = np.stack([prandom.split(k, n_monitors) for k in mkeys[m]])
            res = pfilter(struct, mean_now, J, mon, thresh=thresh, should_trans=True)
            lls[:, m] = res["logLik"].mean(axis=1, dtype=F32)
            run.run(cur, nxt, kt[m], m, tm, llt)
            means[m] = tm.cpu().numpy()
            cur, nxt = nxt, cur
        final = cur
    else:
        final, mt, lt = run.iterate(t0, mkeys, graph=graph)
        torch.cuda.current_stream().synchronize()
        means = mt.cpu().numpy()
        lls = lt.cpu().numpy().T.astype(F32)
    traces = np.concatenate([theta.mean(axis=1, dtype=F32)[:, None], np.transpose(means, (1, 0, 2))], axis=1)
    return lls, model.from_est(traces), model.from_est(final.cpu().numpy())


@dataclass
class PanelStruct:
    """A panel of units sharing one native model (functional/structs.py,
    PanelPompStruct): per-unit series, shared and unit-specific parameter names
    (the panel order is shared names, then unit-specific names).  Units may have
    different lengths."""
    model: object
    units: list
    shared_param_names: list
    unit_param_names: list
    unit_names: list = None
    # panel-level estimation-scale transforms (PanelPompStruct.par_trans,
    # panel/panel.py:146); None: the model's own.  The filter kernels always
    # apply the model's transforms (the reference binds them into the unit
    # mechanics when the Pomp is built, core/pomp.py:271-313).
    par_trans: object = None
    _dev: dict = field(default_factory=dict, repr=False)

    def __post_init__(self):
        names = list(self.shared_param_names) + list(self.unit_param_names)
        if sorted(names) != sorted(self.model.param_names) or len(set(names)) != len(names):
            raise ValueError("shared + unit-specific parameter names must be the model's parameter names")
        if self.unit_names is None:
            self.unit_names = [f"unit{u + 1}" for u in range(len(self.units))]

    @property
    def U(self):
        return len(self.units)

    @property
    def param_names(self):
        return list(self.shared_param_names) + list(self.unit_param_names)

    @property
    def perm(self):
        """Model-order index into the panel order (unit_param_permutations)."""
        names = self.param_names
        return np.array([names.index(n) for n in self.model.param_names], dtype=np.int64)

    def device_series(self):
        return PompStruct.device_series(self)


def panel_pfilter(struct, thetas_panel, J, keys, thresh=0.0, chunk_size=1, CLL=False, ESS=False,
                  filter_mean=False, prediction_mean=False):
    """Panel bootstrap filter (functional/pfilter.py panel_pfilter; pfilter.py:248-304):
    thetas_panel (n, U, n_shared + n_spec) natural scale in the panel order,
    keys (n, U, 2).  Every (set, unit) filter runs in one launch (``chunk_size``
    is accepted for signature compatibility).  Returns ``logLik`` (n, U) and the
    requested diagnostics (n, U, T_max, ...)."""
    th = np.asarray(thetas_panel, dtype=F32)
    n, U = th.shape[0], th.shape[1]
    if U != struct.U:
        raise ValueError("thetas_panel has a different number of units than the struct")
    keys = np.asarray(keys, dtype=np.uint32).reshape(n, U, 2)
    flat = np.ascontiguousarray(th[..., struct.perm].reshape(n * U, -1))
    unit = np.tile(np.arange(U, dtype=np.int32), n)
    out = pfilter(struct, flat, J, keys.reshape(n * U, 1, 2), thresh=thresh, CLL=CLL, ESS=ESS,
                  filter_mean=filter_mean, prediction_mean=prediction_mean, unit=unit)
    res = {}
    for k, v in out.items():
        res[k] = v.reshape((n, U) + v.shape[2:])
    return res


def panel_mif(struct, shared_array, unit_array, rw_sd, M, J, keys, thresh=0.0, n_monitors=0, block=True,
              graph=False):
    """Panel IF2 (functional/mif.py:125-294; core/algorithms/panel_mif.py:17-260).
    shared_array (n, J, n_shared) and unit_array (n, J, U, n_spec) on the natural
    scale, keys (n, 2).  Returns (shared_traces (n, M+1, n_shared+1),
    unit_traces (n, M+1, U, n_spec+1), final shared swarm, final unit swarm),
    natural scale; column 0 of the traces is the log-likelihood (NaN at m = 0).

    Per iteration the unit keys are split(iter_key, U) (panel_mif.py:113) and
    unit u runs _perfilter_internal on [shared, unit u] in the model's order.
    With no shared parameters and block=True the units are independent: every
    (replicate, unit) swarm of an iteration runs in one launch, and ``graph``
    records the M launches as one HIP graph.  Otherwise units run one launch
    after another, carrying the shared swarm (and, for block=False, reordering
    the other units' swarms by the unit's ancestry, panel_mif.py:224-231)."""
    import torch
    model = struct.model
    dev = _native.require_device()
    U, ns, nu = struct.U, len(struct.shared_param_names), len(struct.unit_param_names)
    sh = np.asarray(shared_array, dtype=F32)
    un = np.asarray(unit_array, dtype=F32)
    keys = np.asarray(keys, dtype=np.uint32).reshape(-1, 2)
    n = keys.shape[0]
    if sh.shape[:2] != (n, J) or sh.shape[2] != ns or un.shape != (n, J, U, nu):
        raise ValueError("shared_array must be (n, J, n_shared) and unit_array (n, J, U, n_spec)")
    if struct.model.d != ns + nu:
        raise ValueError("parameter count mismatch")
    perm = struct.perm
    inv = np.argsort(perm)
    pt = model if struct.par_trans is None else struct.par_trans
    rw = rw_sd.canonicalize(model.param_names)  # model order
    # natural -> estimation scale (panel order)
    full = np.concatenate([np.broadcast_to(sh[:, :, None, :], (n, J, U, ns)), un], axis=3)
    full_est = pt.to_est(full[..., perm])[..., inv]
    sh_est = full_est[:, :, 0, :ns].copy() if ns else np.zeros((n, J, 0), F32)
    un_est = full_est[:, :, :, ns:].copy()
    iter_keys = np.stack([prandom.split(k, M + 1)[1:] for k in keys], axis=1)            # (M, n, 2)
    unit_keys = np.stack([[prandom.split(k, U) for k in row] for row in iter_keys])     # (M, n, U, 2)
    sh_tr = np.full((n, M + 1, ns + 1), np.nan, dtype=F32)
    un_tr = np.full((n, M + 1, U, nu + 1), np.nan, dtype=F32)
    sh_tr[:, 0, 1:] = sh_est.mean(axis=1, dtype=F32)
    un_tr[:, 0, :, 1:] = un_est.mean(axis=1, dtype=F32)

    if ns == 0 and block and n_monitors == 0:
        # independent units: filter r * U + u, one launch per iteration
        unit = np.tile(np.arange(U, dtype=np.int32), n)
        th0 = np.ascontiguousarray(np.transpose(un_est, (0, 2, 1, 3))[..., perm].reshape(n * U, J, -1))
        run = _PerRunner(struct, n * U, J, rw, thresh, unit)
        final, mt, lt = run.iterate(torch.from_numpy(th0).to(dev), unit_keys.reshape(M, n * U, 2), graph=graph)
        torch.cuda.current_stream().synchronize()
        means = mt.cpu().numpy()[..., inv].reshape(M, n, U, nu)
        lls = lt.cpu().numpy().reshape(M, n, U).astype(F32)
        un_tr[:, 1:, :, 0] = np.transpose(lls, (1, 0, 2))
        un_tr[:, 1:, :, 1:] = np.transpose(means, (1, 0, 2, 3))
        sh_tr[:, 1:, 0] = lls.sum(axis=2, dtype=F32).T
        un_est = np.transpose(final.cpu().numpy()[..., inv].reshape(n, U, J, nu), (0, 2, 1, 3))
    else:
        perm_t = torch.from_numpy(perm).to(dev)
        inv_t = torch.from_numpy(inv).to(dev)
        sh_t = torch.from_numpy(np.ascontiguousarray(sh_est)).to(dev)
        un_t = torch.from_numpy(np.ascontiguousarray(un_est)).to(dev)
        units = [_PerRunner(struct, n, J, rw, thresh, np.full(n, u, np.int32), ancestry=not block)
                 for u in range(U)]
        tm = torch.empty((n, model.d), dtype=torch.float32, device=dev)
        llt = torch.empty(n, dtype=torch.float64, device=dev)
        for m in range(M):
            total = np.zeros(n, dtype=F32)
            for u in range(U):
                th = torch.cat([sh_t, un_t[:, :, u, :]], dim=2).index_select(2, perm_t).contiguous()
                kt = _keys_tensor(unit_keys[m, :, u], dev)
                if n_monitors >= 1:
                    mean_now = th.mean(dim=1).cpu().numpy()
                    mon = np.stack([prandom.split(k, n_monitors) for k in unit_keys[m, :, u]])
                    sub = PompStruct(model=model, units=[struct.units[u]])
                    res = pfilter(sub, mean_now, J, mon, thresh=thresh, should_trans=True)
                    ll_u = res["logLik"].mean(axis=1, dtype=F32)
                out = torch.empty_like(th)
                units[u].run(th, out, kt, m, tm, llt)
                upd = out.index_select(2, inv_t)
                if n_monitors < 1:
                    ll_u = llt.cpu().numpy().astype(F32)
                if ns:
                    sh_t = upd[:, :, :ns].contiguous()
                if not block:
                    anc = units[u].anc.long()
                    un_t = torch.stack([un_t[r].index_select(0, anc[r]) for r in range(n)])
                un_t = un_t.clone()
                un_t[:, :, u, :] = upd[:, :, ns:]
                un_tr[:, m + 1, u, 0] = ll_u
                un_tr[:, m + 1, u, 1:] = upd[:, :, ns:].mean(dim=1).cpu().numpy()
                total = (total + ll_u).astype(F32)
            sh_tr[:, m + 1, 0] = total
            if ns:
                sh_tr[:, m + 1, 1:] = sh_t.mean(dim=1).cpu().numpy()
        torch.cuda.current_stream().synchronize()
        sh_est = sh_t.cpu().numpy()
        un_est = un_t.cpu().numpy()

    def nat(shared_part, unit_part):
        """est -> natural for arrays (..., ns) and (..., U, nu) in panel order."""
        lead = unit_part.shape[:-2]
        fullp = np.concatenate([np.broadcast_to(shared_part[..., None, :], lead + (U, ns)), unit_part], axis=-1)
        natp = pt.from_est(fullp[..., perm])[..., inv]
        return natp[..., 0, :ns], natp[..., ns:]

    s_nat, u_nat = nat(sh_tr[..., 1:], un_tr[..., 1:])
    sh_tr[..., 1:] = s_nat
    un_tr[..., 1:] = u_nat
    fs, fu = nat(sh_est, un_est)
    return sh_tr, un_tr, np.ascontiguousarray(fs), np.ascontiguousarray(fu)


def mop_grad(struct, thetas_array, J, alpha, keys, K=50, form="batched", unit=None):
    """MOP(alpha) value and gradient by the native adjoint.  thetas (n, d) on the
    estimation scale, keys (n, 2), unit (n,) the panel unit of each row (None:
    a one-unit struct).  Returns (loglik (n,) float64, grad (n, d) float32)
    with grad = d loglik / d theta_est.  The reference's
    ``_mop_internal_mean`` is -loglik / T (core/algorithms/mop.py:148-158)."""
    import torch
    dev = _native.require_device()
    thetas = np.ascontiguousarray(np.asarray(thetas_array, dtype=F32))
    if thetas.ndim == 1:
        thetas = thetas[None]
    keys = np.asarray(keys, dtype=np.uint32).reshape(-1, 2)
    R = keys.shape[0]
    if thetas.shape[0] != R:
        raise ValueError("thetas and keys disagree on the number of rows")
    ll, g = _mop_grad_dev(struct, torch.from_numpy(thetas).to(dev), J, alpha, _keys_tensor(keys, dev), K, form,
                          _unit_tensor(unit, R, dev, len(struct.units)))
    torch.cuda.current_stream().synchronize()
    return ll.cpu().numpy(), g.cpu().numpy()


def _mop_grad_dev(struct, th, J, alpha, kt, K=50, form="batched", unit=None):
    """mp_mop_grad on device tensors: th (R, d) float32 estimation scale, kt
    (R, 2) int32 view of the keys, unit (R,) int32 device tensor of panel units
    or None (a one-unit struct); returns device (loglik (R,) float64, grad
    (R, d) float32), enqueued on the current stream without synchronising."""
    import torch
    for u in struct.units:
        if not np.all(u.nstep == 1):
            raise ValueError("the MOP adjoint needs exactly one Euler step per observation interval")
    dev = th.device
    R = th.shape[0]
    ser = struct.device_series()
    md = struct.model.desc()
    ll = torch.empty(R, dtype=torch.float64, device=dev)
    g = torch.empty((R, struct.model.d), dtype=torch.float32, device=dev)
    fid = _form_id(form)
    n = _native.lib().mp_mop_grad_workspace_size(md, ser.desc, 1 if fid == _native.MP_FORM_PLAIN else R, J, K)
    ws = torch.empty(int(n), dtype=torch.uint8, device=dev)
    _native.check(_native.lib().mp_mop_grad(md, ser.desc, _native.ptr(th), struct.model.d, _native.ptr(kt),
                                            None if unit is None else _native.ptr(unit), R, J, float(alpha),
                                            int(K), fid, _native.ptr(ll), _native.ptr(g), _native.ptr(ws), int(n),
                                            _native.stream_ptr()))
    return ll, g


class _HessBuffers:
    """Device buffers of repeated mp_mop_hess calls on n rows (the train loop)."""

    def __init__(self, struct, n, J, dev):
        import torch
        self.ser = struct.device_series()
        self.md = struct.model.desc()
        self.n, self.J, self.d = n, J, struct.model.d
        self.ll = torch.empty(n, dtype=torch.float64, device=dev)
        self.g = torch.empty((n, self.d), dtype=torch.float32, device=dev)
        self.H = torch.empty((n, self.d, self.d), dtype=torch.float32, device=dev)
        self.nws = int(_native.lib().mp_mop_hess_workspace_size(self.md, self.ser.desc, n, J))
        self.ws = torch.empty(self.nws, dtype=torch.uint8, device=dev)

    def run(self, th, keys, alpha):
        """Hessian (n, d, d) of the log-likelihood at th (n, d) est. scale, keys (n, 2) int32 view."""
        _native.check(_native.lib().mp_mop_hess(
            self.md, self.ser.desc, _native.ptr(th), self.d, _native.ptr(keys), self.n, self.J, float(alpha),
            _native.ptr(self.ll), _native.ptr(self.g), _native.ptr(self.H), _native.ptr(self.ws), self.nws,
            _native.stream_ptr()))
        return self.H


def mop_hess(struct, thetas_array, J, alpha, keys):
    """MOP(alpha) value, gradient and Hessian with respect to the
    estimation-scale parameters (the native forward-mode kernel).  thetas (n, d)
    on the estimation scale, keys (n, 2).  Returns (loglik (n,) float64, grad
    (n, d), hess (n, d, d)) of the log-likelihood; the reference's
    ``_jhess_mop`` is the Hessian of -loglik / T (core/algorithms/ad_helpers.py:56-75)."""
    import torch
    dev = _native.require_device()
    thetas = np.ascontiguousarray(np.asarray(thetas_array, dtype=F32))
    if thetas.ndim == 1:
        thetas = thetas[None]
    keys = np.asarray(keys, dtype=np.uint32).reshape(-1, 2)
    R, d = keys.shape[0], struct.model.d
    if thetas.shape[0] != R:
        raise ValueError("thetas and keys disagree on the number of rows")
    buf = _HessBuffers(struct, R, J, dev)
    buf.run(torch.from_numpy(thetas).to(dev), _keys_tensor(keys, dev), alpha)
    torch.cuda.current_stream().synchronize()
    return buf.ll.cpu().numpy(), buf.g.cpu().numpy(), buf.H.cpu().numpy()


def _cosine_cooling(i, M, c):
    F = np.float32
    return F(F(c) + (F(1) - F(c)) * F(0.5) * (F(1) + np.cos(F(np.pi) * F(i) / F(M), dtype=F)))


def _armijo_step(struct, th, grad, direction, curr, J, keys, thresh, optimizer, eta0):
    """Per-row step sizes of the reference's Armijo backtracking
    (core/algorithms/train.py:442-499, stoch=False, frac=0.5): the loop of a
    row stops at the first trial point whose bootstrap-filter negative
    log-likelihood is not NaN and at most curr + eta c (grad . direction);
    after max_ls_itn halvings it stops at the current size.  All rows still
    searching share one filter launch per round.  float32 like the
    reference."""
    n = th.shape[0]
    etas = np.full(n, F32(eta0), dtype=F32)
    slope = (grad * direction).sum(dim=1).cpu().numpy().astype(F32)
    active = np.ones(n, dtype=bool)
    itn = 0
    thn = th.cpu().numpy()
    dirn = direction.cpu().numpy()
    c = F32(optimizer.c)
    while active.any():
        rows = np.nonzero(active)[0]
        pts = (thn[rows] + etas[rows, None] * dirn[rows]).astype(F32)
        res = pfilter(struct, pts, J, np.asarray(keys, np.uint32)[rows].reshape(-1, 1, 2), thresh=thresh,
                      should_trans=True)
        nxt = (-res["logLik"][:, 0]).astype(F32)
        cont = (nxt > curr[rows] + (etas[rows] * c) * slope[rows]) | np.isnan(nxt)
        go = cont & (itn < optimizer.max_ls_itn)
        etas[rows[go]] = etas[rows[go]] * F32(0.5)
        active[rows[~go]] = False
        itn += 1
    return etas


def train(struct, thetas_array, J, optimizer, M, eta, alpha, keys, alpha_cooling=1.0, thresh=0.0, n_monitors=1,
          K=50):
    """Gradient ascent on the MOP log-likelihood (functional/train.py:11-113;
    core/algorithms/train.py:38-180): per iteration m, alpha_m = 1 - (1 -
    alpha) c_m with cosine cooling c_m, the MOP value and hand-adjoint gradient
    (one launch over all n parameter sets), the optimizer's direction and
    theta += eta[m] direction.  thetas (n, d) are used as the estimation-scale
    start, as the reference passes them through.  Keys per iteration follow the
    reference: (key, sub) = split(key) for the MOP draw, n_monitors further
    keys when n_monitors > 1, then one more split.  With ``optimizer.ls`` the
    step size comes from the reference's Armijo backtracking
    (core/algorithms/train.py:139-162, 442-499): starting from the mean
    learning rate, halve while the bootstrap filter's negative log-likelihood
    at the trial point, drawn with the iteration's MOP key, exceeds the current
    objective plus eta c (grad . direction) or is NaN, at most max_ls_itn
    times.  The parameters and the optimizer state stay on the device for all M
    iterations (line search and FullMatrixAdam's eigendecompositions read back
    every iteration).  Returns (neg_logliks (n, M+1), NaN first, traces (n,
    M+1, d)), float32."""
    import torch
    from .optim import BFGS, SGD, Adam, FullMatrixAdam, Newton, WeightedNewton
    if not isinstance(optimizer, (Adam, SGD, FullMatrixAdam, BFGS, Newton, WeightedNewton)):
        raise ValueError(f"Optimizer '{type(optimizer).__name__}' not supported")
    newton = isinstance(optimizer, (Newton, WeightedNewton))
    if n_monitors < 1 and optimizer.ls:
        raise ValueError("Line search requires at least one monitor")
    dev = _native.require_device()
    th0 = np.ascontiguousarray(np.asarray(thetas_array, dtype=F32))
    if th0.ndim == 1:
        th0 = th0[None]
    keys = np.asarray(keys, dtype=np.uint32).reshape(-1, 2)
    n, d = th0.shape
    if keys.shape[0] != n:
        raise ValueError("thetas and keys disagree on the number of rows")
    for u in struct.units:
        if not np.all(u.nstep == u.nstep[0]):
            raise ValueError("train needs the same number of Euler steps in every observation interval")
    T = struct.units[0].T
    eta_arr = torch.from_numpy(eta.to_array(struct.param_names, M)).to(dev)
    # host key schedule: MOP keys (M, n, 2), Hessian keys and the running keys
    mop_keys = np.zeros((M, n, 2), np.uint32)
    hess_keys = np.zeros((M, n, 2), np.uint32)
    mon_keys = []
    for r in range(n):
        k = keys[r]
        row_mon = []
        for m in range(M):
            k2 = prandom.split(k, 2)
            k, mop_keys[m, r] = k2[0], k2[1]
            if n_monitors > 1:
                ks = prandom.split(k, n_monitors + 1)
                k, sub = ks[0], ks[1:]
                row_mon.append(sub)
            k2 = prandom.split(k, 2)
            k, hess_keys[m, r] = k2[0], k2[1]  # train.py:132
        mon_keys.append(row_mon)
    ser = struct.device_series()
    md = struct.model.desc()
    fid = _native.MP_FORM_BATCHED
    nws = _native.lib().mp_mop_grad_workspace_size(md, ser.desc, n, J, K)
    ws = torch.empty(int(nws), dtype=torch.uint8, device=dev)
    kt = torch.from_numpy(mop_keys.view(np.int32)).to(dev)
    if newton:
        hkt = torch.from_numpy(hess_keys.view(np.int32)).to(dev)
        hws = _HessBuffers(struct, n, J, dev)
    th = torch.from_numpy(th0).to(dev)
    state = optimizer.init_state(th)
    lls = torch.empty((M, n), dtype=torch.float64, device=dev)
    g = torch.empty((n, d), dtype=torch.float32, device=dev)
    traces = [th.clone()]
    nll = np.full((n, M + 1), np.nan, dtype=F32)
    for m in range(M):
        a_m = float(F32(F32(1) - (F32(1) - F32(alpha)) * _cosine_cooling(m, M, alpha_cooling)))
        _native.check(_native.lib().mp_mop_grad(md, ser.desc, _native.ptr(th), d, _native.ptr(kt[m]), None, n, J,
                                                a_m, int(K), fid, _native.ptr(lls[m]), _native.ptr(g),
                                                _native.ptr(ws), int(nws), _native.stream_ptr()))
        grad = -g / np.float32(T)  # gradient of the mean objective, -loglik / T
        if n_monitors > 1:
            means = th.cpu().numpy()
            res = pfilter(struct, means, J, np.stack([mon_keys[r][m] for r in range(n)]), thresh=thresh,
                          should_trans=True)
            nll[:, m + 1] = -res["logLik"].mean(axis=1, dtype=F32)
        if optimizer.clip_norm is not None:
            grad = torch.clamp(grad, -optimizer.clip_norm, optimizer.clip_norm)
        if newton:
            # Hessian of the mean objective -loglik / T at theta, drawn with the
            # iteration's Hessian subkey (train.py:134-141)
            hess = hws.run(th, hkt[m], a_m) / (-np.float32(T))
            direction, state = optimizer.step(grad, state, m, eta_i=eta_arr[m], hess=hess)
        else:
            direction, state = optimizer.step(grad, state, m, eta_i=eta_arr[m])
        if optimizer.scale:
            direction = direction / torch.linalg.norm(direction, dim=-1, keepdim=True)
        if optimizer.ls:
            curr = (-lls[m].cpu().numpy()).astype(F32) if n_monitors == 1 else nll[:, m + 1]
            step = _armijo_step(struct, th, grad, direction, curr, J, mop_keys[m], thresh, optimizer,
                                float(eta_arr[m].mean()))
            th = th + torch.from_numpy(step).to(dev)[:, None] * direction
        else:
            th = th + eta_arr[m][None, :] * direction
        traces.append(th.clone())
    torch.cuda.current_stream().synchronize()
    if n_monitors == 1:
        nll[:, 1:] = (-lls.cpu().numpy().T).astype(F32)
    return nll, torch.stack(traces, dim=1).cpu().numpy()


def panel_mop(struct, shared_array, unit_array, J, alpha, keys, chunk_size=1, grad=False, K=50):
    """Panel MOP objective (core/algorithms/mop.py:183-275,
    ``_chunked_panel_mop_internal`` and its value-and-gradient): shared
    (n_shared,) and unit (U, n_spec) on the estimation scale, keys (U, 2), one
    per unit.  Every unit's MOP filter runs in one launch; the value is
    sum_u (-loglik_u) / (U T), summed chunk by chunk in float32 as the
    reference's scan does.  With ``grad`` returns (value, d value / d shared
    (n_shared,), d value / d unit (U, n_spec)) by the native adjoint.  Units
    must have one length and ``chunk_size`` must divide U (the reference
    reshapes the units into chunks)."""
    import torch
    U, ns, nu = struct.U, len(struct.shared_param_names), len(struct.unit_param_names)
    sh = np.asarray(shared_array, dtype=F32).reshape(ns)
    un = np.asarray(unit_array, dtype=F32).reshape(U, nu)
    keys = np.asarray(keys, dtype=np.uint32).reshape(U, 2)
    cs = int(chunk_size)
    if cs < 1 or U % cs:
        raise ValueError(f"chunk_size {cs} does not divide the {U} units")
    lengths = {u.T for u in struct.units}
    if len(lengths) != 1:
        raise ValueError("the panel MOP needs units with the same number of observations")
    T = lengths.pop()
    perm = struct.perm
    th = np.ascontiguousarray(np.concatenate([np.broadcast_to(sh[None], (U, ns)), un], axis=1)[:, perm])
    unit = np.arange(U, dtype=np.int32)
    if grad:
        ll, g = mop_grad(struct, th, J, alpha, keys, K=K, unit=unit)
    else:
        ll = -mop(struct, th, J, alpha, keys, return_float64=True, unit=unit)
    res = (-ll).astype(F32).reshape(U // cs, cs)
    total = F32(0)
    for c in range(U // cs):
        total = F32(total + res[c].sum(dtype=F32))
    value = F32(total / F32(U * T))
    if not grad:
        return value
    gp = (-g.astype(np.float64)[:, np.argsort(perm)] / (U * T)).astype(F32)  # d value / d theta, panel order
    return value, gp[:, :ns].sum(axis=0, dtype=F32), np.ascontiguousarray(gp[:, ns:])


def panel_train(struct, shared_array, unit_array, J, optimizer, M, eta, alpha, keys, alpha_cooling=1.0,
                chunk_size=1, K=50):
    """Gradient ascent on the panel MOP objective (functional/train.py:112-242;
    core/algorithms/train.py:190-415).  shared_array (n, n_shared) and
    unit_array (n, U, n_spec) natural scale, keys (n, M, U, 2): the MOP key of
    unit u at iteration m.  Per iteration m (alpha_m by cosine cooling) the
    units are taken ``chunk_size`` at a time, in order; each chunk is one
    adjoint launch over its (replicate, unit) filters at the current shared
    values.  The chunk loss is sum_u (-loglik_u) / (chunk_size T); its shared
    gradient sums over the chunk and the unit gradients are scaled by
    chunk_size; the shared optimizer steps with the running chunk count, the
    unit optimizer with m; shared += eta_shared[m] / n_chunks direction, unit
    += eta_spec[m] direction.  Only SGD, Adam and FullMatrixAdam are taken,
    as the reference.  Returns (neg_logliks (n, M+1), NaN first, the mean over
    chunks of the chunk loss times T U; shared history (n, M+1, n_shared);
    unit history (n, M+1, U, n_spec)), natural scale, float32."""
    import torch
    from .optim import SGD, Adam, FullMatrixAdam
    if not isinstance(optimizer, (SGD, Adam, FullMatrixAdam)):
        raise ValueError(f"Optimizer '{type(optimizer).__name__}' not supported for panel train")
    model = struct.model
    U, ns, nu = struct.U, len(struct.shared_param_names), len(struct.unit_param_names)
    sh = np.asarray(shared_array, dtype=F32)
    un = np.asarray(unit_array, dtype=F32)
    if un.ndim == 2:
        sh, un = sh.reshape(1, -1), un[None]
    n = un.shape[0]
    if sh.shape != (n, ns) or un.shape != (n, U, nu):
        raise ValueError("shared_array must be (n, n_shared) and unit_array (n, U, n_spec)")
    keys = np.asarray(keys, dtype=np.uint32)
    if keys.size != n * M * U * 2:
        raise ValueError(f"keys must hold (n, M, U) = ({n}, {M}, {U}) keys")
    cs = max(1, int(chunk_size))
    if U % cs:
        raise ValueError(f"chunk_size {cs} does not divide the {U} units")
    n_chunks = U // cs
    lengths = {u.T for u in struct.units}
    if len(lengths) != 1:
        raise ValueError("panel train needs units with the same number of observations")
    T = lengths.pop()
    for u in struct.units:
        if not np.all(u.nstep == 1):
            raise ValueError("the MOP adjoint needs exactly one Euler step per observation interval")
    dev = _native.require_device()
    perm = struct.perm
    inv = np.argsort(perm)
    pt = model if struct.par_trans is None else struct.par_trans
    eta_sh = torch.from_numpy(np.asarray(eta.to_array(struct.shared_param_names, M), F32).reshape(M, ns)
                              if ns else np.zeros((M, 0), F32)).to(dev)
    eta_un = torch.from_numpy(np.asarray(eta.to_array(struct.unit_param_names, M), F32).reshape(M, nu)
                              if nu else np.zeros((M, 0), F32)).to(dev)
    full = np.concatenate([np.broadcast_to(sh[:, None, :], (n, U, ns)), un], axis=2)
    full_est = pt.to_est(full[..., perm])[..., inv]
    sh_t = torch.from_numpy(np.ascontiguousarray(full_est[:, 0, :ns])).to(dev)
    un_t = torch.from_numpy(np.ascontiguousarray(full_est[:, :, ns:])).to(dev)

    # filter rows of chunk c: (replicate r, unit c cs + j) at row r cs + j
    R = n * cs
    kc = np.ascontiguousarray(keys.reshape(n, M, n_chunks, cs, 2).transpose(1, 2, 0, 3, 4).reshape(
        M, n_chunks, R, 2))
    kt = torch.from_numpy(kc.view(np.int32)).to(dev)
    rows = [torch.from_numpy(np.tile(np.arange(c * cs, (c + 1) * cs, dtype=np.int32), n)).to(dev)
            for c in range(n_chunks)]
    perm_t = torch.from_numpy(perm).to(dev)
    inv_t = torch.from_numpy(inv).to(dev)
    ser = struct.device_series()
    md = model.desc()
    d = model.d
    nws = int(_native.lib().mp_mop_grad_workspace_size(md, ser.desc, R, J, K))
    ws = torch.empty(nws, dtype=torch.uint8, device=dev)
    lls = torch.empty((M, n_chunks, R), dtype=torch.float64, device=dev)
    g = torch.empty((R, d), dtype=torch.float32, device=dev)
    st_sh = optimizer.init_state(sh_t) if ns else None
    st_un = [optimizer.init_state(un_t[:, c * cs:(c + 1) * cs].reshape(R, nu)) if nu else None
             for c in range(n_chunks)]
    scale = float(F32(cs * T))
    sh_hist, un_hist = [sh_t.clone()], [un_t.clone()]
    step = 0
    for m in range(M):
        a_m = float(F32(F32(1) - (F32(1) - F32(alpha)) * _cosine_cooling(m, M, alpha_cooling)))
        for c in range(n_chunks):
            th = torch.cat([sh_t.repeat_interleave(cs, dim=0), un_t[:, c * cs:(c + 1) * cs].reshape(R, nu)],
                           dim=1).index_select(1, perm_t).contiguous()
            _native.check(_native.lib().mp_mop_grad(md, ser.desc, _native.ptr(th), d, _native.ptr(kt[m, c]),
                                                    _native.ptr(rows[c]), R, J, a_m, int(K), _native.MP_FORM_BATCHED,
                                                    _native.ptr(lls[m, c]), _native.ptr(g), _native.ptr(ws), nws,
                                                    _native.stream_ptr()))
            gr = -g.index_select(1, inv_t) / scale  # d chunk loss / d theta, panel order
            g_sh = gr[:, :ns].reshape(n, cs, ns).sum(dim=1)
            g_un = gr[:, ns:] * float(cs)
            if optimizer.clip_norm is not None:
                g_sh = torch.clamp(g_sh, -optimizer.clip_norm, optimizer.clip_norm)
                g_un = torch.clamp(g_un, -optimizer.clip_norm, optimizer.clip_norm)
            if ns:
                d_sh, st_sh = optimizer.step(g_sh, st_sh, step)
                if optimizer.scale:
                    d_sh = d_sh / torch.clamp(torch.linalg.norm(d_sh, dim=-1, keepdim=True), min=1e-8)
                sh_t = sh_t + (eta_sh[m] / float(n_chunks))[None, :] * d_sh
            if nu:
                d_un, st_un[c] = optimizer.step(g_un, st_un[c], m)
                if optimizer.scale:
                    d_un = d_un / torch.clamp(torch.linalg.norm(d_un, dim=-1, keepdim=True), min=1e-8)
                un_t = un_t.clone()
                un_t[:, c * cs:(c + 1) * cs] += (eta_un[m][None, :] * d_un).reshape(n, cs, nu)
            step += 1
        sh_hist.append(sh_t.clone())
        un_hist.append(un_t.clone())
    torch.cuda.current_stream().synchronize()
    # objective: per chunk sum_u (-loglik_u) / (cs T) times T U, float32, mean over chunks
    res = (-lls.cpu().numpy()).astype(F32).reshape(M, n_chunks, n, cs)
    chunk = (res.sum(axis=3, dtype=F32) / F32(cs * T) * F32(T * U)).astype(F32)
    nll = np.full((n, M + 1), np.nan, dtype=F32)
    nll[:, 1:] = chunk.mean(axis=1, dtype=F32).T
    sh_tr = torch.stack(sh_hist, dim=1).cpu().numpy()
    un_tr = torch.stack(un_hist, dim=1).cpu().numpy()
    fullp = np.concatenate([np.broadcast_to(sh_tr[:, :, None, :], (n, M + 1, U, ns)), un_tr], axis=-1)
    natp = pt.from_est(fullp[..., perm])[..., inv]
    return nll, np.ascontiguousarray(natp[:, :, 0, :ns]), np.ascontiguousarray(natp[..., ns:])


class _MopFn:
    """Builds the torch.autograd.Function lazily (torch is only needed on use)."""
    fn = None

    @classmethod
    def get(cls):
        if cls.fn is None:
            import torch

            class MopFunction(torch.autograd.Function):
                @staticmethod
                def forward(ctx, theta_est, struct, J, alpha, keys, K):
                    # device-resident: theta and keys stay on the GPU and nothing
                    # synchronises (keys given as a device int32/uint32 (n, 2)
                    # tensor are used in place; NumPy keys are copied up once)
                    dev = _native.require_device()
                    th = theta_est.detach().to(device=dev, dtype=torch.float32).contiguous()
                    if isinstance(keys, torch.Tensor):
                        kt = keys.to(device=dev).contiguous().view(torch.int32).reshape(-1, 2)
                    else:
                        kt = _keys_tensor(keys, dev)
                    ll, g = _mop_grad_dev(struct, th, J, alpha, kt, K)
                    ctx.save_for_backward(g)
                    return ll.to(theta_est.device)

                @staticmethod
                def backward(ctx, grad_out):
                    (g,) = ctx.saved_tensors
                    gb = grad_out.to(device=g.device, dtype=g.dtype)[:, None] * g
                    return (gb.to(grad_out.device), None, None, None, None, None)

            cls.fn = MopFunction
        return cls.fn


def mop_loglik_autograd(theta_est, struct, J, alpha, keys, K=50):
    """MOP log-likelihood as a differentiable torch function of theta_est (n, d);
    its backward is the native adjoint (the forward runs value + gradient)."""
    return _MopFn.get().apply(theta_est, struct, J, alpha, keys, K)
