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


def _workspace(algo, md, ser, R, J, form, dev):
    import torch
    n = _native.lib().mp_filter_workspace_size(algo, md, ser.desc, R, J, form)
    if n == 0:
        raise _native.NativeError("workspace size query failed")
    return torch.empty(int(n), dtype=torch.uint8, device=dev), int(n)


def pfilter(struct, thetas_array, J, keys, thresh=0.0, CLL=False, ESS=False, filter_mean=False,
            prediction_mean=False, form="batched", should_trans=False):
    """Bootstrap particle filter.  thetas (n, d) natural scale, keys (n, reps, 2).
    Returns a dict with ``logLik`` (n, reps) and the requested diagnostics."""
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
    ser = struct.device_series()
    md = struct.model.desc()
    th = torch.from_numpy(np.repeat(thetas, reps, axis=0)).to(dev)
    kt = _keys_tensor(keys.reshape(R, 2), dev)
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
        md, ser.desc, _native.ptr(th), struct.model.d, _native.ptr(kt), None, R, J, float(thresh), flags,
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


def mop(struct, thetas_array, J, alpha, keys, form="batched", return_float64=False):
    """MOP(alpha) objective: thetas (n, d) on the estimation scale, keys (n, 2).
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
    _native.check(_native.lib().mp_mop(md, ser.desc, _native.ptr(th), struct.model.d, _native.ptr(kt), None,
                                       R, J, float(alpha), fid, _native.ptr(ll), _native.ptr(ws), nws,
                                       _native.stream_ptr()))
    torch.cuda.current_stream().synchronize()
    v = -ll.cpu().numpy()
    return v if return_float64 else v.astype(F32)


def perfilter(struct, thetas_est, keys, J, rw_sd, m, thresh=0.0, unit=None):
    """One IF2 iteration for R swarms: thetas_est (R, J, d), keys (R, 2).
    Returns (thetas_out (R, J, d), theta_mean (R, d), loglik (R,)), all est. scale."""
    import torch
    dev = _native.require_device()
    thetas = np.ascontiguousarray(np.asarray(thetas_est, dtype=F32))
    R = thetas.shape[0]
    ser = struct.device_series()
    md = struct.model.desc()
    tin = torch.from_numpy(thetas).to(dev)
    tout = torch.empty_like(tin)
    tmean = torch.empty((R, struct.model.d), dtype=torch.float32, device=dev)
    ll = torch.empty(R, dtype=torch.float64, device=dev)
    kt = _keys_tensor(keys, dev)
    sig = torch.from_numpy(np.ascontiguousarray(rw_sd.sigmas_array)).to(dev)
    sig0 = torch.from_numpy(np.ascontiguousarray(rw_sd.sigmas_init_array)).to(dev)
    ut = None if unit is None else torch.from_numpy(np.asarray(unit, dtype=np.int32)).to(dev)
    ws, nws = _workspace(_native.MP_ALGO_PER, md, ser, R, J, _native.MP_FORM_BATCHED, dev)
    _native.check(_native.lib().mp_perfilter(
        md, ser.desc, _native.ptr(tin), _native.ptr(kt), _native.ptr(ut), R, J, float(max(0.0, thresh)),
        _native.ptr(sig), _native.ptr(sig0), float(rw_sd.factor), int(m), _native.ptr(tout),
        _native.ptr(tmean), _native.ptr(ll), _native.ptr(ws), nws, _native.stream_ptr()))
    torch.cuda.current_stream().synchronize()
    return tout.cpu().numpy(), tmean.cpu().numpy(), ll.cpu().numpy()


def mif(struct, thetas_array, rw_sd, M, J, keys, thresh=0.0, n_monitors=0, unit=None):
    """IF2 (functional/mif.py:15-122): thetas (n, J, d) natural scale, keys (n, 2).
    Returns (logliks (n, M), traces (n, M+1, d) natural, final swarm (n, J, d) natural)."""
    model = struct.model
    thetas = np.asarray(thetas_array, dtype=F32)
    keys = np.asarray(keys, dtype=np.uint32).reshape(-1, 2)
    n = keys.shape[0]
    rw = rw_sd.canonicalize(model.param_names)
    theta = model.to_est(thetas)
    traces = [theta.mean(axis=1, dtype=F32)]
    lls = np.zeros((n, M), dtype=F32)
    mkeys = np.stack([prandom.split(k, M + 1)[1:] for k in keys])  # (n, M, 2)
    for m in range(M):
        new, mean, ll_per = perfilter(struct, theta, mkeys[:, m], J, rw, m, thresh, unit=unit)
        if n_monitors >= 1:
            means = theta.mean(axis=1, dtype=F32)
            mon = np.stack([prandom.split(k, n_monitors) for k in mkeys[:, m]])
            res = pfilter(struct, means, J, mon, thresh=thresh, should_trans=True)
            lls[:, m] = res["logLik"].mean(axis=1, dtype=F32)
        else:
            lls[:, m] = ll_per.astype(F32)
        traces.append(mean.astype(F32))
        theta = new
    traces = np.stack(traces, axis=1)
    return lls, model.from_est(traces), model.from_est(theta)


def mop_grad(struct, thetas_array, J, alpha, keys, K=50, form="batched"):
    """MOP(alpha) value and gradient by the native adjoint.  thetas (n, d) on the
    estimation scale, keys (n, 2).  Returns (loglik (n,) float64, grad (n, d)
    float32) with grad = d loglik / d theta_est.  The reference's
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
    for u in struct.units:
        if not np.all(u.nstep == 1):
            raise ValueError("the MOP adjoint needs exactly one Euler step per observation interval")
    ser = struct.device_series()
    md = struct.model.desc()
    th = torch.from_numpy(thetas).to(dev)
    kt = _keys_tensor(keys, dev)
    ll = torch.empty(R, dtype=torch.float64, device=dev)
    g = torch.empty((R, struct.model.d), dtype=torch.float32, device=dev)
    fid = _form_id(form)
    n = _native.lib().mp_mop_grad_workspace_size(md, ser.desc, 1 if fid == _native.MP_FORM_PLAIN else R, J, K)
    ws = torch.empty(int(n), dtype=torch.uint8, device=dev)
    _native.check(_native.lib().mp_mop_grad(md, ser.desc, _native.ptr(th), struct.model.d, _native.ptr(kt), R, J,
                                            float(alpha), int(K), fid, _native.ptr(ll), _native.ptr(g),
                                            _native.ptr(ws), int(n), _native.stream_ptr()))
    torch.cuda.current_stream().synchronize()
    return ll.cpu().numpy(), g.cpu().numpy()


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
                    ll, g = mop_grad(struct, theta_est.detach().cpu().numpy(), J, alpha, keys, K)
                    ctx.save_for_backward(torch.from_numpy(g).to(theta_est.device))
                    return torch.from_numpy(ll).to(device=theta_est.device, dtype=torch.float64)

                @staticmethod
                def backward(ctx, grad_out):
                    (g,) = ctx.saved_tensors
                    return (grad_out[:, None].to(g.dtype) * g, None, None, None, None, None)

            cls.fn = MopFunction
        return cls.fn


def mop_loglik_autograd(theta_est, struct, J, alpha, keys, K=50):
    """MOP log-likelihood as a differentiable torch function of theta_est (n, d);
    its backward is the native adjoint (the forward runs value + gradient)."""
    return _MopFn.get().apply(theta_est, struct, J, alpha, keys, K)
