"""MOP(alpha) gradient oracle: a float64 torch restatement of the reference's
``_mop_internal`` / ``_mop_step`` (pypomp/core/algorithms/mop.py:19-135) with the
reference's stop-gradient construction ``wF = (wP + m - sg(m))[a]`` (mop.py:125),
differentiated by torch autograd -- the stand-in for ``jax.grad`` of the
reference's ``_mop_internal_mean`` (core/algorithms/ad_helpers.py:6-8), which
cannot run here.  TEST ORACLE ONLY.

The random numbers and the resampling ancestry come from a NumPy oracle run
(oracle/filters.py ``mop(..., return_trace=True)``) and are held fixed, exactly
as the reference's autodiff treats them (integer indices carry no gradient).

Pins (tests/test_oracle_mop_grad.py): the value at theta equals the NumPy
oracle's MOP value; autograd equals central finite differences of the
frozen-ancestry objective (the MOP gradient is exactly that derivative).
"""

import numpy as np
import torch

from . import filters as fl
from . import threefry as tf


def sv_from_est(th):
    mu, phi, sig, rho, V0, S0 = th.unbind(-1)
    return mu, torch.tanh(phi), torch.exp(sig), torch.tanh(rho), V0, torch.exp(S0)


def sv_noise(series, key, J):
    """rinit normals (J,) and step normals (T, J) drawn by the pfilter key tree."""
    K = tf.split(key, J + 1)
    run, pk = K[0], K[1:]
    z0 = tf.normal_many(pk, 1)[:, 0]
    zs = []
    for t in range(series.T):
        k2 = tf.split(run)
        run, sub = k2[0], k2[1]
        sk = tf.split(sub)[1]
        zs.append(tf.normal_many(tf.split(sk, J), 1)[:, 0])
        run = tf.split(run)[0]
    return z0, np.stack(zs)


def sv_mop_objective(theta_est, ys, z0, zs, anc, alpha, frozen_m=None):
    """MOP log-likelihood (float64 torch).  anc[t] = ancestors chosen after step t.
    With ``frozen_m`` (T, J) the stop-gradient terms use those fixed values (the
    frozen-ancestry objective whose gradient at theta0 is the MOP gradient)."""
    mu, phi, sig, rho, V0, S0 = sv_from_est(theta_est)
    J = z0.shape[0]
    v = V0 + S0 * torch.as_tensor(z0, dtype=torch.float64)
    wF = torch.full((J,), float(np.log(1.0 / J)), dtype=torch.float64)
    ll = torch.zeros((), dtype=torch.float64)
    omr2 = 1 - rho * rho
    for t in range(zs.shape[0]):
        wP = alpha * wF
        if t > 0:
            v = v[torch.as_tensor(anc[t - 1], dtype=torch.long)]
        z = torch.as_tensor(zs[t], dtype=torch.float64)
        v = mu + phi * (v - mu) + sig * z
        u = float(ys[t]) * torch.exp(-0.5 * v) - rho * z
        m = -0.5 * np.log(2 * np.pi) - 0.5 * torch.log(omr2) - 0.5 * v - 0.5 * u * u / omr2
        ll = ll + torch.logsumexp(wP + m, 0) - torch.logsumexp(wP, 0)
        if t < zs.shape[0] - 1:
            a = torch.as_tensor(anc[t], dtype=torch.long)
            sg = m.detach() if frozen_m is None else torch.as_tensor(frozen_m[t], dtype=torch.float64)
            wF = (wP + m - sg)[a]
    return ll


def sv_mop_value_and_grad(theta_est, ys, key, J, alpha):
    """MOP value (oracle forward), its float64 re-evaluation on the oracle's ancestry
    and the autograd gradient; also the noise and ancestry used."""
    from .sv import SVModel
    model = SVModel()
    ser = fl.make_series(ys, np.arange(1, len(ys) + 1.0), 0.0, nstep=1)
    val, tr = fl.mop(model, theta_est, ser, key, J, alpha, return_trace=True)
    z0, zs = sv_noise(ser, key, J)
    anc = np.stack(tr["ancestors"])
    th = torch.tensor(np.asarray(theta_est, np.float64), requires_grad=True)
    ll = sv_mop_objective(th, ys[:, 0], z0, zs, anc, alpha)
    ll.backward()
    return float(val), ll, th, z0, zs, anc


def sv_mop_reference_grad(theta_est, ys, key, J, alpha):
    """(value, autograd gradient, trace) at theta_est with the oracle's ancestry."""
    val, ll, th, z0, zs, anc = sv_mop_value_and_grad(theta_est, ys, key, J, alpha)
    with torch.no_grad():
        frozen_m = np.stack([_sv_meas(th, ys[:, 0], z0, zs, anc, t) for t in range(zs.shape[0])])
    return float(val), float(ll.detach()), th.grad.numpy().copy(), dict(z0=z0, zs=zs, anc=anc, frozen_m=frozen_m)


def _sv_meas(th, ys, z0, zs, anc, t_want):
    mu, phi, sig, rho, V0, S0 = sv_from_est(th)
    v = V0 + S0 * torch.as_tensor(z0, dtype=torch.float64)
    for t in range(t_want + 1):
        if t > 0:
            v = v[torch.as_tensor(anc[t - 1], dtype=torch.long)]
        z = torch.as_tensor(zs[t], dtype=torch.float64)
        v = mu + phi * (v - mu) + sig * z
    u = float(ys[t_want]) * torch.exp(-0.5 * v) - rho * z
    omr2 = 1 - rho * rho
    return (-0.5 * np.log(2 * np.pi) - 0.5 * torch.log(omr2) - 0.5 * v - 0.5 * u * u / omr2).numpy()


def frozen_fd_grad(theta_est, ys, trace, alpha, h=1e-3):
    """Central differences of the frozen-ancestry objective (step h on the estimation scale)."""
    g = np.zeros(len(theta_est))
    for p in range(len(theta_est)):
        vals = []
        for sgn in (1, -1):
            th = np.asarray(theta_est, np.float64).copy()
            th[p] += sgn * h
            with torch.no_grad():
                vals.append(float(sv_mop_objective(torch.tensor(th), ys[:, 0], trace["z0"], trace["zs"],
                                                   trace["anc"], alpha, frozen_m=trace["frozen_m"])))
        g[p] = (vals[0] - vals[1]) / (2 * h)
    return g
