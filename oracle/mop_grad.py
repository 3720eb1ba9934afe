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


def sv_mop_objective(theta_est, ys, z0, zs, anc, alpha, frozen_m=None, record=None):
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
        if record is not None:
            record.append(m.detach().clone())
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


# ---- linear Gaussian ------------------------------------------------------

def lg_noise(series, key, J, dx):
    """rinit normals (J, dx) and step normals (T, J, dx) of the filter key tree."""
    K = tf.split(key, J + 1)
    run, pk = K[0], K[1:]
    z0 = tf.normal_many(pk, dx)
    zs = []
    for t in range(series.T):
        k2 = tf.split(run)
        run, sub = k2[0], k2[1]
        sk = tf.split(sub)[1]
        zs.append(tf.normal_many(tf.split(sk, J), dx))
        run = tf.split(run)[0]
    return z0, np.stack(zs)


def lg_unpack_torch(th, spec):
    """theta (d,) float64 torch, natural scale -> A, C, LQ, LR, X0, L0."""
    dx, dy = spec.dx, spec.dy
    o = 0
    A = th[o:o + dx * dx].reshape(dx, dx); o += dx * dx
    C = th[o:o + dy * dx].reshape(dy, dx); o += dy * dx

    def tril(o, d):
        rows = []
        for i in range(d):
            row = [th[o + i * (i + 1) // 2 + j] if j <= i else th.new_zeros(()) for j in range(d)]
            rows.append(torch.stack(row))
        return torch.stack(rows), o + d * (d + 1) // 2

    LQ, o = tril(o, dx)
    LR, o = tril(o, dy)
    X0 = th[o:o + dx]; o += dx
    L0 = tril(o, dx)[0] if spec.separate_init else LQ
    return A, C, LQ, LR, X0, L0


def lg_from_est_torch(th_est, spec, identity=False):
    if identity:
        return th_est
    dx, dy = spec.dx, spec.dy
    out = th_est.clone()
    base = dx * dx + dy * dx
    for i in range(dx):
        o = base + i * (i + 1) // 2 + i
        out = out.index_put((torch.tensor([o]),), torch.exp(th_est[o:o + 1]))
    for i in range(dy):
        o = base + dx * (dx + 1) // 2 + i * (i + 1) // 2 + i
        out = out.index_put((torch.tensor([o]),), torch.exp(th_est[o:o + 1]))
    return out


def lg_mop_objective(theta_est, ys, z0, zs, anc, alpha, spec, identity=False, frozen_m=None, record=None):
    """MOP log-likelihood of the LG model (float64 torch), ancestry and noise fixed;
    ``frozen_m`` (T, J) fixes the stop-gradient values as in sv_mop_objective."""
    th = lg_from_est_torch(theta_est, spec, identity)
    A, C, LQ, LR, X0, L0 = lg_unpack_torch(th, spec)
    J = z0.shape[0]
    x = X0 + torch.as_tensor(z0, dtype=torch.float64) @ L0.T
    S = LR @ LR.T
    Lc = torch.linalg.cholesky(S)
    logdet = torch.log(torch.diagonal(Lc)).sum()
    wF = torch.full((J,), float(np.log(1.0 / J)), dtype=torch.float64)
    ll = torch.zeros((), dtype=torch.float64)
    dy = spec.dy
    for t in range(zs.shape[0]):
        wP = alpha * wF
        if t > 0:
            x = x[torch.as_tensor(anc[t - 1], dtype=torch.long)]
        x = x @ A.T + torch.as_tensor(zs[t], dtype=torch.float64) @ LQ.T
        r = torch.as_tensor(ys[t], dtype=torch.float64)[None, :] - x @ C.T
        s = torch.linalg.solve_triangular(Lc, r.T, upper=False).T
        m = -0.5 * (s * s).sum(1) - 0.5 * dy * np.log(2 * np.pi) - logdet
        if record is not None:
            record.append(m.detach().clone())
        ll = ll + torch.logsumexp(wP + m, 0) - torch.logsumexp(wP, 0)
        if t < zs.shape[0] - 1:
            a = torch.as_tensor(anc[t], dtype=torch.long)
            sg = m.detach() if frozen_m is None else torch.as_tensor(frozen_m[t], dtype=torch.float64)
            wF = (wP + m - sg)[a]
    return ll


def lg_mop_value_and_grad(model, theta_est, ys, key, J, alpha, identity=False):
    """MOP value (NumPy oracle), float64 re-evaluation and autograd gradient for
    an oracle LG model (oracle/lg.py); ``identity`` treats theta as natural."""
    ser = fl.make_series(ys, np.arange(1, len(ys) + 1.0), 0.0, nstep=1)
    m = fl.IdentityTrans(model) if identity else model
    val, tr = fl.mop(m, theta_est, ser, key, J, alpha, return_trace=True)
    z0, zs = lg_noise(ser, key, J, model.spec.dx)
    anc = np.stack(tr["ancestors"])
    th = torch.tensor(np.asarray(theta_est, np.float64), requires_grad=True)
    ll = lg_mop_objective(th, ys, z0, zs, anc, alpha, model.spec, identity)
    ll.backward()
    return float(val), float(ll.detach()), th.grad.numpy().copy()


# ---- Hessian (jax.hessian of _mop_internal_mean, ad_helpers.py:8, 56-75) ----

def _frozen_densities(objective, th0, T):
    """The densities m_t at th0 (T, J), recorded by one evaluation."""
    rec = []
    with torch.no_grad():
        objective(th0, rec)
    return np.stack([r.numpy() for r in rec])


def hessian_and_fd(objective, theta_est, T, h=1e-6, fd=True):
    """(value, gradient, Hessian) of ``objective(th, frozen_m=None)`` at
    theta_est by torch's double backward -- the stop-gradient values are the
    densities at theta_est, held fixed, as the tangent of stop_gradient is zero
    in jax.hessian's forward-over-reverse -- and the central differences of the
    frozen objective's autograd gradient (step h), the pin of the Hessian."""
    th0 = torch.tensor(np.asarray(theta_est, np.float64))
    frozen = _frozen_densities(lambda th, rec: objective(th, record=rec), th0, T)
    f = lambda th: objective(th, frozen_m=frozen)  # noqa: E731
    H = torch.autograd.functional.hessian(f, th0).numpy()
    thg = th0.clone().requires_grad_(True)
    val = f(thg)
    val.backward()
    g = thg.grad.numpy().copy()

    def grad_at(th):
        t = torch.tensor(th, requires_grad=True)
        f(t).backward()
        return t.grad.numpy().copy()

    if not fd:
        return float(val.detach()), g, H, None
    d = th0.shape[0]
    fd = np.zeros((d, d))
    for j in range(d):
        e = np.zeros(d)
        e[j] = h
        fd[:, j] = (grad_at(th0.numpy() + e) - grad_at(th0.numpy() - e)) / (2 * h)
    return float(val.detach()), g, H, fd


def sv_hessian_case(theta_est, ys, key, J, alpha):
    """The SV MOP objective at theta_est with the oracle's ancestry and noise:
    a callable objective(th, frozen_m=None, record=None) for hessian_and_fd."""
    from .sv import SVModel
    ser = fl.make_series(ys, np.arange(1, len(ys) + 1.0), 0.0, nstep=1)
    _, tr = fl.mop(SVModel(), theta_est, ser, key, J, alpha, return_trace=True)
    z0, zs = sv_noise(ser, key, J)
    anc = np.stack(tr["ancestors"])

    def objective(th, frozen_m=None, record=None):
        return sv_mop_objective(th, ys[:, 0], z0, zs, anc, alpha, frozen_m=frozen_m, record=record)

    return objective


def lg_hessian_case(model, theta_est, ys, key, J, alpha):
    """The LG MOP objective (oracle LG model) as in sv_hessian_case."""
    ser = fl.make_series(ys, np.arange(1, len(ys) + 1.0), 0.0, nstep=1)
    _, tr = fl.mop(model, theta_est, ser, key, J, alpha, return_trace=True)
    z0, zs = lg_noise(ser, key, J, model.spec.dx)
    anc = np.stack(tr["ancestors"])

    def objective(th, frozen_m=None, record=None):
        return lg_mop_objective(th, ys, z0, zs, anc, alpha, model.spec, frozen_m=frozen_m, record=record)

    return objective


# ---- train (core/algorithms/train.py:38-180) ----------------------------

def cosine_cooling(i, M, c):
    """``_cosine_cooling`` (core/algorithms/helpers.py:302-314), float32."""
    F32 = np.float32
    return F32(F32(c) + (F32(1) - F32(c)) * F32(0.5) * (F32(1) + np.cos(F32(np.pi) * F32(i) / F32(M), dtype=F32)))


def fmadam_direction(grad, m1, v1, step, beta1=0.9, beta2=0.999, eps=1e-4):
    """Full-matrix Adam (core/optimizer.py:157-213) on one parameter vector:
    returns (direction, m, F)."""
    F32 = np.float32
    m1 = (F32(beta1) * m1 + F32(1 - beta1) * grad).astype(F32)
    mh = (m1 / (F32(1) - F32(beta1) ** F32(step + 1))).astype(F32)
    v1 = (F32(beta2) * v1 + F32(1 - beta2) * np.outer(grad, grad)).astype(F32)
    fh = (v1 / (F32(1) - F32(beta2) ** F32(step + 1))).astype(F32)
    ev, vec = np.linalg.eigh(fh)
    inv = (F32(1) / np.sqrt(np.maximum(ev, F32(0)) + F32(eps))).astype(F32)
    fis = ((vec * inv[None, :]) @ vec.T).astype(F32)
    return (-(fis @ mh)).astype(F32), m1, v1


def bfgs_direction(grad, hess, prev, step, eta_i):
    """BFGS on the inverse Hessian (core/optimizer.py:216-267) on one parameter
    vector, the previous direction taken as -prev as the reference does:
    returns (direction, H, grad)."""
    F32 = np.float32
    if step <= 1:
        return (-grad).astype(F32), hess, grad
    s = (F32(np.mean(np.asarray(eta_i, F32), dtype=F32)) * (-prev)).astype(F32)
    y = (grad - prev).astype(F32)
    rho = F32(1) / F32(np.dot(y, s))
    hy = (hess @ y).astype(F32)
    yhy = F32(np.dot(y, hy))
    new = (hess - rho * np.outer(s, hy) - rho * np.outer(hy, s)
           + (rho * (rho * yhy + F32(1))) * np.outer(s, s)).astype(F32)
    if not np.isfinite(rho):
        new = hess
    return (-(new @ grad)).astype(F32), new, grad


def armijo(obj, curr, pt, grad, direction, eta, c=0.1, max_itn=10, frac=0.5):
    """Step size of the reference's backtracking line search
    (core/algorithms/train.py:442-499, stoch=False)."""
    F32 = np.float32
    eta = F32(eta)
    slope = F32(np.sum(grad * direction, dtype=F32))
    itn = 0
    while True:
        nxt = F32(obj((pt + eta * direction).astype(F32)))
        cont = (nxt > curr + (eta * F32(c)) * slope) or np.isnan(nxt)
        if not (cont and itn < max_itn):
            return eta
        eta = F32(eta * F32(frac))
        itn += 1


def pinv_hermitian(H):
    """``jnp.linalg.pinv(H, hermitian=True)`` (core/optimizer.py:285, 322):
    singular values at or below 10 d eps32 times the largest are dropped."""
    d = H.shape[-1]
    return np.linalg.pinv(np.asarray(H, np.float64), rcond=10.0 * d * float(np.finfo(np.float32).eps),
                          hermitian=True)


def newton_weight(step):
    """WeightedNewton's history weight (core/optimizer.py:315-318), float32:
    i^log(i) / (i + 1)^log(i + 1)."""
    F32 = np.float32
    i = F32(step)
    return F32(np.power(i, np.log(i, dtype=F32), dtype=F32)
               / np.power(i + F32(1), np.log(i + F32(1), dtype=F32), dtype=F32))


def train(value_and_grad, theta_est0, key, T, M, eta, alpha, alpha_cooling=1.0, beta1=0.9, beta2=0.999,
          eps=1e-8, clip_norm=None, optimizer="adam", ls=None, c=0.1, max_ls_itn=10, hessian=None):
    """``_train_internal`` / ``_train_scan_step`` (core/algorithms/train.py:38-180)
    with one monitor (n_monitors = 1: the MOP value of the same draw is the
    reported objective) and the optimizer ``optimizer``: "adam"
    (core/optimizer.py:119-153), "sgd", "fmadam" (full-matrix Adam, epsilon
    1e-4 unless ``eps`` is given) or "bfgs".  ``value_and_grad(theta, key,
    alpha)`` returns (loglik, d loglik / d theta).  eta (M, d).  ``ls``: None,
    or ``neg_loglik_pf(theta, key)``, the bootstrap filter's negative
    log-likelihood with the iteration's MOP key, for the Armijo line search.
    "newton" and "wnewton" (core/optimizer.py:270-330) take ``hessian(theta,
    key, alpha)``, the MOP log-likelihood's Hessian, drawn with the
    iteration's Hessian subkey (train.py:132-141).
    Returns (neg_logliks (M+1,), NaN first, traces (M+1, d)), float32 like the
    reference."""
    F32 = np.float32
    th = np.asarray(theta_est0, F32).copy()
    d = th.shape[0]
    m1 = np.zeros_like(th)
    v1 = np.zeros_like(th) if optimizer != "fmadam" else np.zeros((d, d), F32)
    hess, prev = np.eye(d, dtype=F32), np.zeros_like(th)
    prev_h = np.eye(d, dtype=F32)
    nll, traces = [np.nan], [th.copy()]
    for m in range(M):
        a_m = F32(F32(1) - (F32(1) - F32(alpha)) * cosine_cooling(m, M, alpha_cooling))
        k2 = tf.split(key)
        key, sub = k2[0], k2[1]
        ll, g = value_and_grad(th, sub, float(a_m))
        grad = (-np.asarray(g, np.float64) / T).astype(F32)  # gradient of the mean objective
        nll.append(F32(-ll))
        if clip_norm is not None:
            grad = np.clip(grad, -clip_norm, clip_norm).astype(F32)
        k2 = tf.split(key)
        key, hkey = k2[0], k2[1]  # the Hessian subkey
        if optimizer in ("newton", "wnewton"):
            Hm = (-np.asarray(hessian(th, hkey, float(a_m)), np.float64) / T).astype(F32)  # of the mean objective
            if optimizer == "newton" or m == 0:
                Hw = Hm
            else:
                wt = newton_weight(m)
                Hw = (wt * prev_h + (F32(1) - wt) * Hm).astype(F32)
            direction = (-(pinv_hermitian(Hw) @ grad.astype(np.float64))).astype(F32)
            prev_h = Hm
        elif optimizer == "adam":
            m1 = (F32(beta1) * m1 + F32(1 - beta1) * grad).astype(F32)
            v1 = (F32(beta2) * v1 + F32(1 - beta2) * grad * grad).astype(F32)
            mh = (m1 / (F32(1) - F32(beta1) ** F32(m + 1))).astype(F32)
            vh = (v1 / (F32(1) - F32(beta2) ** F32(m + 1))).astype(F32)
            direction = (-mh / (np.sqrt(vh) + F32(eps))).astype(F32)
        elif optimizer == "sgd":
            direction = -grad
        elif optimizer == "fmadam":
            direction, m1, v1 = fmadam_direction(grad, m1, v1, m, beta1, beta2, 1e-4 if eps == 1e-8 else eps)
        elif optimizer == "bfgs":
            direction, hess, prev = bfgs_direction(grad, hess, prev, m, eta[m])
        else:
            raise ValueError(optimizer)
        if ls is not None:
            step = armijo(lambda x: ls(x, sub), F32(-ll), th, grad, direction,
                          F32(np.mean(np.asarray(eta[m], F32), dtype=F32)), c, max_ls_itn)
            th = (th + step * direction).astype(F32)
        else:
            th = (th + np.asarray(eta[m], F32) * direction).astype(F32)
        traces.append(th.copy())
    return np.array(nll, dtype=F32), np.stack(traces)


def _adam_direction(grad, m1, v1, step, beta1=0.9, beta2=0.999, eps=1e-8):
    """Adam (core/optimizer.py:119-153) on an array of any shape, float32."""
    F32 = np.float32
    m1 = (F32(beta1) * m1 + F32(1 - beta1) * grad).astype(F32)
    v1 = (F32(beta2) * v1 + F32(1 - beta2) * grad * grad).astype(F32)
    mh = (m1 / (F32(1) - F32(beta1) ** F32(step + 1))).astype(F32)
    vh = (v1 / (F32(1) - F32(beta2) ** F32(step + 1))).astype(F32)
    return (-mh / (np.sqrt(vh) + F32(eps))).astype(F32), m1, v1


def panel_train(value_and_grad, shared0, unit0, perm, keys, T, M, eta_shared, eta_unit, alpha, alpha_cooling=1.0,
                chunk_size=1, optimizer="adam", clip_norm=None, scale=False):
    """One replicate of ``_panel_train_internal`` (core/algorithms/train.py:
    190-415) on the estimation scale.  shared0 (ns,), unit0 (U, nu) in the
    panel order (shared names, then unit names); ``perm`` maps the panel order
    to the model's (theta_model = theta_panel[perm]); keys (M, U, 2), the MOP
    key of unit u at iteration m.  ``value_and_grad(theta_model, key, alpha,
    u)`` returns (loglik_u, d loglik_u / d theta_model).

    Per iteration m, alpha_m = 1 - (1 - alpha) c_m (cosine cooling), and per
    chunk of ``chunk_size`` units, in order: every unit of the chunk is
    differentiated at the current shared values; the chunk loss is
    sum_u (-loglik_u) / (chunk_size T), its shared gradient the sum over the
    chunk, the unit gradients scaled by chunk_size (train.py:348-351); the
    shared optimizer steps with the running chunk count, the unit optimizer
    with m (train.py:357-360); shared += eta_shared[m] / n_chunks dir,
    unit += eta_unit[m] dir.  The iteration's objective is the mean over
    chunks of the chunk loss times T U.  Returns (neg_logliks (M+1,), NaN
    first, shared (M+1, ns), unit (M+1, U, nu)), float32."""
    F32 = np.float32
    sh = np.asarray(shared0, F32).copy()
    un = np.asarray(unit0, F32).copy()
    U, nu = un.shape
    ns = sh.shape[0]
    cs = int(chunk_size)
    n_chunks = U // cs
    inv = np.argsort(perm)

    def init(shape_rows, d):
        if optimizer == "fmadam":
            return np.zeros(shape_rows + (d,), F32), np.zeros(shape_rows + (d, d), F32)
        return np.zeros(shape_rows + (d,), F32), np.zeros(shape_rows + (d,), F32)

    def direction(grad, state, step):
        m1, v1 = state
        if optimizer == "sgd":
            return -grad, state
        if optimizer == "adam":
            dr, m1, v1 = _adam_direction(grad, m1, v1, step)
            return dr, (m1, v1)
        if optimizer == "fmadam":
            out = [fmadam_direction(grad[i], m1[i], v1[i], step) for i in range(grad.shape[0])]
            return (np.stack([o[0] for o in out]), (np.stack([o[1] for o in out]), np.stack([o[2] for o in out])))
        raise ValueError(f"Optimizer '{optimizer}' not supported for panel train")

    st_sh = init((1,), ns)
    st_un = [init((cs,), nu) for _ in range(n_chunks)]
    step = 0
    nll, sh_tr, un_tr = [np.nan], [sh.copy()], [un.copy()]
    for m in range(M):
        a_m = F32(F32(1) - (F32(1) - F32(alpha)) * cosine_cooling(m, M, alpha_cooling))
        chunk_nll = []
        for c in range(n_chunks):
            units = range(c * cs, (c + 1) * cs)
            res, g_sh, g_un = [], np.zeros(ns), []
            for u in units:
                th = np.concatenate([sh, un[u]])[perm]
                ll, g = value_and_grad(th, keys[m, u], float(a_m), u)
                gp = np.asarray(g, np.float64)[inv]           # d loglik / d theta_panel
                res.append(F32(-ll))
                g_sh = g_sh - gp[:ns] / (cs * T)
                g_un.append(-gp[ns:] / (cs * T))
            loss = F32(np.sum(np.array(res, F32), dtype=F32) / F32(cs * T))
            chunk_nll.append(F32(loss * F32(T * U)))
            g_sh = g_sh.astype(F32)[None]
            g_un = (np.array(g_un) * cs).astype(F32)
            if clip_norm is not None:
                g_sh = np.clip(g_sh, -clip_norm, clip_norm).astype(F32)
                g_un = np.clip(g_un, -clip_norm, clip_norm).astype(F32)
            d_sh, st_sh = direction(g_sh, st_sh, step) if ns else (g_sh, st_sh)
            d_un, st_un[c] = direction(g_un, st_un[c], m) if nu else (g_un, st_un[c])
            if scale:
                d_sh = (d_sh / np.maximum(np.linalg.norm(d_sh, axis=-1, keepdims=True), F32(1e-8))).astype(F32)
                d_un = (d_un / np.maximum(np.linalg.norm(d_un, axis=-1, keepdims=True), F32(1e-8))).astype(F32)
            sh = (sh + (np.asarray(eta_shared[m], F32) / F32(n_chunks)) * d_sh[0]).astype(F32)
            un[c * cs:(c + 1) * cs] = (un[c * cs:(c + 1) * cs] + np.asarray(eta_unit[m], F32) * d_un).astype(F32)
            step += 1
        nll.append(F32(np.mean(np.array(chunk_nll, F32), dtype=F32)))
        sh_tr.append(sh.copy())
        un_tr.append(un.copy())
    return np.array(nll, F32), np.stack(sh_tr), np.stack(un_tr)


def panel_mop(value_and_grad, shared, unit, perm, keys, T, chunk_size=1):
    """``_chunked_panel_mop_internal`` (core/algorithms/mop.py:183-270) and its
    gradient (``_vg_chunked_panel_mop_internal``, :273-275) on the estimation
    scale: theta_u = [shared, unit_u][perm] in the model's order, res_u =
    -loglik_u (float32), the chunk sums added in order in float32 and the
    total divided by U T.  ``value_and_grad(theta_model, key, u)`` returns
    (loglik_u, d loglik_u / d theta_model).  Returns (value, d/d shared,
    d/d unit)."""
    F32 = np.float32
    sh = np.asarray(shared, F32)
    un = np.asarray(unit, F32)
    U, ns = un.shape[0], sh.shape[0]
    inv = np.argsort(perm)
    res, g_sh, g_un = [], np.zeros(ns), []
    for u in range(U):
        ll, g = value_and_grad(np.concatenate([sh, un[u]])[perm], keys[u], u)
        gp = -np.asarray(g, np.float64)[inv] / (U * T)
        res.append(F32(-ll))
        g_sh = g_sh + gp[:ns]
        g_un.append(gp[ns:])
    res = np.array(res, F32).reshape(U // chunk_size, chunk_size)
    total = F32(0)
    for c in range(res.shape[0]):
        total = F32(total + res[c].sum(dtype=F32))
    return F32(total / F32(U * T)), g_sh.astype(F32), np.array(g_un).astype(F32)
