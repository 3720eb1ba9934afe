"""Optimizers and learning-rate schedules of ``train`` (core/optimizer.py,
core/learning_rate.py).  The update runs on device tensors (float32, as the
reference), so an M-step run synchronises with the host only at its end."""

from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class Optimizer:
    clip_norm: float = None
    scale: bool = False
    ls: bool = False
    c: float = 0.1
    max_ls_itn: int = 10

    def init_state(self, theta):
        raise NotImplementedError

    def step(self, grad, state, step_num, eta_i=None):
        """(direction, new state) for gradients ``grad`` (n, d) of the mean
        negative log-likelihood; ``eta_i`` (d,) is the iteration's learning
        rates (BFGS uses their mean)."""
        raise NotImplementedError


@dataclass(frozen=True)
class SGD(Optimizer):
    """direction = -grad (core/optimizer.py SGD)."""

    def init_state(self, theta):
        return ()

    def step(self, grad, state, step_num, eta_i=None):
        return -grad, state


@dataclass(frozen=True)
class Adam(Optimizer):
    """Adam (core/optimizer.py:119-153)."""
    beta1: float = 0.9
    beta2: float = 0.999
    epsilon: float = 1e-8

    def init_state(self, theta):
        import torch
        return torch.zeros_like(theta), torch.zeros_like(theta)

    def step(self, grad, state, step_num, eta_i=None):
        # float32 arithmetic; the constants 1 - beta are formed in double and
        # rounded once, as JAX does with Python-float operands
        import torch
        F = np.float32
        m, v = state
        m = float(F(self.beta1)) * m + float(F(1 - self.beta1)) * grad
        v = float(F(self.beta2)) * v + float(F(1 - self.beta2)) * (grad * grad)
        c1 = F(1) - np.power(F(self.beta1), F(step_num + 1), dtype=F)
        c2 = F(1) - np.power(F(self.beta2), F(step_num + 1), dtype=F)
        mh = m / float(c1)
        vh = v / float(c2)
        return -mh / (torch.sqrt(vh) + float(F(self.epsilon))), (m, v)


@dataclass(frozen=True)
class FullMatrixAdam(Optimizer):
    """Full-matrix Adam (core/optimizer.py:157-213): first moment m and the
    running second-moment matrix F of the gradients, bias-corrected, and the
    direction -(F_hat)^(-1/2) m_hat with the inverse square root taken through
    an eigendecomposition, eigenvalues floored at 0 and epsilon added.  The
    d x d eigendecompositions run on the host in float32."""
    beta1: float = 0.9
    beta2: float = 0.999
    epsilon: float = 1e-4

    def init_state(self, theta):
        import torch
        n, d = theta.shape
        return torch.zeros_like(theta), torch.zeros((n, d, d), dtype=theta.dtype, device=theta.device)

    def step(self, grad, state, step_num, eta_i=None):
        import torch
        F = np.float32
        m, v = state
        m = float(F(self.beta1)) * m + float(F(1 - self.beta1)) * grad
        mh = m / float(F(1) - np.power(F(self.beta1), F(step_num + 1), dtype=F))
        v = float(F(self.beta2)) * v + float(F(1 - self.beta2)) * (grad[:, :, None] * grad[:, None, :])
        vh = (v / float(F(1) - np.power(F(self.beta2), F(step_num + 1), dtype=F))).cpu()
        evals, evecs = torch.linalg.eigh(vh)
        inv = 1.0 / torch.sqrt(torch.clamp(evals, min=0.0) + float(F(self.epsilon)))
        fis = (evecs * inv[:, None, :]) @ evecs.transpose(1, 2)
        direction = -(fis.to(grad.device) @ mh[:, :, None])[:, :, 0]
        return direction, (m, v)


@dataclass(frozen=True)
class BFGS(Optimizer):
    """Quasi-Newton BFGS on the inverse Hessian (core/optimizer.py:216-267).
    From the third iteration (step_num > 1): s = mean(eta_i) (-g_prev), y = g -
    g_prev, rho = 1 / (y . s), H <- H - rho s (Hy)^T - rho (Hy) s^T + rho (rho
    y^T H y + 1) s s^T (kept when rho is not finite), direction -H g; before,
    direction -g with H unchanged.  The previous direction is taken as
    -g_prev, as the reference does."""

    def init_state(self, theta):
        import torch
        n, d = theta.shape
        eye = torch.eye(d, dtype=theta.dtype, device=theta.device)
        return eye.expand(n, d, d).clone(), torch.zeros_like(theta)

    def step(self, grad, state, step_num, eta_i=None):
        import torch
        if eta_i is None:
            raise ValueError("BFGS optimizer requires eta_i")
        hess, prev = state
        if step_num <= 1:
            return -grad, (hess, grad)
        s = eta_i.mean() * (-prev)                             # (n, d)
        y = grad - prev
        rho = 1.0 / (y * s).sum(dim=1)                         # (n,)
        hy = (hess @ y[:, :, None])[:, :, 0]
        yhy = (y * hy).sum(dim=1)
        r = rho[:, None, None]
        new = (hess - r * (s[:, :, None] * hy[:, None, :]) - r * (hy[:, :, None] * s[:, None, :])
               + (r * (rho * yhy + 1.0)[:, None, None]) * (s[:, :, None] * s[:, None, :]))
        new = torch.where(torch.isfinite(rho)[:, None, None], new, hess)
        return -(new @ grad[:, :, None])[:, :, 0], (new, grad)


def _pinv_hermitian(H):
    """``jnp.linalg.pinv(H, hermitian=True)`` over a batch (n, d, d): singular
    values at or below 10 d eps32 times the largest are dropped."""
    import torch
    d = H.shape[-1]
    return torch.linalg.pinv(H, rtol=10.0 * d * float(np.finfo(np.float32).eps), hermitian=True)


@dataclass(frozen=True)
class Newton(Optimizer):
    """Newton step on the MOP Hessian (core/optimizer.py:270-289): direction =
    -pinv(H) grad with H the Hessian of the mean objective, from the native
    Hessian kernel (``train`` passes it as ``hess``)."""
    needs_hessian = True

    def init_state(self, theta):
        return ()

    def step(self, grad, state, step_num, eta_i=None, hess=None):
        if hess is None:
            raise ValueError("Newton optimizer requires compute_hessian_fn")
        return -(_pinv_hermitian(hess) @ grad[:, :, None])[:, :, 0], ()


@dataclass(frozen=True)
class WeightedNewton(Optimizer):
    """Weighted Newton step (core/optimizer.py:292-330): from the second step
    on, H_w = w H_prev + (1 - w) H with w = i^log(i) / (i + 1)^log(i + 1),
    i the step number, and H_prev the previous step's Hessian (the identity
    before the first); direction = -pinv(H_w) grad."""
    needs_hessian = True

    def init_state(self, theta):
        import torch
        n, d = theta.shape
        return (torch.eye(d, dtype=theta.dtype, device=theta.device).expand(n, d, d).clone(),)

    def step(self, grad, state, step_num, eta_i=None, hess=None):
        if hess is None:
            raise ValueError("WeightedNewton optimizer requires compute_hessian_fn")
        (prev,) = state
        if step_num == 0:
            hw = hess
        else:
            F = np.float32
            i = F(step_num)
            wt = F(np.power(i, np.log(i, dtype=F), dtype=F) / np.power(i + F(1), np.log(i + F(1), dtype=F), dtype=F))
            hw = float(wt) * prev + float(F(1) - wt) * hess
        return -(_pinv_hermitian(hw) @ grad[:, :, None])[:, :, 0], (hess,)


class LearningRate:
    """Per-parameter learning rates, constant or per-iteration schedules
    (core/learning_rate.py)."""

    def __init__(self, rates):
        if not isinstance(rates, dict) or not rates:
            raise ValueError("rates must be a non-empty dictionary")
        clean = {}
        for n, v in rates.items():
            if not isinstance(n, str):
                raise ValueError("All keys in rates must be strings")
            arr = np.asarray(v, dtype=np.float64)
            if arr.ndim > 1 or np.any(arr < 0) or not np.all(np.isfinite(arr)):
                raise ValueError(f"learning rate of '{n}' must be a non-negative float or 1-D schedule")
            clean[n] = arr
        self.param_names = tuple(clean)
        self.rates = clean

    def to_array(self, param_names, M):
        """(M, d) float32 schedule in the order ``param_names``."""
        if set(param_names) != set(self.param_names):
            raise ValueError("LearningRate parameter names must match the model's parameter names, "
                             f"got {sorted(self.param_names)}, expected {sorted(param_names)}")
        cols = []
        for n in param_names:
            r = self.rates[n]
            if r.ndim == 0:
                cols.append(np.full(M, float(r)))
            else:
                if r.shape[0] < M:
                    raise ValueError(f"schedule of '{n}' is shorter than M = {M}")
                cols.append(r[:M])
        return np.stack(cols, axis=1).astype(np.float32)
