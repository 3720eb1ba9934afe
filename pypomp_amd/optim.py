"""Optimizers and learning-rate schedules of ``train`` (core/optimizer.py,
core/learning_rate.py).  The update runs on device tensors (float32, as the
reference), so an M-step run synchronises with the host only at its end."""

from collections.abc import Mapping
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
    """Per-parameter learning rates (core/learning_rate.py:11-365): each a
    constant or a schedule of one length M, kept as ``rates_all_arr`` of shape
    (d,) or (M, d) in the order the rates were given; ``to_array`` gives the
    (M, k) float32 schedule of any k of the names, in the order asked, and the
    decay methods multiply every rate by a per-iteration factor."""

    def __init__(self, rates):
        if not isinstance(rates, Mapping):
            raise ValueError("rates must be a Mapping (e.g., dict)")
        vals, lengths = [], set()
        for n, v in rates.items():
            if not isinstance(n, str):
                raise ValueError("All keys in rates must be strings (parameter names)")
            if isinstance(v, (int, float, np.number)):
                vals.append(float(v))
            elif isinstance(v, (list, np.ndarray)):
                a = np.asarray(v, dtype=np.float64)
                if a.ndim != 1:
                    raise ValueError(f"Learning rate schedule for '{n}' must be 1D")
                vals.append(a)
                lengths.add(a.size)
            else:
                raise TypeError(f"Learning rate for '{n}' must be float or 1D sequence, got {type(v).__name__}")
        if len(lengths) > 1:
            raise ValueError(f"All 1D learning rate schedules must have the same length. Got lengths: "
                             f"{sorted(lengths)}")
        self.param_names = tuple(rates.keys())
        if not vals:
            self.rates_all_arr = np.empty((0,), np.float64)
        elif lengths:
            L = lengths.pop()
            self.rates_all_arr = np.column_stack([np.full(L, v) if isinstance(v, float) else v for v in vals])
        else:
            self.rates_all_arr = np.asarray(vals, dtype=np.float64)

    @classmethod
    def _of(cls, names, arr):
        out = object.__new__(cls)
        out.param_names = tuple(names)
        out.rates_all_arr = arr
        return out

    def _column(self, i):
        a = self.rates_all_arr
        if a.ndim == 1:
            return float(a[i])
        col = a[:, i]
        return float(col[0]) if col.size and np.all(col == col[0]) else col

    def _varying(self):
        a = self.rates_all_arr
        for i, n in enumerate(self.param_names):
            if a.ndim == 2 and not np.all(a[:, i] == a[0, i]):
                return n
        return self.param_names[0] if self.param_names else "PARAM_NAMES_NONE"

    def to_array(self, param_names, M):
        """(M, len(param_names)) float32 schedule in the order ``param_names``;
        every name must have a rate, a schedule must have length M."""
        idx = {n: i for i, n in enumerate(self.param_names)}
        for n in param_names:
            if n not in idx:
                raise ValueError(f"Parameter '{n}' not found in learning rates")
        order = np.array([idx[n] for n in param_names], dtype=np.int64)
        a = self.rates_all_arr
        if a.ndim == 1:
            out = np.tile(a[order], (max(M, 1), 1))
        else:
            if a.shape[0] != M:
                raise ValueError(f"Learning rate schedule for '{self._varying()}' has length {a.shape[0]}, "
                                 f"expected M={M}")
            out = a[:, order]
        return np.ascontiguousarray(out, dtype=np.float32)

    def _decayed(self, factor, M, kind):
        a = self.rates_all_arr
        if a.ndim == 2 and a.shape[0] != M:
            raise ValueError(f"Cannot apply {kind} decay of length {M} to schedule of length {a.shape[0]} for "
                             f"'{self._varying()}'")
        return LearningRate._of(self.param_names, factor[:, None] * (a[None, :] if a.ndim == 1 else a))

    def cosine_decay(self, final_factor, M):
        """Rates times f + (1 - f) (1 + cos(pi m / M)) / 2, m = 0 .. M-1."""
        if not 0 <= final_factor <= 1:
            raise ValueError("final_factor should be between 0 and 1")
        m = np.arange(M)
        return self._decayed(final_factor + (1.0 - final_factor) * 0.5 * (1.0 + np.cos(np.pi * m / M)), M,
                             "cosine")

    def geometric_decay(self, decay_rate, M):
        """Rates times decay_rate^m."""
        if not 0 <= decay_rate <= 1:
            raise ValueError("decay_rate should be between 0 and 1")
        return self._decayed(decay_rate ** np.arange(M), M, "geometric")

    def linear_decay(self, final_factor, M):
        """Rates times a linear ramp from 1 to final_factor over M iterations."""
        if not 0 <= final_factor <= 1:
            raise ValueError("final_factor should be between 0 and 1")
        return self._decayed(np.linspace(1.0, final_factor, M), M, "linear")

    def hyperbolic_decay(self, decay_rate, M):
        """Rates times 1 / (1 + decay_rate m)."""
        if decay_rate < 0:
            raise ValueError("decay_rate should be non-negative")
        return self._decayed(1.0 / (1.0 + decay_rate * np.arange(M)), M, "hyperbolic")

    @property
    def rates(self):
        return {n: self._column(i) for i, n in enumerate(self.param_names)}

    def __getitem__(self, name):
        if name not in self.param_names:
            raise KeyError(f"Parameter '{name}' not found in learning rates.")
        return self._column(self.param_names.index(name))

    def __contains__(self, name):
        return name in self.param_names

    def __len__(self):
        return len(self.param_names)

    def __iter__(self):
        return iter(self.param_names)

    def keys(self):
        return self.rates.keys()

    def values(self):
        return self.rates.values()

    def items(self):
        return self.rates.items()

    def get(self, name, default=None):
        return self[name] if name in self.param_names else default

    def __eq__(self, other):
        return (isinstance(other, LearningRate) and self.param_names == other.param_names
                and np.array_equal(self.rates_all_arr, other.rates_all_arr))

    __hash__ = None

    def __str__(self):
        parts = []
        for n, v in self.rates.items():
            if isinstance(v, float):
                parts.append(f"'{n}': {v:.4g}")
            elif v.size == 0:
                parts.append(f"'{n}': []")
            elif v.size <= 5:
                parts.append(f"'{n}': [" + ", ".join(f"{x:.4g}" for x in v) + "]")
            else:
                parts.append(f"'{n}': [{v[0]:.4g} ... {v[-1]:.4g}] (len={v.size})")
        return "LearningRate(\n    " + "\n    ".join(parts) + "\n)"

    __repr__ = __str__

