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

    def step(self, grad, state, step_num):
        raise NotImplementedError


@dataclass(frozen=True)
class SGD(Optimizer):
    """direction = -grad (core/optimizer.py SGD)."""

    def init_state(self, theta):
        return ()

    def step(self, grad, state, step_num):
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

    def step(self, grad, state, step_num):
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
