"""Random-walk sigmas and cooling for IF2 (pypomp/core/rw_sigma.py).

Only the geometric schedule is evaluated natively (the reference's default,
rw_sigma.py:112-113); ``cooling_factor`` gives the host value of any schedule.
"""

import math

import numpy as np


class RWSigma:
    def __init__(self, sigmas, init_names=()):
        if not isinstance(sigmas, dict):
            raise ValueError("sigmas must be a dictionary")
        if not isinstance(init_names, (list, tuple)):
            raise ValueError("init_names must be a list or tuple")
        clean = {}
        for n, v in sigmas.items():
            if not isinstance(n, str):
                raise ValueError("All keys in sigmas must be strings")
            if isinstance(v, bool):
                raise ValueError(f"Value for '{n}' must be a float, got bool")
            v = float(v)
            if v < 0:
                raise ValueError("All values in sigmas dictionary must be non-negative")
            clean[n] = v
        if len(init_names) != len(set(init_names)):
            raise ValueError("Duplicate names found in init_names")
        if not all(n in clean for n in init_names):
            raise ValueError("All init_names names must be in sigmas dictionary")
        self.param_names = tuple(clean)
        self.init_names = tuple(init_names)
        self.sigmas_all = np.array([clean[n] for n in self.param_names], dtype=float)
        self.init_mask = np.array([1.0 if n in set(init_names) else 0.0 for n in self.param_names])
        self.cooling_type = "geometric"
        self.a = 0.5

    def copy(self):
        o = RWSigma.__new__(RWSigma)
        o.__dict__.update({k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in self.__dict__.items()})
        return o

    def geometric_cooling(self, a):
        if not (0 <= a <= 1):
            raise ValueError("a should be between 0 and 1")
        o = self.copy()
        o.cooling_type, o.a = "geometric", float(a)
        return o

    def canonicalize(self, names):
        names = tuple(names)
        if set(names) != set(self.param_names):
            raise ValueError("RWSigma parameter names must match canonical_param_names up to reordering. "
                             f"Got {sorted(self.param_names)}, expected {sorted(names)}.")
        idx = {n: i for i, n in enumerate(self.param_names)}
        order = np.array([idx[n] for n in names], dtype=int)
        o = self.copy()
        o.param_names = names
        o.sigmas_all = self.sigmas_all[order]
        o.init_mask = self.init_mask[order]
        return o

    @property
    def sigmas_array(self):
        return (self.sigmas_all.astype(np.float32) * (np.float32(1) - self.init_mask.astype(np.float32)))

    @property
    def sigmas_init_array(self):
        return self.sigmas_all.astype(np.float32) * self.init_mask.astype(np.float32)

    @property
    def factor(self):
        """Geometric cooling base a^(1/50) as float32 (rw_sigma.py:210-213)."""
        return np.float32(self.a ** (1.0 / 50.0))

    def cooling_factor(self, nt, m, ntimes):
        frac = np.float32(np.float32(nt) / np.float32(ntimes) + np.float32(m))
        return np.float32(np.power(self.factor, frac, dtype=np.float32))
