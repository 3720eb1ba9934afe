"""Random-walk sigmas and cooling for IF2 (pypomp/core/rw_sigma.py).

The geometric schedule (the reference's default, rw_sigma.py:112-113) is
evaluated in the IF2 kernel; the cosine, hyperbolic and custom schedules
(rw_sigma.py:193-225, 308-390) are evaluated here in float32, one factor per
observation of the iteration, and handed to the kernel as a table.
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
        self.c = self.M = self.s = None
        self.custom_fn = None

    def __getstate__(self):
        # a custom schedule is a user function: cloudpickle carries it by value,
        # as the reference pickles it (rw_sigma.py:392-410)
        state = dict(self.__dict__)
        if state.get("custom_fn") is not None:
            import cloudpickle
            state["custom_fn"] = ("cloudpickle", cloudpickle.dumps(state["custom_fn"]))
        return state

    def __setstate__(self, state):
        fn = state.get("custom_fn")
        if isinstance(fn, tuple) and len(fn) == 2 and fn[0] == "cloudpickle":
            import cloudpickle
            state = dict(state, custom_fn=cloudpickle.loads(fn[1]))
        state.setdefault("c", None)
        state.setdefault("M", None)
        state.setdefault("s", None)
        state.setdefault("custom_fn", None)
        self.__dict__.update(state)

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

    def cosine_cooling(self, c, M):
        """c + (1 - c) (1 + cos(pi t / M)) / 2 at t = m + nt / ntimes (rw_sigma.py:308-334)."""
        if not (0 <= c <= 1):
            raise ValueError("c should be between 0 and 1")
        if M <= 0:
            raise ValueError("M must be positive")
        o = self.copy()
        o.cooling_type, o.c, o.M = "cosine", float(c), int(M)
        return o

    def hyperbolic_cooling(self, s):
        """1 / (1 + s t) at t = m + nt / ntimes (rw_sigma.py:336-359)."""
        if s < 0:
            raise ValueError("s must be non-negative")
        o = self.copy()
        o.cooling_type, o.s = "hyperbolic", float(s)
        return o

    def custom_cooling(self, cooling_fn):
        """A user schedule ``cooling_fn(nt, m, ntimes) -> float`` (rw_sigma.py:361-376)."""
        o = self.copy()
        o.cooling_type, o.custom_fn = "custom", cooling_fn
        return o

    def cooled(self, factor):
        """Every standard deviation scaled by ``factor`` (rw_sigma.py:378-385)."""
        if factor < 0:
            raise ValueError("factor must be >= 0")
        o = self.copy()
        o.sigmas_all = np.asarray(self.sigmas_all) * factor
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
        """The schedule at sub-step nt of iteration m (rw_sigma.py:193-225),
        float32 like the reference's traced arithmetic (nt, m may be arrays)."""
        F = np.float32
        frac = (np.asarray(nt, dtype=F) / F(ntimes) + F(m)).astype(F)
        if self.cooling_type == "geometric":
            return np.power(self.factor, frac, dtype=F)
        if self.cooling_type == "cosine":
            progress = (frac / F(self.M)).astype(F)
            return (F(self.c) + F((1.0 - self.c) * 0.5) * (F(1) + np.cos(F(np.pi) * progress, dtype=F))).astype(F)
        if self.cooling_type == "hyperbolic":
            return (F(1) / (F(1) + F(self.s) * frac)).astype(F)
        if self.cooling_type == "custom":
            nts = np.atleast_1d(np.asarray(nt))
            out = np.array([F(self.custom_fn(int(k), int(m), int(ntimes))) for k in nts], dtype=F)
            return out if np.ndim(nt) else out[0]
        return np.ones_like(frac)

    @property
    def geometric(self):
        return self.cooling_type == "geometric"

    def cooling_table(self, m, T_units):
        """Factors of iteration m at every observation row of a series of units
        with lengths T_units (the kernel's table), or None for the geometric
        schedule, which the kernel evaluates itself."""
        if self.geometric:
            return None
        return np.concatenate([np.atleast_1d(self.cooling_factor(np.arange(int(T)), m, int(T)))
                               for T in T_units]).astype(np.float32)
