"""Observation series, Euler time grid and their device upload.

The grid follows ``_calc_steps`` / ``_num_euler_steps`` / ``_calc_ys_covars``
(pypomp/core/algorithms/helpers.py:155-299): per observation interval either a
fixed ``nstep`` or the number of ``dt`` steps, the flat array of step sizes
and the covariates interpolated onto the grid.
"""

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _native


def calc_steps(times0, dt=None, nstep=None):
    """(nstep_array, dt_array) per interval of ``times0`` (helpers.py:158-215).

    With ``dt`` the step size is carried from interval to interval as the
    reference's loop does: each interval starts from the previous interval's
    adjusted step (``(t2 - t1) / n``), not from the caller's ``dt``.  A
    zero-length interval leaves a step of 0, so a later positive interval
    divides by zero and raises ``ZeroDivisionError``, as in the reference."""
    if dt is not None and nstep is not None:
        raise ValueError("Only nstep or dt can be provided, not both")
    if dt is None and nstep is None:
        raise ValueError("Either dt or nstep must be provided")
    tol = np.sqrt(np.finfo(float).eps)
    ns, dts = [], []
    step = None if dt is None else float(dt)
    for i in range(len(times0) - 1):
        t1, t2 = float(times0[i]), float(times0[i + 1])
        if nstep is not None:
            n, step = int(nstep), (t2 - t1) / int(nstep)
        elif t1 >= t2:
            n, step = 0, 0.0
        elif t1 + step >= t2:
            n, step = 1, t2 - t1
        else:
            n = int(np.ceil((t2 - t1) / step / (1 + tol)))
            step = (t2 - t1) / n
        ns.append(n)
        dts.append(step)
    return np.array(ns, dtype=np.int64), np.array(dts, dtype=np.float64)


def interp_covars(t, ctimes, covars, order="linear"):
    """Covariates at times ``t`` (helpers.py:218-262): linear interpolation
    between the bracketing covariate times, extrapolating linearly beyond the
    first and last, or the right-continuous step function ("constant").
    float64 arithmetic, in the reference's operation order."""
    if covars is None or ctimes is None:
        return None
    if order not in ("linear", "constant"):
        raise ValueError(f"Unsupported interpolation order: '{order}'. Must be 'linear' or 'constant'.")
    scalar = np.isscalar(t)
    ta = np.atleast_1d(np.asarray(t, dtype=np.float64))
    ctimes = np.asarray(ctimes, dtype=np.float64)
    covars = np.asarray(covars, dtype=np.float64)
    if order == "linear":
        hi = np.clip(np.searchsorted(ctimes, ta, side="left"), 1, len(ctimes) - 1)
        lo = hi - 1
        frac = (ta - ctimes[lo]) / (ctimes[hi] - ctimes[lo])
        if covars.ndim > 1:
            frac = frac[:, None]
        out = covars[lo] + (covars[hi] - covars[lo]) * frac
    else:
        idx = np.clip(np.searchsorted(ctimes, ta, side="right"), 1, len(ctimes)) - 1
        out = covars[idx]
    return out.ravel() if scalar else out


def grid_covars(times0, ctimes, covars, nstep, dts, order="linear"):
    """Covariates on the Euler grid plus the last observation time
    (helpers.py:265-299, covars_extended): the sub-step start times
    times0[i] + k dt_i of every interval, then times0[-1]; (E + 1, dc) float32."""
    if covars is None or ctimes is None:
        return None
    steps = [times0[i] + np.arange(nstep[i]) * dts[i] for i in range(len(nstep))]
    grid = np.concatenate(steps + [[times0[-1]]])
    out = interp_covars(grid, ctimes, covars, order)
    return np.asarray(out, dtype=np.float32).reshape(grid.shape[0], -1)


@dataclass
class Series:
    """One unit's data on the host: ys (T, dy) float32, its Euler grid and its
    covariates on that grid (``covars`` (E + 1, dc) float32, or None)."""
    ys: np.ndarray
    times: np.ndarray
    t0: float
    nstep: np.ndarray
    dt_ext: np.ndarray
    covars: np.ndarray = None

    @classmethod
    def build(cls, ys, times, t0, nstep=None, dt=None, ctimes=None, covars=None, order="linear"):
        """``ctimes`` (n,) and ``covars`` (n, dc): the covariate table, interpolated
        onto the grid as the reference's _calc_ys_covars does."""
        ys = np.asarray(ys, dtype=np.float32)
        if ys.ndim == 1:
            ys = ys[:, None]
        times = np.asarray(times, dtype=np.float64)
        if ys.shape[0] != times.shape[0]:
            raise ValueError("ys and times must have the same length")
        times0 = np.concatenate([[float(t0)], times])
        ns, dts = calc_steps(times0, dt=dt, nstep=nstep)
        cv = None
        if covars is not None:
            covars = np.asarray(covars, dtype=np.float64)
            if covars.ndim == 1:
                covars = covars[:, None]
            cv = grid_covars(times0, ctimes, covars, ns, dts, order)
        return cls(ys=ys, times=times.astype(np.float32), t0=float(t0), nstep=ns,
                   dt_ext=np.repeat(dts, ns).astype(np.float32), covars=cv)

    @property
    def dc(self):
        return 0 if self.covars is None else self.covars.shape[1]

    @property
    def T(self):
        return self.ys.shape[0]


class DeviceSeries:
    """Concatenated series of U units resident on the device, plus the C descriptor."""

    def __init__(self, units, device):
        import torch
        units = list(units)
        if not units:
            raise ValueError("at least one series is required")
        dy = units[0].ys.shape[1]
        t0 = units[0].t0
        for u in units:
            if u.ys.shape[1] != dy:
                raise ValueError("all units must have the same observation dimension")
            if u.t0 != t0:
                raise ValueError("all units must share t0")
        ys = np.concatenate([u.ys for u in units]).astype(np.float32)
        times = np.concatenate([u.times for u in units]).astype(np.float32)
        nstep = np.concatenate([u.nstep for u in units]).astype(np.int32)
        tidx0 = np.concatenate([np.concatenate([[0], np.cumsum(u.nstep)[:-1]]) for u in units]).astype(np.int32)
        dt_ext = np.concatenate([u.dt_ext for u in units] + [np.zeros(1, np.float32)]).astype(np.float32)
        obs_off = np.concatenate([[0], np.cumsum([u.T for u in units])]).astype(np.int64)
        dt_off = np.concatenate([[0], np.cumsum([u.dt_ext.shape[0] for u in units])[:-1]]).astype(np.int64)
        dcs = {u.dc for u in units}
        if len(dcs) != 1:
            raise ValueError("all units must have the same covariates")
        self.dc = dc = dcs.pop()
        if dc:
            covars = np.concatenate([u.covars for u in units]).astype(np.float32)
            cv_off = np.concatenate([[0], np.cumsum([u.covars.shape[0] for u in units])[:-1]]).astype(np.int64)
        else:
            covars, cv_off = np.zeros((1, 1), np.float32), np.zeros(1, np.int64)
        self.T_units = np.array([u.T for u in units], dtype=np.int64)
        self.U = len(units)
        self.Tmax = int(self.T_units.max())
        self.max_euler = int(max(int(u.nstep.sum()) for u in units))
        self.dy = dy
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)
        self.max_nstep = int(max(int(u.nstep.max()) if u.nstep.size else 0 for u in units))
        self._bufs = dict(ys=t(ys), times=t(times), nstep=t(nstep), tidx0=t(tidx0), dt_ext=t(dt_ext),
                          obs_off=t(obs_off), dt_off=t(dt_off), covars=t(covars), cv_off=t(cv_off))
        b = self._bufs
        self.desc = _native.SeriesDesc(
            ys=b["ys"].data_ptr(), times=b["times"].data_ptr(), nstep=b["nstep"].data_ptr(),
            tidx0=b["tidx0"].data_ptr(), dt_ext=b["dt_ext"].data_ptr(), obs_off=b["obs_off"].data_ptr(),
            dt_off=b["dt_off"].data_ptr(), t0=C.c_float(t0), U=self.U, Tmax=self.Tmax,
            max_euler=self.max_euler, covars=b["covars"].data_ptr() if dc else None,
            cv_off=b["cv_off"].data_ptr() if dc else None, dc=dc, max_nstep=self.max_nstep)
