"""Observation series, Euler time grid and their device upload.

The grid follows ``_calc_steps`` / ``_num_euler_steps`` / ``_calc_ys_covars``
(pypomp/core/algorithms/helpers.py:155-299): per observation interval either a
fixed ``nstep`` or the number of ``dt`` steps, and the flat array of step sizes.
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


@dataclass
class Series:
    """One unit's data on the host: ys (T, dy) float32 and its Euler grid."""
    ys: np.ndarray
    times: np.ndarray
    t0: float
    nstep: np.ndarray
    dt_ext: np.ndarray

    @classmethod
    def build(cls, ys, times, t0, nstep=None, dt=None):
        ys = np.asarray(ys, dtype=np.float32)
        if ys.ndim == 1:
            ys = ys[:, None]
        times = np.asarray(times, dtype=np.float64)
        if ys.shape[0] != times.shape[0]:
            raise ValueError("ys and times must have the same length")
        times0 = np.concatenate([[float(t0)], times])
        ns, dts = calc_steps(times0, dt=dt, nstep=nstep)
        return cls(ys=ys, times=times.astype(np.float32), t0=float(t0), nstep=ns,
                   dt_ext=np.repeat(dts, ns).astype(np.float32))

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
        self.T_units = np.array([u.T for u in units], dtype=np.int64)
        self.U = len(units)
        self.Tmax = int(self.T_units.max())
        self.max_euler = int(max(int(u.nstep.sum()) for u in units))
        self.dy = dy
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)
        self._bufs = dict(ys=t(ys), times=t(times), nstep=t(nstep), tidx0=t(tidx0), dt_ext=t(dt_ext),
                          obs_off=t(obs_off), dt_off=t(dt_off))
        b = self._bufs
        self.desc = _native.SeriesDesc(
            ys=b["ys"].data_ptr(), times=b["times"].data_ptr(), nstep=b["nstep"].data_ptr(),
            tidx0=b["tidx0"].data_ptr(), dt_ext=b["dt_ext"].data_ptr(), obs_off=b["obs_off"].data_ptr(),
            dt_off=b["dt_off"].data_ptr(), t0=C.c_float(t0), U=self.U, Tmax=self.Tmax,
            max_euler=self.max_euler)
