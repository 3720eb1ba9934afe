"""Host-side aggregation over replicates: log-mean-exp and its jackknife SE
(pypomp/maths.py:24-157), float64 NumPy."""

import numpy as np


def logmeanexp(x, axis=None, ignore_nan=False):
    x = np.asarray(x, dtype=float)
    if ignore_nan:
        with np.errstate(all="ignore"):
            mx = np.nanmax(x, axis=axis, keepdims=True)
            mask = np.isnan(x)
            xs = np.where(mask, -np.inf, x)
            cnt = np.sum(~mask, axis=axis, keepdims=True)
            res = np.log(np.sum(np.exp(xs - mx), axis=axis, keepdims=True) / cnt) + mx
    else:
        if axis is None and x.size == 0:
            return np.nan
        mx = np.max(x, axis=axis, keepdims=True)
        with np.errstate(all="ignore"):
            res = np.log(np.mean(np.exp(x - mx), axis=axis, keepdims=True)) + mx
    if axis is None:
        return res.item()
    return np.squeeze(res, axis=axis)


def logmeanexp_se(x, axis=None, ignore_nan=False):
    """Leave-one-out jackknife SE of ``logmeanexp`` (NumPy std convention, ddof=0)."""
    if axis is not None:
        return np.apply_along_axis(logmeanexp_se, axis, x, ignore_nan=ignore_nan)
    x = np.asarray(x, dtype=float)
    if ignore_nan:
        x = x[~np.isnan(x)]
    n = x.size
    if n <= 1:
        return np.nan
    mx = np.max(x)
    s = np.exp(x - mx)
    tot = np.sum(s)
    with np.errstate(all="ignore"):
        jack = np.log((tot - s) / (n - 1)) + mx
    if np.sum(x == mx) == 1:
        i = int(np.argmax(x))
        if tot - s[i] <= 0:
            jack[i] = logmeanexp(np.delete(x, i))
    return float(np.sqrt(n - 1) * np.std(jack, ddof=0))
