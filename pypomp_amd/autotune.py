"""Launch-shape autotuner for the engine's kernels.

The library exposes its launch shapes as process-wide settings
(``mp_tune_set`` / ``mp_tune_get``: plain scan-tile rows and threads, plain
propagate-block rows and threads, the plain XCD-local chunk order, batched
workgroup size, MOP-adjoint workgroup size).  ``sweep`` times a caller-supplied workload
under every combination of candidate values with HIP events on the current
stream, keeps the fastest, applies it and caches it per (workload key, device),
in memory and optionally in a JSON file (``PYPOMP_AMD_TUNE_CACHE``).  The
reference has no counterpart: XLA chooses its own launch shapes.

Tuning never changes results: every setting computes the same bits.  The
resampler's fixed-point prefix sums and the log-likelihood's exp-sum (their
exact integer total) do not depend on the evaluation order.
"""

import itertools
import json
import os
import threading

from . import _native

PARAMS = {
    "plain_scan_rows": (_native.MP_TUNE_PLAIN_SCAN_ROWS, (4, 8, 16)),
    "plain_prop_rows": (_native.MP_TUNE_PLAIN_PROP_ROWS, (1, 2, 4)),
    "batched_block": (_native.MP_TUNE_BATCHED_BLOCK, (256, 512, 768)),
    "grad_block": (_native.MP_TUNE_GRAD_BLOCK, (256, 512)),
    "plain_scan_threads": (_native.MP_TUNE_PLAIN_SCAN_THREADS, (256, 512, 1024)),
    "plain_prop_threads": (_native.MP_TUNE_PLAIN_PROP_THREADS, (256, 512, 1024)),
    # 1: propagate blocks take the particle chunks of the scan tiles on their own XCD
    "plain_xcd_map": (_native.MP_TUNE_PLAIN_XCD_MAP, (1, 0)),
    # replicate filters of the plain form run at once, each on its own stream
    "plain_lanes": (_native.MP_TUNE_PLAIN_LANES, (4, 1, 2, 3)),
}

_lock = threading.Lock()
_cache = {}


def get(name):
    return int(_native.lib().mp_tune_get(PARAMS[name][0]))


def set(name, value):  # noqa: A001 - mirrors mp_tune_set
    _native.check(_native.lib().mp_tune_set(PARAMS[name][0], int(value)))


def _device_name():
    import torch
    return torch.cuda.get_device_name(torch.cuda.current_device())


def _cache_file():
    return os.environ.get("PYPOMP_AMD_TUNE_CACHE", "")


def _load_file():
    path = _cache_file()
    if path and os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def _store_file(key, cfg):
    path = _cache_file()
    if not path:
        return
    data = _load_file()
    data[key] = cfg
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def time_call(run, repeats=2):
    """Median milliseconds of ``run()`` (one warm-up call first), HIP events on
    the current stream."""
    import torch
    run()
    ts = []
    for _ in range(repeats):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        run()
        b.record()
        b.synchronize()
        ts.append(a.elapsed_time(b))
    ts.sort()
    return ts[len(ts) // 2]


def sweep(run, params, key=None, repeats=2):
    """Time ``run`` under every combination of ``params`` ({name: candidates});
    apply and return the fastest {name: value} with its time in ms.  With a
    ``key`` the result is cached (and reused without timing) per device."""
    names = list(params)
    full_key = None if key is None else f"{_device_name()}|{key}|{','.join(names)}"
    with _lock:
        if full_key is not None:
            if full_key not in _cache:
                disk = _load_file()
                if full_key in disk:
                    _cache[full_key] = disk[full_key]
            if full_key in _cache:
                best = _cache[full_key]
                for n in names:
                    set(n, best[n])
                return dict(best)
        saved = {n: get(n) for n in names}
        best = None
        try:
            for combo in itertools.product(*(params[n] for n in names)):
                for n, v in zip(names, combo):
                    set(n, v)
                ms = time_call(run, repeats)
                if best is None or ms < best["ms"]:
                    best = dict(zip(names, combo), ms=ms)
        except Exception:
            for n, v in saved.items():
                set(n, v)
            raise
        for n in names:
            set(n, best[n])
        if full_key is not None:
            _cache[full_key] = dict(best)
            _store_file(full_key, best)
        return dict(best)


def tune_pfilter(struct, theta, J, R, form="batched", key=None):
    """Tune the launch shape of ``mp_pfilter`` for R filters of J particles on
    this struct (parameters ``theta`` (d,), natural scale) and apply it."""
    import numpy as np
    import torch
    from . import random as prandom
    from .functional import _form_id, _keys_tensor, _workspace
    dev = _native.require_device()
    fid = _form_id(form)
    ser = struct.device_series()
    md = struct.model.desc()
    th = torch.from_numpy(np.ascontiguousarray(np.asarray(theta, np.float32))).to(dev)
    kt = _keys_tensor(prandom.split(prandom.key(0), R), dev)
    ll = torch.empty(R, dtype=torch.float64, device=dev)
    if fid == _native.MP_FORM_PLAIN:
        names = ("plain_scan_rows", "plain_prop_rows") + (("plain_lanes",) if R > 1 else ())
    else:
        names = ("batched_block",)
    holder = {}

    def run():
        # the workspace size depends on the launch shape: query per setting
        ws, nws = _workspace(_native.MP_ALGO_PF, md, ser, R, J, fid, dev)
        holder["ws"] = ws
        _native.check(_native.lib().mp_pfilter(
            md, ser.desc, _native.ptr(th), 0, _native.ptr(kt), None, R, J, 0.0, 0, fid, _native.ptr(ll),
            None, None, None, None, _native.ptr(ws), nws, _native.stream_ptr()))

    k = None if key is None else f"pfilter|{key}|J={J}|R={R}|{form}"
    return sweep(run, {n: PARAMS[n][1] for n in names}, key=k)
