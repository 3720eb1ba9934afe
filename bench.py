"""Benchmark of the particle-filter hot path on one MI355X.

Default workload (the throughput configuration): bootstrap particle filter on
the 2-D linear-Gaussian model, J = 1,000,000 particles, T = 200 observations
(seed-2 series), 64 replicate filters launched one after another in the plain
form (each filter spread over the whole GPU).  One step = the 64 filters.
Inputs are resident on the GPU before timing starts.

Prints one JSON line (see the repository's bench contract) with the
particle-steps/s value, the roofline of the dominant kernel (HIP events on the
launch stream) and the CPU baseline (the NumPy oracle on a bounded sample).
"""

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "particle-steps/s (pfilter & MOP-grad, batched vs plain GPU); loglik rel err"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def lg2_struct(seed):
    import pypomp_amd.functional as F
    from pypomp_amd.models import LGModel
    from pypomp_amd.series import Series
    from tests.helpers import workloads
    w = workloads()
    ys, theta = w[f"lg2_seed{seed}_ys"], w[f"lg2_seed{seed}_theta"]
    st = F.PompStruct(model=LGModel(2, 2, separate_init=True),
                      units=[Series.build(ys, np.arange(1, ys.shape[0] + 1.0), 0.0, nstep=1)])
    return st, theta, ys


class PlainLG:
    """Config: LG 2-D, J=1e6, T=200, 64 replicates one at a time (plain form)."""

    def __init__(self, J=1_000_000, reps=64):
        import torch
        from pypomp_amd import _native
        from pypomp_amd import random as prandom
        self.dev = _native.require_device()
        self.struct, self.theta, self.ys = lg2_struct(2)
        self.J, self.R, self.T = J, reps, self.ys.shape[0]
        self.ser = self.struct.device_series()
        self.md = self.struct.model.desc()
        keys = prandom.split(prandom.key(2), reps)
        self.keys = torch.from_numpy(np.ascontiguousarray(keys).view(np.int32)).to(self.dev)
        self.th = torch.from_numpy(np.repeat(self.theta[None], reps, 0)).to(self.dev)
        self.ll = torch.empty(reps, dtype=torch.float64, device=self.dev)
        n = _native.lib().mp_filter_workspace_size(_native.MP_ALGO_PF, self.md, self.ser.desc, reps, J,
                                                   _native.MP_FORM_PLAIN)
        self.ws = torch.empty(int(n), dtype=torch.uint8, device=self.dev)
        self.nws = int(n)

    @property
    def particle_steps(self):
        return self.J * self.T * self.R

    def step(self):
        from pypomp_amd import _native
        _native.check(_native.lib().mp_pfilter(
            self.md, self.ser.desc, _native.ptr(self.th), self.struct.model.d, _native.ptr(self.keys), None,
            self.R, self.J, 0.0, 0, _native.MP_FORM_PLAIN, _native.ptr(self.ll), None, None, None, None,
            _native.ptr(self.ws), self.nws, _native.stream_ptr()))

    # algorithmic bytes per launch (4-byte words; S = 2 state words)
    def kernel_bytes(self):
        S = 2
        return {"plain_prop": self.J * 4 * (2 + S + S + 1),  # C read (8B), gather S, write S, lw
                "plain_scan": self.J * (4 + 8)}              # lw read, C write


def kalman_exact(ys, theta):
    from oracle import kalman
    from oracle import lg as olg
    spec = olg.LGSpec(2, 2, separate_init=True)
    A, C, LQ, LR, X0, L0 = (np.asarray(m, np.float64) for m in spec.unpack(theta))
    return kalman.kalman_loglik(ys, A, C, LQ @ LQ.T, LR @ LR.T, X0, L0 @ L0.T)


def parity_check():
    """Config-1 workload (J=1000, T=50, seed-1 series): GPU vs oracle relative error."""
    import pypomp_amd.functional as F
    from oracle import filters as fl, lg as olg, threefry as tf
    st, theta, ys = lg2_struct(1)
    keys = tf.split(tf.key(4321), 2)
    got = F.pfilter(st, theta[None], 1000, keys.reshape(1, 2, 2))["logLik64"][0]
    model = olg.LGModel(olg.LGSpec(2, 2, separate_init=True))
    ser = fl.make_series(ys, np.arange(1, ys.shape[0] + 1.0), 0.0, nstep=1)
    errs = []
    for r in range(2):
        exp = float(fl.pfilter(model, theta, ser, keys[r], 1000)["logLik"])
        errs.append(abs(float(got[r]) - exp) / abs(exp))
    return max(errs)


def cpu_baseline(J=200_000, T=12):
    """NumPy oracle bootstrap filter, single thread, on a bounded sample of the workload."""
    from oracle import filters as fl, lg as olg, threefry as tf
    st, theta, ys = lg2_struct(2)
    model = olg.LGModel(olg.LGSpec(2, 2, separate_init=True))
    ser = fl.make_series(ys[:T], np.arange(1, T + 1.0), 0.0, nstep=1)
    t0 = time.perf_counter()
    fl.pfilter(model, theta, ser, tf.split(tf.key(2), 1)[0], J)
    dt = time.perf_counter() - t0
    return {"value": J * T / dt, "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "sample": f"NumPy oracle pfilter, LG 2-D, J={J}, T={T}, 1 replicate, single thread ({dt:.1f} s)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--reps", type=int, default=64)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (replicate filters shard "
                                   "across GPUs with no communication, see the design notes)"}))
        sys.exit(2)

    import torch
    from pypomp_amd import _native
    torch.cuda.set_device(0)
    wl = PlainLG(reps=args.reps)
    log(f"workload LG 2-D plain: J={wl.J} T={wl.T} reps={wl.R}")
    for _ in range(args.warmup):
        wl.step()
    torch.cuda.synchronize()
    log("warmup done")
    t0 = time.perf_counter()
    for i in range(args.steps):
        wl.step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    log(f"timed {args.steps} steps in {elapsed:.3f} s")
    ll = wl.ll.cpu().numpy()
    value = wl.particle_steps * args.steps / elapsed

    # per-kernel durations from HIP events on the launch stream (one extra step)
    _native.profile_enable(True)
    wl.step()
    torch.cuda.synchronize()
    prof = _native.profile_read()
    _native.profile_enable(False)
    kb = wl.kernel_bytes()
    kernels = {}
    for k in ("plain_prop", "plain_scan"):
        ms, n = prof[k]
        avg = ms / max(n, 1)
        kernels[k] = {"avg_us": avg * 1e3, "launches": n, "bytes": kb[k],
                      "GBps": kb[k] / (avg * 1e-3) / 1e9 if avg > 0 else None}
    dom = max(kernels, key=lambda k: prof[k][0])
    d = kernels[dom]
    roofline = {"bound": "hbm", "achieved": d["GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": d["GBps"] / HBM_PEAK_GBS, "traffic": None, "kernel": dom,
                "bytes_per_launch": d["bytes"], "avg_launch_us": d["avg_us"]}
    log(f"kernels {json.dumps(kernels)}")

    exact = kalman_exact(wl.ys, wl.theta)
    rel = parity_check()
    cpu = None if args.no_cpu_baseline else cpu_baseline()
    out = {
        "metric": METRIC, "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "pfilter LG 2-D, J=1e6, T=200, 64 replicates one at a time (plain form)",
                   "J": wl.J, "T": wl.T, "replicates": wl.R, "form": "plain"},
        "roofline": roofline, "cpu_baseline": cpu,
        "loglik_rel_err": rel,
        "loglik": {"mean": float(ll.mean()), "sd": float(ll.std(ddof=1)), "kalman_fp64": exact},
        "kernels": kernels,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
