"""Benchmark of the particle-filter hot path on one MI355X.

Default workload (the throughput configuration): bootstrap particle filter on
the 2-D linear-Gaussian model, J = 1,000,000 particles, T = 200 observations
(seed-2 series), 64 replicate filters in the plain form (each filter spread
over the whole GPU), run in lanes: two filters at a time, each on its own
stream, so that the latency-bound scan of one overlaps the VALU-bound
propagate of the other.  One step = the 64 filters.  Inputs are resident on
the GPU before timing starts.

Prints one JSON line with the particle-steps/s value, the roofline of the
dominant kernel (HIP events on the launch stream, one extra step with one lane
so that each launch has the GPU to itself) and of the whole filter step in the
timed lanes, the CPU baseline (the C
restatement of the filter on a bounded sample of the same workload), the
replicates' log-likelihood mean and sd against the exact FP64 Kalman value
(stored with the data) and, under "extra", the batched workloads (SEIR
pfilter, MOP gradient batched and plain, panel IF2).
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
# HBM-side bytes per launch at J = 1e6 from rocprofv3 FETCH_SIZE / WRITE_SIZE
# passes (profiles/plain_lg2_J1e6_traffic.md): FETCH_SIZE x 2 + WRITE_SIZE x 1, the
# factors that make a 1 GiB device copy read 1 GiB and write 1 GiB.
MEASURED_TRAFFIC_J1E6 = {"plain_prop": 30.60e6, "plain_scan": 12.24e6,
                        # the shapes taken with three or four lanes
                        "plain_prop_wide": 30.64e6, "plain_scan_wide": 12.35e6}
# HBM bytes per launch of the SV MOP-gradient kernel at J=1000, T=1000, K=50, 512
# filters (FETCH_SIZE x 2 + WRITE_SIZE, copy-calibrated)
MEASURED_TRAFFIC_MOP = 19.03e9
# HBM bytes per launch of the batched SEIR filter at J=20000, T=260, 256 filters
# (profiles/seir_batched_J20000_T260_R256_kernel_trace.md)
MEASURED_TRAFFIC_SEIR = 86.33e9
# HBM bytes per launch (one IF2 iteration over 800 swarms) of the panel IF2 kernel
# (profiles/panel_if2_seir_U100_J2000_M50_kernel_trace.md)
MEASURED_TRAFFIC_PANEL = 61.02e9
# VALU instructions one lane executes per particle-step (SQ_INSTS_VALU x 64 /
# particle-steps per dispatch, rocprofv3 --pmc; profiles/valu_plain_lg2_seir_mopgrad.md).
# The kernels are VALU- or latency-bound rather than HBM-bound, so each roofline
# object also carries the VALU issue rate implied by the live kernel time.
VALU_INSTR_PER_PSTEP = {"plain_prop": 656.0, "plain_scan": 120.0, "seir": 9171.0, "mop_grad": 762.0,
                        "panel_if2": 13709.0,
                        # the plain kernels in the shapes taken with three or four lanes
                        # (propagate 256 x 4, scan 256 x 16)
                        "plain_prop_wide": 593.0, "plain_scan_wide": 66.0}
# The VALU issue ceiling, measured in this run by the calibration probes
# (valu_probe): the best of five instruction mixes (integer add/xor with FMAs
# issues about 1.7 wave64 instructions per CU-cycle on gfx950, above one per
# cycle) and the threefry mix (add/rotate/xor, about 1.05 per CU-cycle), the
# ceiling of the random-number-bound kernels.  Set by main() before any
# workload; the single-issue figure 256 CUs x 2.4 GHz stands in until then.
VALU_PEAK = {"best": 256 * 2.4, "threefry": 256 * 2.4, "source": "single issue, 256 CUs x 2.4 GHz"}


def valu_rate(instr_per_pstep, psteps, seconds):
    """VALU issue rate (wave64 Ginstr/s) of a kernel that ran `psteps` particle-steps
    in `seconds`, from its measured instruction count per particle-step, against
    the calibrated ceilings."""
    if not seconds or instr_per_pstep is None:
        return None
    g = instr_per_pstep * psteps / 64.0 / seconds / 1e9
    return {"achieved": g, "peak": VALU_PEAK["best"], "unit": "Ginstr/s (wave64)", "frac": g / VALU_PEAK["best"],
            "threefry_mix_peak": VALU_PEAK["threefry"], "frac_of_threefry_mix": g / VALU_PEAK["threefry"],
            "peak_source": VALU_PEAK["source"], "instr_per_particle_step": instr_per_pstep}


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
    """Config: LG 2-D, J=1e6, T=200, 64 replicates (plain form, in the library's
    lanes: two filters at a time by default)."""

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

    # Algorithmic bytes per launch: 16 (S + 1) bytes per particle-step in 4-byte
    # words (propagate read + write 2S, gather read + write 2S, log-weight write 1
    # and read 1, ancestor write 1 and read 1), split between the two kernels by
    # stage: k_plain_prop propagates, gathers, writes the log-weight and reads the
    # ancestor (4S + 2 words); k_plain_scan reads the log-weight and writes the
    # ancestor (prefix) (2 words).  S = 2.
    def kernel_bytes(self):
        S = 2
        return {"plain_prop": self.J * 4 * (4 * S + 2),
                "plain_scan": self.J * 4 * 2}


class BatchedSEIR:
    """Config: SEIR (7 Euler-binomial sub-steps/week), J=20000, T=260, 256 replicate
    filters in one persistent launch (batched form)."""

    def __init__(self, J=20000, reps=256, form="batched"):
        import torch
        import pypomp_amd.functional as F
        from pypomp_amd import _native
        from pypomp_amd import random as prandom
        from pypomp_amd.models import SEIRModel
        from pypomp_amd.series import Series
        from tests.helpers import workloads
        self.dev = _native.require_device()
        w = workloads()
        ys, self.theta = w["seir_seed3_ys"], w["seir_seed3_theta"]
        self.struct = F.PompStruct(model=SEIRModel(), units=[Series.build(ys, np.arange(1, len(ys) + 1.0), 0.0,
                                                                          nstep=7)])
        self.J, self.R, self.T = J, reps, len(ys)
        self.ser = self.struct.device_series()
        self.md = self.struct.model.desc()
        keys = prandom.split(prandom.key(3), reps)
        self.keys = torch.from_numpy(np.ascontiguousarray(keys).view(np.int32)).to(self.dev)
        self.th = torch.from_numpy(np.ascontiguousarray(self.theta)).to(self.dev)
        self.ll = torch.empty(reps, dtype=torch.float64, device=self.dev)
        self.fid = _native.MP_FORM_PLAIN if form == "plain" else _native.MP_FORM_BATCHED
        n = _native.lib().mp_filter_workspace_size(_native.MP_ALGO_PF, self.md, self.ser.desc, reps, J, self.fid)
        self.ws = torch.empty(int(n), dtype=torch.uint8, device=self.dev)
        self.nws = int(n)
        self.form = form

    @property
    def particle_steps(self):
        return self.J * self.T * self.R

    def step(self):
        from pypomp_amd import _native
        _native.check(_native.lib().mp_pfilter(
            self.md, self.ser.desc, _native.ptr(self.th), 0, _native.ptr(self.keys), None, self.R, self.J, 0.0,
            0, self.fid, _native.ptr(self.ll), None, None, None, None, _native.ptr(self.ws),
            self.nws, _native.stream_ptr()))

    def bytes_per_particle_step(self):
        return 16 * (5 + 1)

    valu_instr_per_pstep = VALU_INSTR_PER_PSTEP["seir"]

    @property
    def measured_traffic(self):
        full = self.J == 20000 and self.R == 256 and self.T == 260 and self.form == "batched"
        return MEASURED_TRAFFIC_SEIR if full else None


class MopGradSV:
    """Config: SV MOP(0.97) value + adjoint gradient (6 parameters), J=1000, T=1000,
    K=50 checkpoints, 512 replicates; batched (one launch) or plain (one filter per
    launch)."""

    def __init__(self, form="batched", reps=512, J=1000, K=50):
        import torch
        import pypomp_amd.functional as F
        from pypomp_amd import _native
        from pypomp_amd import random as prandom
        from pypomp_amd.models import SVModel
        from pypomp_amd.series import Series
        from tests.helpers import workloads
        self.dev = _native.require_device()
        w = workloads()
        ys = w["sv_seed4_ys"]
        model = SVModel()
        self.struct = F.PompStruct(model=model, units=[Series.build(ys, np.arange(1, len(ys) + 1.0), 0.0,
                                                                    nstep=1)])
        self.theta_est = model.to_est(w["sv_seed4_theta"])
        self.J, self.R, self.T, self.K, self.form = J, reps, len(ys), K, form
        self.ser = self.struct.device_series()
        self.md = model.desc()
        keys = prandom.split(prandom.key(4), reps)
        self.keys = torch.from_numpy(np.ascontiguousarray(keys).view(np.int32)).to(self.dev)
        self.th = torch.from_numpy(np.ascontiguousarray(self.theta_est)).to(self.dev)
        self.ll = torch.empty(reps, dtype=torch.float64, device=self.dev)
        self.g = torch.empty((reps, 6), dtype=torch.float32, device=self.dev)
        self.fid = _native.MP_FORM_PLAIN if form == "plain" else _native.MP_FORM_BATCHED
        n = _native.lib().mp_mop_grad_workspace_size(self.md, self.ser.desc, 1 if form == "plain" else reps, J, K)
        self.ws = torch.empty(int(n), dtype=torch.uint8, device=self.dev)
        self.nws = int(n)

    @property
    def particle_steps(self):
        return self.J * self.T * self.R

    def step(self):
        from pypomp_amd import _native
        _native.check(_native.lib().mp_mop_grad(
            self.md, self.ser.desc, _native.ptr(self.th), 0, _native.ptr(self.keys), None, self.R, self.J, 0.97,
            self.K,
            self.fid, _native.ptr(self.ll), _native.ptr(self.g), _native.ptr(self.ws), self.nws,
            _native.stream_ptr()))

    def bytes_per_particle_step(self):
        return 3 * 16 * (2 + 1 + 1)  # forward, recompute, reverse; S = v, eta, weightsF

    @property
    def valu_instr_per_pstep(self):
        return VALU_INSTR_PER_PSTEP["mop_grad"]

    @property
    def measured_traffic(self):
        full = (self.form == "batched" and self.R == 512 and self.J == 1000 and self.T == 1000 and self.K == 50)
        return MEASURED_TRAFFIC_MOP if full else None


class PanelIF2:
    """Config: IF2 on a panel of 100 SEIR units of uneven length (T_u in
    [52, 520] weeks, seed-5 data, unit-specific beta), J=2000, 8 chains per unit,
    M=50 iterations, random-walk sd 0.02 on the estimation scale (S0f, E0, I0
    perturbed at t0 only), geometric cooling 0.5.  One step = the whole M=50
    run: 50 launches of 800 swarms, recorded as one HIP graph and replayed."""

    def __init__(self, J=2000, chains=8, M=50, graph=True):
        import torch
        import pypomp_amd as pp
        import pypomp_amd.functional as F
        from pypomp_amd import random as prandom
        from pypomp_amd.models import SEIRModel
        from pypomp_amd.series import Series
        from tests.helpers import workloads
        w = workloads()
        off, beta, ys = w["panel_off"], w["panel_beta"], w["panel_ys"]
        U = len(beta)
        model = SEIRModel()
        units = [Series.build(ys[off[u]:off[u + 1]][:, None], np.arange(1, off[u + 1] - off[u] + 1.0), 0.0, nstep=7)
                 for u in range(U)]
        self.struct = F.PompStruct(model=model, units=units)
        names = model.param_names
        rw = pp.RWSigma({p: 0.02 for p in names}, init_names=["S0f", "E0", "I0"]).geometric_cooling(0.5)
        rw = rw.canonicalize(names)
        theta = np.array([pp.models.SEIR_DEFAULTS[p] for p in names], dtype=np.float32)
        th = np.repeat(theta[None], U, 0)
        th[:, 0] = beta
        th0 = np.repeat(model.to_est(th)[None, :, None, :], chains, 0)          # (chains, U, 1, d)
        th0 = np.ascontiguousarray(np.broadcast_to(th0, (chains, U, J, len(names))).reshape(chains * U, J, -1))
        unit = np.tile(np.arange(U, dtype=np.int32), chains)
        R = chains * U
        keys = prandom.split(prandom.key(5), chains)
        ik = np.stack([prandom.split(k, M + 1)[1:] for k in keys], axis=1)      # (M, chains, 2)
        self.mkeys = np.stack([[prandom.split(k, U) for k in row] for row in ik]).reshape(M, R, 2)
        self.run = F._PerRunner(self.struct, R, J, rw, 0.0, unit)
        self.t0 = torch.from_numpy(th0).to(self.run.dev)
        self.start = self.t0.clone()
        self.J, self.R, self.M, self.U = J, R, M, U
        self.sumT = int(off[-1])
        self.graph = None
        self.use_graph = graph

    @property
    def particle_steps(self):
        return self.J * self.M * (self.R // self.U) * self.sumT

    def step(self):
        if not self.use_graph:  # eager launches (for counter collection)
            self.start.copy_(self.t0)
            self.out = self.run.iterate(self.start, self.mkeys, graph=False)
        elif self.graph is None:
            self.out = self.run.iterate(self.start, self.mkeys, graph=True)
            self.graph = self.run.graph
        else:
            self.start.copy_(self.t0)
            self.graph.replay()

    def bytes_per_particle_step(self):
        return 16 * (5 + 10 + 1)

    valu_instr_per_pstep = VALU_INSTR_PER_PSTEP["panel_if2"]

    @property
    def measured_traffic(self):
        return MEASURED_TRAFFIC_PANEL if (self.J == 2000 and self.R == 800) else None


def time_workload(wl, steps, warmup):
    import torch
    from pypomp_amd import _native
    for _ in range(warmup):
        wl.step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        wl.step()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    _native.profile_enable(True)
    wl.step()
    torch.cuda.synchronize()
    prof = _native.profile_read()
    _native.profile_enable(False)
    ms, n = prof["batched"]
    if n == 0 and prof["plain_prop"][1] > 0:  # plain form: propagate + scan per observation
        ms, n = prof["plain_prop"][0] + prof["plain_scan"][0], prof["plain_prop"][1]
    if n == 0:  # launches inside a replayed graph are not event-bracketed
        ms, n = el / steps * 1e3, getattr(wl, "M", 1)
    avg = ms / max(n, 1)
    gbps = wl.particle_steps * wl.bytes_per_particle_step() / (ms * 1e-3) / 1e9 if ms > 0 else None
    return {"value": wl.particle_steps * steps / el, "unit": "particle-steps/s", "seconds_per_step": el / steps,
            "kernel_avg_ms": avg, "launches_per_step": n,
            "roofline": {"bound": "hbm", "achieved": gbps, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": (gbps / HBM_PEAK_GBS) if gbps else None,
                         "traffic": getattr(wl, "measured_traffic", None),
                         "bytes_per_particle_step": wl.bytes_per_particle_step(),
                         "valu": valu_rate(getattr(wl, "valu_instr_per_pstep", None), wl.particle_steps,
                                           ms * 1e-3)}}


def kalman_loglik(seed):
    """Exact FP64 Kalman log-likelihood of the LG series of this seed (a fixture
    written with the data, tests/golden/kalman_loglik.json)."""
    path = Path(__file__).resolve().parent / "tests" / "golden" / "kalman_loglik.json"
    return json.loads(path.read_text())[f"lg2_seed{seed}"]


def host_info():
    """The CPU the baselines ran on: model name (lscpu's), logical CPUs, the ones
    this process may use, and the OpenMP threads the C restatement starts."""
    from oracle import cpu
    model = None
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"model": model, "nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)),
            "omp_threads": cpu.num_threads()}


def median_runs(run, runs=5):
    """Median of ``runs`` timed calls of ``run`` (after one untimed warm-up)."""
    run()
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        run()
        times.append(time.perf_counter() - t0)
    return float(np.median(times)), times


def cpu_baseline(target_s=2.5):
    """C restatement of the bootstrap filter (oracle/cpu_pf.c, OpenMP, one replicate
    per thread as the reference's CPU sharding) on a bounded sample of the workload:
    as many J=1e6 replicates as threads, T calibrated to ~target_s seconds per run;
    the median of 5 runs after a warm-up."""
    from oracle import cpu
    from oracle import threefry as tf
    st, theta, ys = lg2_struct(2)
    cores = cpu.num_threads()
    J = 1_000_000
    keys = tf.split(tf.key(2), cores)
    t0 = time.perf_counter()
    cpu.lg_pfilter(theta, 2, ys[:2], J, keys)
    per_step = (time.perf_counter() - t0) / 2
    T = int(max(2, min(ys.shape[0], target_s / max(per_step, 1e-6))))
    dt, times = median_runs(lambda: cpu.lg_pfilter(theta, 2, ys[:T], J, keys))
    return {"value": cores * J * T / dt, "unit": "particle-steps/s", "cores": cores, "kind": "port",
            "sample": f"C restatement (oracle/cpu_pf.c, -O3 -march=x86-64-v3, OpenMP), LG 2-D pfilter, "
                      f"{cores} replicates x J={J} x T={T}, median of 5 runs ({dt:.2f} s each)",
            "runs_s": times, "host": host_info()}


def cpu_mop_grad_baseline(J=1000, T=1000):
    """C restatement of the MOP(0.97) value and hand adjoint gradient
    (oracle/cpu_pf.c cpu_sv_mop_grad, pinned against the float64 autograd
    oracle), OpenMP with one replicate per thread, two replicates per thread;
    the median of 5 runs after a warm-up."""
    from oracle import cpu
    from oracle import sv as osv
    from oracle import threefry as tf
    from tests.helpers import workloads
    ys = workloads()["sv_seed4_ys"][:T]
    th = osv.SVModel().to_est(osv.default_theta())
    cores = cpu.num_threads()
    reps = 2 * cores
    keys = tf.split(tf.key(4), reps)
    dt, times = median_runs(lambda: cpu.sv_mop_grad(th, ys, J, 0.97, keys))
    return {"value": reps * J * T / dt, "unit": "particle-steps/s", "cores": cores, "kind": "port",
            "sample": f"C restatement with the hand adjoint (oracle/cpu_pf.c, -O3 -march=x86-64-v3, OpenMP), "
                      f"MOP(0.97) value + gradient, SV, J={J}, T={T}, {reps} replicates, median of 5 runs "
                      f"({dt:.2f} s each)", "runs_s": times, "host": host_info()}


# VALU instructions one wave issues per probe iteration (rocprofv3 SQ_INSTS_VALU
# / waves / iterations, profiles/valu_peak_calibration.md): kind 0 FMA stream,
# 1 threefry integer mix, 2 FMA + exp2.
PROBE_VALU_PER_ITER = {0: 8.0, 1: 24.0, 2: 16.0, 3: 23.0, 4: 15.0}
PROBE_KINDS = {0: "fma", 1: "int_threefry", 2: "fma_exp2", 3: "fma_int", 4: "int_add_xor"}


def valu_probe(iters=100_000, blocks=2048):
    """The VALU issue calibration kernels (mp_valu_probe) timed with HIP events:
    instructions per second from the per-iteration counts the counters measured."""
    import torch
    from pypomp_amd import _native
    out = torch.empty(blocks * 256, dtype=torch.float32, device="cuda")
    res = {}
    for kind, name in PROBE_KINDS.items():
        _native.check(_native.lib().mp_valu_probe(kind, iters, blocks, _native.ptr(out), _native.stream_ptr()))
        torch.cuda.synchronize()
        _native.profile_enable(True)
        _native.check(_native.lib().mp_valu_probe(kind, iters, blocks, _native.ptr(out), _native.stream_ptr()))
        torch.cuda.synchronize()
        ms, n = _native.profile_read()["valu_probe"]
        _native.profile_enable(False)
        waves = blocks * 256 // 64
        g = PROBE_VALU_PER_ITER[kind] * iters * waves / (ms * 1e-3) / 1e9
        res[name] = {"ms": ms, "Ginstr_per_s": g, "per_cu_cycle": g / (256 * 2.4)}
    return res


def run_only(args):
    """One workload, --warmup untimed + --steps timed passes, nothing else (the
    command profilers wrap).  ``copy`` is a 1 GiB device-to-device copy (the
    calibration kernel for the FETCH_SIZE / WRITE_SIZE counters)."""
    import torch
    if args.only == "copy":
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        step = lambda: dst.copy_(src)  # noqa: E731
        work, unit = src.numel() * 4, "bytes copied"
    elif args.only == "valu_peak":
        print(json.dumps({"only": "valu_peak", "probe": valu_probe()}), flush=True)
        return
    else:
        wl = {"plain": lambda: PlainLG(reps=args.reps), "seir": BatchedSEIR, "mop_batched": MopGradSV,
              "mop_plain": lambda: MopGradSV("plain", reps=16), "panel": PanelIF2,
              "seir_plain": lambda: BatchedSEIR(reps=32, form="plain"),
              "panel_eager": lambda: PanelIF2(M=args.reps if args.reps < 50 else 50, graph=False)}[args.only]()
        step, work, unit = wl.step, wl.particle_steps, "particle-steps"
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    print(json.dumps({"only": args.only, "steps": args.steps, "seconds": el, "rate": work * args.steps / el,
                      "unit": unit + "/s"}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--reps", type=int, default=64)
    ap.add_argument("--no-extra", action="store_true", help="skip the batched workloads")
    ap.add_argument("--lanes", type=int, default=None,
                    help="plain-form filters run at once, each on its own stream (default: the library's, 2)")
    ap.add_argument("--only", choices=["plain", "seir", "seir_plain", "mop_batched", "mop_plain", "panel",
                                       "panel_eager", "copy", "valu_peak"],
                    help="run one workload --steps times and print a short JSON line (for profilers)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print(json.dumps({"error": "the engine runs on one GPU; --gpus must be 1 (replicate filters shard "
                                   "across GPUs with no communication, see the design notes)"}))
        sys.exit(2)

    import torch
    from pypomp_amd import _native, autotune
    torch.cuda.set_device(0)
    if args.lanes is not None:
        autotune.set("plain_lanes", args.lanes)
    if args.only:
        return run_only(args)
    log("VALU issue calibration probes")
    calib = valu_probe()
    VALU_PEAK.update(best=max(v["Ginstr_per_s"] for v in calib.values()),
                     threefry=calib["int_threefry"]["Ginstr_per_s"],
                     source="mp_valu_probe in this run: best of 5 instruction mixes; threefry mix")
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
    lanes = autotune.get("plain_lanes")
    # the same workload with the filters strictly one after another (one lane;
    # every launch carries one filter in both cases)
    seq = None
    if lanes > 1:
        autotune.set("plain_lanes", 1)
        try:
            wl.step()
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            for i in range(args.steps):
                wl.step()
            torch.cuda.synchronize()
            el1 = time.perf_counter() - t1
        finally:
            autotune.set("plain_lanes", lanes)
        seq = {"value": wl.particle_steps * args.steps / el1, "ms_per_step": el1 / args.steps * 1e3, "lanes": 1}
        log(f"one lane: {args.steps} steps in {el1:.3f} s")

    # Per-kernel durations from HIP events on each launch's dispatch: one extra
    # step with one lane, where every launch has the GPU to itself (the kernel
    # roofline), and one in the timed lanes, where the filters' launches overlap
    # (their durations then include the time they share the GPU).

    def profiled_step(nl):
        autotune.set("plain_lanes", nl)
        try:
            _native.profile_enable(True)
            wl.step()
            torch.cuda.synchronize()
            return _native.profile_read()
        finally:
            _native.profile_enable(False)
            autotune.set("plain_lanes", lanes)

    kb = wl.kernel_bytes()
    J, T, R = wl.J, wl.T, wl.R

    def kernel_table(prof):
        out = {}
        for k in ("plain_prop", "plain_scan"):
            ms, n = prof[k]
            avg = ms / max(n, 1)
            out[k] = {"avg_us": avg * 1e3, "launches": n, "bytes": kb[k],
                      "GBps": kb[k] / (avg * 1e-3) / 1e9 if avg > 0 else None,
                      "valu": valu_rate(VALU_INSTR_PER_PSTEP[k], wl.J, avg * 1e-3)}
        return out

    prof = profiled_step(1)
    kernels = kernel_table(prof)
    kernels_lanes = kernel_table(profiled_step(lanes)) if lanes > 1 else kernels
    dom = max(kernels, key=lambda k: prof[k][0])
    d = kernels[dom]
    # the whole filter step in the timed lanes: both kernels' algorithmic bytes
    # and VALU instructions per particle-step (of the launch shapes those lanes
    # take when the shapes are automatic) over the timed region
    fsteps = T * R * args.steps
    wide = lanes >= 3 and all(autotune.get(n) == 0 for n in (
        "plain_scan_rows", "plain_prop_rows", "plain_scan_threads", "plain_prop_threads"))
    sfx = "_wide" if wide else ""
    step_instr = VALU_INSTR_PER_PSTEP["plain_prop" + sfx] + VALU_INSTR_PER_PSTEP["plain_scan" + sfx]
    step_bytes = sum(kb.values())
    step_gbps = step_bytes * fsteps / elapsed / 1e9
    roofline = {"bound": "hbm", "achieved": d["GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": d["GBps"] / HBM_PEAK_GBS,
                "traffic": MEASURED_TRAFFIC_J1E6[dom] if J == 1_000_000 else None, "kernel": dom,
                "bytes_per_launch": d["bytes"], "avg_launch_us": d["avg_us"], "measured_with_lanes": 1,
                "valu": valu_rate(VALU_INSTR_PER_PSTEP[dom], J, d["avg_us"] * 1e-6),
                "filter_step": {"lanes": lanes, "us_per_filter_step": elapsed / fsteps * 1e6,
                                "bytes_per_filter_step": step_bytes, "achieved": step_gbps,
                                "traffic_per_filter_step": (MEASURED_TRAFFIC_J1E6["plain_prop" + sfx]
                                                            + MEASURED_TRAFFIC_J1E6["plain_scan" + sfx])
                                if J == 1_000_000 else None,
                                "frac": step_gbps / HBM_PEAK_GBS,
                                "valu": valu_rate(step_instr, J * fsteps, elapsed)}}
    log(f"kernels (one lane) {json.dumps(kernels)}")
    log(f"kernels ({lanes} lanes) {json.dumps(kernels_lanes)}")

    exact = kalman_loglik(2)
    rel = abs(float(ll.mean()) - exact) / abs(exact)
    del wl
    torch.cuda.empty_cache()
    extra = {}
    if not args.no_extra:
        log("SEIR batched")
        sb = time_workload(BatchedSEIR(), 1, 1)
        log(f"SEIR plain (32 of the filters, {lanes} at a time)")
        sp = time_workload(BatchedSEIR(reps=32, form="plain"), 1, 0)
        log("SEIR plain (32 of the filters, one at a time)")
        autotune.set("plain_lanes", 1)
        try:
            sp1 = time_workload(BatchedSEIR(reps=32, form="plain"), 1, 0)
        finally:
            autotune.set("plain_lanes", lanes)
        sb["vs_plain"] = sb["value"] / sp["value"]
        sb["vs_plain_one_lane"] = sb["value"] / sp1["value"]
        sp["lanes"], sp1["lanes"] = lanes, 1
        extra["seir_pfilter_batched"] = sb
        extra["seir_pfilter_plain"] = sp
        extra["seir_pfilter_plain_one_lane"] = sp1
        log("MOP grad batched")
        mb = time_workload(MopGradSV("batched"), 2, 1)
        log("MOP grad plain (16 filters, one launch each)")
        mp_ = time_workload(MopGradSV("plain", reps=16), 1, 0)
        mb["vs_plain"] = mb["value"] / mp_["value"]
        extra["mop_grad_batched"] = mb
        extra["mop_grad_plain"] = mp_
        log("panel IF2 (100 SEIR units, 8 chains, J=2000, M=50, one HIP graph)")
        extra["panel_if2_graph"] = time_workload(PanelIF2(), 1, 1)

    cpu = None if args.no_cpu_baseline else cpu_baseline()
    if cpu is not None and not args.no_extra:
        cg = cpu_mop_grad_baseline()
        extra["mop_grad_cpu_baseline"] = cg
        extra["mop_grad_batched"]["vs_cpu"] = extra["mop_grad_batched"]["value"] / cg["value"]
    out = {
        "metric": METRIC, "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"pfilter LG 2-D, J=1e6, T=200, 64 replicates (plain form: one filter per "
                               f"launch, {lanes} in flight on {lanes} streams)",
                   "J": J, "T": T, "replicates": R, "form": "plain", "lanes": lanes},
        # the same 64 filters strictly one after another
        "one_filter_at_a_time": seq,
        "roofline": roofline, "cpu_baseline": cpu,
        # Monte Carlo error of the replicates' mean log-likelihood against the exact
        # FP64 Kalman value (an accuracy check, not parity with the reference)
        "loglik_mc_rel_err_vs_kalman": rel,
        "loglik": {"mean": float(ll.mean()), "sd": float(ll.std(ddof=1)), "kalman_fp64": exact},
        "kernels": kernels,
        "kernels_in_lanes": kernels_lanes,
        "valu_calibration": calib,
        "extra": extra,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
