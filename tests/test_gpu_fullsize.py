"""GPU tests at the benchmarked workloads' full sizes, through properties that do
not depend on the size: the exact Kalman likelihood within Monte Carlo error,
the per-observation terms summing to the log-likelihood, equal results from
the plain and batched execution forms, and a MOP gradient that does not depend
on the checkpoint interval.  (Bit-level parity with the oracle is tested at
sizes the oracle finishes in seconds.)"""

import json
from pathlib import Path

import numpy as np
import pytest

from oracle import threefry as tf
from tests.helpers import workloads

pytestmark = pytest.mark.gpu


def lg2_seed2():
    import pypomp_amd.functional as F
    from pypomp_amd.models import LGModel
    from pypomp_amd.series import Series
    w = workloads()
    ys, theta = w["lg2_seed2_ys"], w["lg2_seed2_theta"]
    struct = F.PompStruct(model=LGModel(2, 2, separate_init=True),
                          units=[Series.build(ys, np.arange(1, ys.shape[0] + 1.0), 0.0, nstep=1)])
    return struct, theta


def test_lg2_million_particles_kalman_and_cll_sum():
    # LG 2-D, J = 1e6, T = 200 (seed-2 series), plain form, four replicates
    import pypomp_amd.functional as F
    struct, theta = lg2_seed2()
    reps = 4
    out = F.pfilter(struct, theta[None], 1_000_000, tf.split(tf.key(2024), reps).reshape(1, reps, 2),
                    CLL=True, form="plain")
    ll = out["logLik64"].ravel()
    path = Path(__file__).resolve().parent / "golden" / "kalman_loglik.json"
    exact = json.loads(path.read_text())["lg2_seed2"]
    # the sd of one replicate's log-likelihood at J = 1e6 is about 0.03
    assert np.all(np.abs(ll - exact) < 0.2), (ll, exact)
    assert abs(ll.mean() - exact) < 0.1, (ll.mean(), exact)
    # the float32 per-observation terms add up to the float64 log-likelihood
    np.testing.assert_allclose(out["CLL"].astype(np.float64).sum(axis=-1).ravel(), ll, rtol=2e-6)


def test_lg2_million_particles_plain_equals_batched():
    # one filter spread over the GPU and the same filter in one workgroup take
    # the same ancestors and give the same float64 log-likelihood
    import pypomp_amd.functional as F
    struct, theta = lg2_seed2()
    keys = tf.split(tf.key(7), 1).reshape(1, 1, 2)
    plain = F.pfilter(struct, theta[None], 1_000_000, keys, form="plain")
    batched = F.pfilter(struct, theta[None], 1_000_000, keys, form="batched")
    np.testing.assert_array_equal(plain["logLik64"], batched["logLik64"])


def test_seir_full_size_plain_equals_batched():
    # SEIR, J = 20000, T = 260 (seed-3 series, 7 Euler-binomial sub-steps per week)
    import pypomp_amd.functional as F
    from pypomp_amd.models import SEIRModel
    from pypomp_amd.series import Series
    w = workloads()
    ys, theta = w["seir_seed3_ys"], w["seir_seed3_theta"]
    struct = F.PompStruct(model=SEIRModel(), units=[Series.build(ys, np.arange(1, len(ys) + 1.0), 0.0, nstep=7)])
    keys = tf.split(tf.key(3), 2).reshape(1, 2, 2)
    plain = F.pfilter(struct, theta[None], 20000, keys, form="plain")
    batched = F.pfilter(struct, theta[None], 20000, keys, form="batched")
    np.testing.assert_array_equal(plain["logLik64"], batched["logLik64"])
    assert np.all(np.isfinite(plain["logLik64"]))


def test_sv_mop_gradient_full_size_checkpoint_independent():
    # SV MOP(0.97) value and gradient, J = 1000, T = 1000, 512 filters: the
    # reverse sweep recomputes K-step blocks from checkpoints, and the result
    # does not depend on K
    import pypomp_amd.functional as F
    from pypomp_amd.models import SVModel
    from pypomp_amd.series import Series
    w = workloads()
    ys = w["sv_seed4_ys"]
    model = SVModel()
    struct = F.PompStruct(model=model, units=[Series.build(ys, np.arange(1, len(ys) + 1.0), 0.0, nstep=1)])
    R = 512
    th = np.repeat(model.to_est(w["sv_seed4_theta"])[None], R, 0)
    keys = tf.split(tf.key(4), R)
    ll50, g50 = F.mop_grad(struct, th, 1000, 0.97, keys, K=50)
    ll250, g250 = F.mop_grad(struct, th, 1000, 0.97, keys, K=250)
    np.testing.assert_array_equal(ll50, ll250)
    np.testing.assert_array_equal(g50, g250)
    assert np.all(np.isfinite(g50)) and np.all(np.isfinite(ll50))
    # replicates differ only by their keys: the estimates scatter around one value
    assert ll50.std() < 0.05 * abs(ll50.mean())


def test_sv_mop_gradient_full_size_matches_autograd_oracle():
    # the MOP(0.97) adjoint at the benchmarked size (T = 1000, J = 1000, 512
    # filters in one launch): two of the rows against torch autograd of the
    # oracle's estimator on the same ancestry; the value is bit-exact
    import pypomp_amd.functional as F
    from oracle import mop_grad as mg
    from pypomp_amd.models import SVModel
    from pypomp_amd.series import Series
    w = workloads()
    ys = w["sv_seed4_ys"]
    model = SVModel()
    struct = F.PompStruct(model=model, units=[Series.build(ys, np.arange(1, len(ys) + 1.0), 0.0, nstep=1)])
    R = 512
    th = model.to_est(w["sv_seed4_theta"])
    keys = tf.split(tf.key(4), R)
    ll, g = F.mop_grad(struct, np.repeat(th[None], R, 0), 1000, 0.97, keys, K=50)
    for r in (0, 311):
        print(f"sv full-size gradient: row {r}", flush=True)
        val, _, gref, _ = mg.sv_mop_reference_grad(th, ys, keys[r], 1000, 0.97)
        assert ll[r] == val
        scale = np.abs(gref).max()
        np.testing.assert_allclose(g[r], gref, rtol=1e-2, atol=1e-3 * scale)


def test_panel_if2_full_size_graph_equals_eager_and_units_are_independent():
    # the panel IF2 workload at full size (100 SEIR units of 52..520 weeks,
    # J = 2000, 8 chains) for two iterations: the HIP-graph run equals the
    # eager one bit for bit, and with no shared parameters each unit's swarm
    # and log-likelihoods equal IF2 on that unit alone with the unit's keys
    # (core/algorithms/panel_mif.py:113, 216-235)
    import pypomp_amd as pp
    import pypomp_amd.functional as F
    from pypomp_amd import random as prandom
    from pypomp_amd.models import SEIR_DEFAULTS, SEIRModel
    from pypomp_amd.series import Series
    w = workloads()
    off, beta, ys = w["panel_off"], w["panel_beta"], w["panel_ys"]
    U, J, chains, M = len(beta), 2000, 8, 2
    model = SEIRModel()
    names = model.param_names
    units = [Series.build(ys[off[u]:off[u + 1]][:, None], np.arange(1, off[u + 1] - off[u] + 1.0), 0.0, nstep=7)
             for u in range(U)]
    struct = F.PanelStruct(model=model, units=units, shared_param_names=[], unit_param_names=names)
    rw = pp.RWSigma({p: 0.02 for p in names}, init_names=["S0f", "E0", "I0"]).geometric_cooling(0.5)
    th = np.repeat(np.array([SEIR_DEFAULTS[p] for p in names], np.float32)[None], U, 0)
    th[:, 0] = beta
    un = np.ascontiguousarray(np.broadcast_to(th[None, None], (chains, J, U, len(names))))
    sh = np.zeros((chains, J, 0), np.float32)
    keys = prandom.split(prandom.key(5), chains)
    print("panel IF2 full size: eager", flush=True)
    eager = F.panel_mif(struct, sh, un, rw, M, J, keys, graph=False)
    print("panel IF2 full size: graph", flush=True)
    graph = F.panel_mif(struct, sh, un, rw, M, J, keys, graph=True)
    for a, b in zip(eager, graph):
        np.testing.assert_array_equal(a, b)
    st, ut, fs, fu = eager
    assert np.all(np.isfinite(ut[:, 1:, :, 0]))
    rwc = rw.canonicalize(names)
    ik = np.stack([prandom.split(k, M + 1)[1:] for k in keys], axis=1)   # (M, chains, 2)
    for r, u in ((0, 0), (3, 57), (7, U - 1)):
        print(f"panel IF2 full size: chain {r} unit {u} alone", flush=True)
        one = F.PompStruct(model=model, units=[units[u]])
        sw = model.to_est(un[r][:, u, :])
        for m in range(M):
            k = prandom.split(ik[m, r], U)[u]
            out, _, ll = F.perfilter(one, sw[None], k[None], J, rwc, m)
            sw = out[0]
            assert np.float32(ll[0]) == ut[r, m + 1, u, 0], (r, u, m)
        np.testing.assert_array_equal(model.from_est(sw), fu[r][:, u, :])
