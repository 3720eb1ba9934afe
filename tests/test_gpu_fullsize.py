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
