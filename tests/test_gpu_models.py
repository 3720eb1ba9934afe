"""GPU parity of the SV (MOP) and SEIR models against the oracle at small sizes."""

import numpy as np
import pytest

from oracle import filters as fl
from oracle import seir as oseir, sv as osv, threefry as tf
from tests.helpers import workloads

pytestmark = pytest.mark.gpu
F32 = np.float32


def sv_struct(T=None):
    import pypomp_amd.functional as F
    from pypomp_amd.models import SVModel
    from pypomp_amd.series import Series
    w = workloads()
    ys = w["sv_seed4_ys"][:T]
    return F.PompStruct(model=SVModel(), units=[Series.build(ys, np.arange(1, len(ys) + 1.0), 0.0, nstep=1)]), \
        w["sv_seed4_theta"], ys


def seir_struct(T=None):
    import pypomp_amd.functional as F
    from pypomp_amd.models import SEIRModel
    from pypomp_amd.series import Series
    w = workloads()
    ys = w["seir_seed3_ys"][:T]
    return F.PompStruct(model=SEIRModel(), units=[Series.build(ys, np.arange(1, len(ys) + 1.0), 0.0, nstep=7)]), \
        w["seir_seed3_theta"], ys


@pytest.mark.parametrize("form", ["batched", "plain"])
def test_sv_mop_matches_oracle(form):
    import pypomp_amd.functional as F
    struct, theta, ys = sv_struct(100)
    model = osv.SVModel()
    th_est = model.to_est(theta)
    keys = tf.split(tf.key(41), 2)
    got = F.mop(struct, np.stack([th_est, th_est]), 500, 0.97, keys, form=form, return_float64=True)
    ser = fl.make_series(ys, np.arange(1, len(ys) + 1.0), 0.0, nstep=1)
    for r in range(2):
        exp = fl.mop(model, th_est, ser, keys[r], 500, 0.97)
        assert -got[r] == exp, (r, -got[r], exp)


@pytest.mark.parametrize("block", [768, 1024, 0])
def test_seir_workgroup_sizes_keep_the_bits(block):
    # filters of 2000 particles take 1024-thread workgroups by default (0);
    # 768 and 1024 give the same bits, and the oracle's
    import pypomp_amd.functional as F
    from pypomp_amd import autotune
    struct, theta, ys = seir_struct(12)
    keys = tf.split(tf.key(17), 2)
    saved = autotune.get("batched_block")
    try:
        autotune.set("batched_block", block)
        out = F.pfilter(struct, theta[None], 2000, keys.reshape(1, 2, 2), CLL=True)
        mop = F.mop(struct, np.stack([oseir.SEIRModel().to_est(theta)] * 2), 2000, 0.9, keys, return_float64=True)
    finally:
        autotune.set("batched_block", saved)
    model = oseir.SEIRModel()
    ser = fl.make_series(ys, np.arange(1, len(ys) + 1.0), 0.0, nstep=7)
    for r in range(2):
        o = fl.pfilter(model, theta, ser, keys[r], 2000, diagnostics=True)
        assert float(out["logLik64"][0, r]) == o["logLik64"], (block, r)
        np.testing.assert_array_equal(out["CLL"][0, r], np.array(o["CLL"], dtype=np.float32))
        assert -mop[r] == fl.mop(model, model.to_est(theta), ser, keys[r], 2000, 0.9), (block, r)


def test_seir_pfilter_matches_oracle():
    import pypomp_amd.functional as F
    struct, theta, ys = seir_struct(20)
    keys = tf.split(tf.key(13), 2)
    out = F.pfilter(struct, theta[None], 256, keys.reshape(1, 2, 2), CLL=True, prediction_mean=True)
    model = oseir.SEIRModel()
    ser = fl.make_series(ys, np.arange(1, len(ys) + 1.0), 0.0, nstep=7)
    for r in range(2):
        o = fl.pfilter(model, theta, ser, keys[r], 256, diagnostics=True)
        assert float(out["logLik64"][0, r]) == o["logLik64"]
        # integer counts follow the same trajectories
        np.testing.assert_allclose(out["prediction_mean"][0, r], np.array(o["prediction_mean"]), rtol=1e-6)


def test_seir_counts_are_conserved():
    import pypomp_amd.functional as F
    struct, theta, _ = seir_struct(30)
    out = F.pfilter(struct, theta[None], 1000, tf.split(tf.key(2), 1).reshape(1, 1, 2), prediction_mean=True)
    pm = out["prediction_mean"][0, 0]
    np.testing.assert_allclose(pm[:, :4].sum(axis=1), 1e6, rtol=1e-6)
