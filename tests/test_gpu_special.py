"""GPU parity of the filters' numerical special cases against the oracle:
the weight collapse of ``_normalize_weights`` (every log-weight -inf:
loglik term -inf, uniform normalised weights, core/algorithms/helpers.py:49-57)
and IF2's density substitution (``nan_to_num(dmeas, nan=log(1e-18))``, with
+-inf clipped to the largest float32, core/algorithms/mif.py:205-210)."""

import numpy as np
import pytest

from oracle import filters as fl
from oracle import lg as olg, threefry as tf
from tests.helpers import lg1_data, lg1_theta

pytestmark = pytest.mark.gpu
F32 = np.float32
LG_NAMES = ["A11", "C11", "Q11", "R11", "X0_1"]


def lg1_struct_with(ys):
    import pypomp_amd.functional as F
    from pypomp_amd.models import LGModel
    from pypomp_amd.series import Series
    return F.PompStruct(model=LGModel(1, 1), units=[Series.build(ys, np.arange(1, ys.shape[0] + 1.0), 0.0, nstep=1)])


def outlier_data(pos, value, T=6):
    ys = np.concatenate([lg1_data(0), lg1_data(1)])[:T].astype(F32)
    ys[pos, 0] = F32(value)
    return ys


@pytest.mark.parametrize("form", ["batched", "plain"])
@pytest.mark.parametrize("pos", [0, 2])
@pytest.mark.parametrize("thresh", [0.0, 0.5])
def test_weight_collapse_matches_oracle(form, pos, thresh):
    # y = 1e30 makes (y - x)^2 overflow: every particle's density is -inf at
    # observation pos, the filter collapses and recovers at the next step
    import pypomp_amd.functional as F
    ys = outlier_data(pos, 1e30)
    struct = lg1_struct_with(ys)
    J = 100
    keys = tf.split(tf.key(77), 2)
    print(f"collapse: form={form} pos={pos} thresh={thresh} J={J}", flush=True)
    flags = dict(CLL=True) if form == "plain" else dict(CLL=True, ESS=True, filter_mean=True)
    out = F.pfilter(struct, lg1_theta()[None], J, keys.reshape(1, 2, 2), thresh=thresh, form=form, **flags)
    model = olg.LGModel(olg.LGSpec(1, 1))
    ser = fl.make_series(ys, np.arange(1, ys.shape[0] + 1.0), 0.0, nstep=1)
    for r in range(2):
        o = fl.pfilter(model, lg1_theta(), ser, keys[r], J, thresh=thresh, diagnostics=True)
        assert o["CLL"][pos] == -np.inf and o["logLik64"] == -np.inf
        assert out["logLik64"][0, r] == -np.inf
        # the collapsed step and every step after it, bit for bit
        np.testing.assert_array_equal(out["CLL"][0, r], np.array(o["CLL"], dtype=F32))
        assert np.all(np.isfinite(out["CLL"][0, r][pos + 1:]))
        if form == "batched":
            # uniform weights after the collapse: ESS = J
            assert out["ESS"][0, r][pos] == pytest.approx(J, rel=1e-6)
            np.testing.assert_allclose(out["ESS"][0, r], np.array(o["ESS"]), rtol=1e-5)
            np.testing.assert_allclose(out["filter_mean"][0, r], np.array(o["filter_mean"]), rtol=1e-5, atol=1e-6)


def test_collapse_plain_equals_batched():
    import pypomp_amd.functional as F
    ys = outlier_data(3, -1e30, T=8)
    struct = lg1_struct_with(ys)
    keys = tf.split(tf.key(78), 3).reshape(1, 3, 2)
    a = F.pfilter(struct, lg1_theta()[None], 3000, keys, CLL=True, form="batched")
    b = F.pfilter(struct, lg1_theta()[None], 3000, keys, CLL=True, form="plain")
    np.testing.assert_array_equal(a["CLL"], b["CLL"])
    np.testing.assert_array_equal(a["logLik64"], b["logLik64"])


@pytest.mark.parametrize("bad", [np.nan, 1e30])
def test_perfilter_density_substitution_matches_oracle(bad):
    # NaN densities become log(1e-18), -inf ones the most negative float32: the
    # swarm, its ancestry and the log-likelihood follow the oracle
    import pypomp_amd as pp
    import pypomp_amd.functional as F
    ys = outlier_data(1, bad)
    ys[3, 0] = F32(np.nan)
    struct = lg1_struct_with(ys)
    J = 150
    rng = np.random.default_rng(5)
    th = (lg1_theta()[None] + F32(0.01) * rng.standard_normal((J, 5)).astype(F32)).astype(F32)
    model = olg.LGModel(olg.LGSpec(1, 1))
    th_est = model.to_est(th)
    rw = pp.RWSigma({n: 0.02 for n in LG_NAMES}).geometric_cooling(0.5)
    k = tf.key(19)
    print(f"perfilter substitution: bad={bad} J={J}", flush=True)
    out, mean, ll, anc = F.perfilter(struct, th_est[None], k[None], J, rw, 2, ancestry=True)
    orw = fl.RWSigma(np.full(5, 0.02), np.zeros(5), a=0.5)
    ser = fl.make_series(ys, np.arange(1, ys.shape[0] + 1.0), 0.0, nstep=1)
    eout, ell, eanc = fl.perfilter(model, 2, th_est, ser, k, J, orw, return_ancestry=True)
    np.testing.assert_array_equal(anc[0], eanc)
    np.testing.assert_array_equal(out[0], eout)
    assert np.isfinite(float(ell)) and np.isfinite(float(ll[0]))
    assert abs(float(ll[0]) - float(ell)) <= 1e-5 * abs(float(ell))
    if bad != bad:
        # two NaN observations contribute log(1e-18) each, up to the
        # normalisation: the log-likelihood is far below the clean one's
        assert float(ll[0]) < 2 * np.log(1e-18) + 10
