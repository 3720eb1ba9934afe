"""GPU parity of covariate-driven filtering and of the @vectorized rproc key
mode against the oracle, on the reference's S&P 500 model (pypomp/models/
spx.py) and data: covariates reach rinit, every Euler sub-step and dmeas at
the reference's grid indices (core/model_mechanics.py:894-906,
core/algorithms/pfilter.py:28-35, 124-128), and the sub-step key is either
split per particle or used whole (model_mechanics.py:553-588)."""

import numpy as np
import pytest

from oracle import filters as fl
from oracle import spx as ospx
from oracle import threefry as tf

pytestmark = pytest.mark.gpu
F32 = np.float32


def spx_data(T):
    d = np.load("tests/golden/spx.npz")
    return d["y"][:T], d["time"][:T]


def structs(T, vectorized, nstep=1, coarse=False):
    """(product struct, oracle model, oracle series) over the first T returns.
    ``coarse``: covariates on a coarser time table than the observations, so
    the sub-steps read linearly interpolated values."""
    import pypomp_amd.functional as F
    from pypomp_amd.models import SPXModel
    from pypomp_amd.series import Series
    ys, times = spx_data(T)
    ct, cv = ospx.covariate_table(ys, times)
    if coarse:
        ct, cv = ct[::3], cv[::3]
    ser = Series.build(ys, times, 0.0, nstep=nstep, ctimes=ct, covars=cv)
    oser = fl.make_series(ys, times, 0.0, nstep=nstep, ctimes=ct, covars=cv)
    np.testing.assert_array_equal(ser.covars, oser.covars)
    return F.PompStruct(model=SPXModel(vectorized=vectorized), units=[ser]), ospx.SPXModel(vectorized), oser


@pytest.mark.parametrize("vectorized", [True, False])
@pytest.mark.parametrize("form", ["batched", "plain"])
def test_spx_pfilter_matches_oracle(vectorized, form):
    import pypomp_amd.functional as F
    struct, model, oser = structs(60, vectorized)
    th = ospx.default_theta()
    keys = tf.split(tf.key(31), 2)
    print(f"spx pfilter: vectorized={vectorized} form={form}", flush=True)
    out = F.pfilter(struct, th[None], 400, keys.reshape(1, 2, 2), CLL=True, form=form)
    for r in range(2):
        o = fl.pfilter(model, th, oser, keys[r], 400, diagnostics=True)
        assert np.isfinite(o["logLik64"])
        assert float(out["logLik64"][0, r]) == o["logLik64"], (r, out["logLik64"][0, r], o["logLik64"])
        np.testing.assert_array_equal(out["CLL"][0, r], np.array(o["CLL"], dtype=F32))


def test_key_modes_differ_and_forms_agree():
    import pypomp_amd.functional as F
    th = ospx.default_theta()
    keys = tf.split(tf.key(32), 3).reshape(1, 3, 2)
    ll = {}
    for v in (True, False):
        struct, _, _ = structs(200, v)
        a = F.pfilter(struct, th[None], 5000, keys, form="batched")["logLik64"]
        b = F.pfilter(struct, th[None], 5000, keys, form="plain")["logLik64"]
        np.testing.assert_array_equal(a, b)
        ll[v] = a
    assert not np.array_equal(ll[True], ll[False])
    # two estimates of one likelihood: within Monte Carlo error of each other
    assert abs(float(ll[True].mean()) - float(ll[False].mean())) < 3.0


@pytest.mark.parametrize("vectorized", [True, False])
def test_spx_substeps_read_interpolated_covariates(vectorized):
    # four Euler sub-steps per interval on a covariate table three times
    # coarser than the data: rows tidx0 + k and tidx0 + nstep of the grid
    import pypomp_amd.functional as F
    struct, model, oser = structs(40, vectorized, nstep=4, coarse=True)
    assert oser.covars.shape == (161, 1)
    th = ospx.default_theta()
    keys = tf.split(tf.key(33), 2)
    for form in ("batched", "plain"):
        out = F.pfilter(struct, th[None], 300, keys.reshape(1, 2, 2), CLL=True, form=form)
        for r in range(2):
            o = fl.pfilter(model, th, oser, keys[r], 300, diagnostics=True)
            assert float(out["logLik64"][0, r]) == o["logLik64"], (form, r)
            np.testing.assert_array_equal(out["CLL"][0, r], np.array(o["CLL"], dtype=F32))


@pytest.mark.parametrize("vectorized", [True, False])
def test_spx_mop_matches_oracle(vectorized):
    import pypomp_amd.functional as F
    struct, model, oser = structs(50, vectorized)
    th_est = model.to_est(ospx.default_theta())
    keys = tf.split(tf.key(34), 2)
    got = F.mop(struct, np.stack([th_est, th_est]), 300, 0.9, keys, return_float64=True)
    for r in range(2):
        assert -got[r] == fl.mop(model, th_est, oser, keys[r], 300, 0.9)


@pytest.mark.parametrize("vectorized", [True, False])
def test_spx_perfilter_matches_oracle(vectorized):
    # IF2 perturbs every particle's parameters: the @vectorized step sees them
    # per particle (model_mechanics.py:553-568, the batched-theta path)
    import pypomp_amd as pp
    import pypomp_amd.functional as F
    struct, model, oser = structs(40, vectorized)
    J = 200
    rng = np.random.default_rng(8)
    th = np.repeat(ospx.default_theta()[None], J, 0)
    th_est = (model.to_est(th) + F32(0.02) * rng.standard_normal(th.shape).astype(F32)).astype(F32)
    rw = pp.RWSigma({n: 0.02 for n in ospx.PARAM_NAMES}, init_names=["V_0"]).geometric_cooling(0.5)
    k = tf.key(35)
    out, mean, ll, anc = F.perfilter(struct, th_est[None], k[None], J, rw, 1, ancestry=True)
    init = np.array([n == "V_0" for n in ospx.PARAM_NAMES])
    orw = fl.RWSigma(np.full(6, 0.02), init.astype(np.float64), a=0.5)
    eout, ell, eanc = fl.perfilter(model, 1, th_est, oser, k, J, orw, return_ancestry=True)
    np.testing.assert_array_equal(anc[0], eanc)
    np.testing.assert_array_equal(out[0], eout)
    assert abs(float(ll[0]) - float(ell)) <= 1e-5 * abs(float(ell))


@pytest.mark.parametrize("vectorized", [True, False])
def test_spx_simulate_matches_oracle(vectorized):
    import pypomp_amd.functional as F
    struct, model, oser = structs(30, vectorized, nstep=2)
    th = ospx.default_theta()
    keys = tf.split(tf.key(36), 2)
    X, Y = F.simulate(struct, np.stack([th, th]), 16, keys)
    for r in range(2):
        eX, eY = fl.simulate(model, th, oser, keys[r], nsim=16)
        np.testing.assert_array_equal(X[r], np.transpose(eX, (1, 0, 2)))
        np.testing.assert_array_equal(Y[r], np.transpose(eY, (1, 0, 2)))


def test_spx_pomp_object_api_equals_the_functional_call():
    import pypomp_amd.functional as F
    from pypomp_amd import random as prandom
    from pypomp_amd.models import spx
    ys, times = spx_data(80)
    P = spx(ys, times, vectorized=True)
    P.pfilter(J=500, key=prandom.key(41), reps=2)
    res = P.results()
    new = tf.split(tf.key(41))[1]
    keys = tf.split(new, 2)
    direct = F.pfilter(P.to_struct(), P._theta_array(), 500, keys.reshape(1, 2, 2))
    np.testing.assert_array_equal(P.results_history[-1]["logLiks"].reshape(-1), direct["logLik"].reshape(-1))
    assert np.isfinite(float(res["logLik"].iloc[0]))
