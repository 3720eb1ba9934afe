"""Host-side logic (no GPU): the Euler time grid, replicate aggregation,
random-walk sigmas, key derivation and argument errors -- cases as the
reference's own tests (tests/test_maths.py, tests/test_core/test_helpers.py)."""

import warnings

import numpy as np
import pytest

from oracle import threefry as tf
from pypomp_amd import RWSigma
from pypomp_amd.maths import logmeanexp, logmeanexp_se
from pypomp_amd.series import Series, calc_steps


def test_logmeanexp_matches_r_pomp():
    # R: pomp::logmeanexp(c(100,101,102,103,104), se=TRUE) -> 102.8424765, 0.7510094
    x = np.array([100.0, 101.0, 102.0, 103.0, 104.0])
    assert abs(logmeanexp(x) - 102.8424765) < 1e-6
    assert abs(logmeanexp_se(x) - 0.7510094) < 1e-6


@pytest.mark.parametrize("arr, lme_nan, se_nan", [
    (np.array([100.0]), False, True),
    (np.array([]), True, True),
    (np.array([np.nan, np.nan]), True, True),
    (np.array([100, 101, -np.inf]), False, False),
    (np.array([1, np.nan, -np.inf]), True, True),
    (np.array([42.0]), False, True),
])
def test_logmeanexp_edge_cases(arr, lme_nan, se_nan):
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        lme, se = logmeanexp(arr), logmeanexp_se(arr)
    assert np.isnan(lme) == lme_nan and (lme_nan or np.isfinite(lme))
    assert np.isnan(se) == se_nan and (se_nan or np.isfinite(se))


def test_logmeanexp_large_values():
    arr = np.array([1e10, 1e10 + 1, 1e10 + 2])
    assert np.isclose(logmeanexp(arr), 1e10 + logmeanexp(np.array([0.0, 1, 2])), atol=1e-5)


def test_euler_grid_dt_and_nstep():
    times0 = np.array([-1.0, 0, 0.5, 1, 1.5, 2, 2.5, 3, 3.5, 4, 4.5, 5])
    ns, dts = calc_steps(times0, dt=0.5)
    assert ns.tolist() == [2] + [1] * 10 and np.allclose(dts, [0.5] * 11)
    # the reference carries each interval's adjusted step into the next one
    # (helpers.py:205-214 rebinds dt): [0, 1] takes 4 steps of 0.25, so [1, 3]
    # takes 8 steps of 0.25, not ceil(2 / 0.3) = 7
    ns, dts = calc_steps(np.array([0.0, 1.0, 3.0]), dt=0.3)
    assert ns.tolist() == [4, 8] and np.allclose(dts, [0.25, 0.25])
    # a short first interval shrinks the step of every later one
    ns, dts = calc_steps(np.array([0.0, 0.1, 5.1]), dt=1.0)
    assert ns.tolist() == [1, 50] and np.allclose(dts, [0.1, 0.1])
    # a zero-length interval leaves a step of 0: the next positive interval
    # divides by zero, as in the reference
    with pytest.raises(ZeroDivisionError):
        calc_steps(np.array([0.0, 1.0, 1.0, 3.0]), dt=0.3)
    ns, dts = calc_steps(np.array([0.0, 1.0, 1.0]), dt=0.3)
    assert ns.tolist() == [4, 0] and np.allclose(dts, [0.25, 0.0])
    # the reference's own grid case (tests/test_core/test_helpers.py:10-31)
    ns, dts = calc_steps(np.array([-1.0] + [0.5 * i for i in range(11)]), dt=0.5)
    assert ns.tolist() == [2] + [1] * 10 and np.allclose(np.repeat(dts, ns), 0.5)
    ns, dts = calc_steps(np.array([0.0, 1.0, 3.0]), nstep=7)
    assert ns.tolist() == [7, 7] and np.allclose(dts, [1 / 7, 2 / 7])
    with pytest.raises(ValueError):
        calc_steps(times0, dt=0.5, nstep=2)
    with pytest.raises(ValueError):
        calc_steps(times0)
    s = Series.build(np.zeros((3, 1)), [1.0, 2.0, 4.0], 0.0, dt=0.5)
    assert s.nstep.tolist() == [2, 2, 4] and s.dt_ext.shape == (8,)
    assert np.isclose(s.dt_ext.astype(np.float64).sum(), 4.0)


@pytest.mark.parametrize("times,t0,dt", [([1.0, 3.0], 0.0, 0.3), ([0.1, 5.1], 0.0, 1.0),
                                         ([0.7, 1.9, 2.0, 7.3], 0.0, 0.45), ([1.0, 1.0], 0.0, 0.3)])
def test_oracle_euler_grid_equals_the_host_grid(times, t0, dt):
    from oracle import filters as fl
    s = Series.build(np.zeros((len(times), 1)), times, t0, dt=dt)
    o = fl.make_series(np.zeros((len(times), 1)), times, t0, dt=dt)
    assert s.nstep.tolist() == o.nstep.tolist()
    np.testing.assert_array_equal(s.dt_ext, o.dt_ext)


def test_rwsigma_validation_and_order():
    rw = RWSigma({"a": 0.1, "b": 0.2, "c": 0.3}, init_names=["c"]).geometric_cooling(0.5)
    c = rw.canonicalize(["c", "a", "b"])
    assert np.allclose(c.sigmas_array, [0, 0.1, 0.2]) and np.allclose(c.sigmas_init_array, [0.3, 0, 0])
    assert np.float32(c.factor) == np.float32(0.5 ** (1 / 50))
    # cooling factor at nt = 0 of iteration m is factor^m
    assert abs(float(c.cooling_factor(0, 3, 10)) - 0.5 ** (3 / 50)) < 1e-6
    for bad in ({"a": -1.0}, {"a": True}, {1: 0.1}):
        with pytest.raises(ValueError):
            RWSigma(bad)
    with pytest.raises(ValueError):
        RWSigma({"a": 0.1}, init_names=["b"])
    with pytest.raises(ValueError):
        rw.canonicalize(["a", "b"])
    with pytest.raises(ValueError):
        rw.geometric_cooling(1.5)


def test_host_key_derivation_matches_oracle():
    from pypomp_amd import random as prandom
    k = prandom.key(20260811)
    np.testing.assert_array_equal(np.asarray(k, np.uint32), tf.key(20260811))
    np.testing.assert_array_equal(prandom.split(k, 5), tf.split(tf.key(20260811), 5))


def test_pomp_argument_errors():
    import pypomp_amd as pp
    from pypomp_amd.models import LG
    P = LG(ys=np.zeros((4, 2), np.float32))
    with pytest.raises(ValueError):
        P.pfilter(J=0, key=prandom_key())
    with pytest.raises(ValueError):
        P.mif(J=10, M=1, key=prandom_key(), rw_sd=pp.RWSigma({"nope": 0.1}))


def prandom_key():
    from pypomp_amd import random as prandom
    return prandom.key(1)


def test_host_from_est_equals_oracle_transform():
    # the host transform runs the device code on the CPU: bit-identical to the
    # oracle's deterministic restatement
    from oracle import lg as olg, seir as oseir, sv as osv
    from pypomp_amd.models import LGModel, SEIRModel, SVModel
    rng = np.random.default_rng(4)
    for hm, om in ((LGModel(2, 2, separate_init=True), olg.LGModel(olg.LGSpec(2, 2, separate_init=True))),
                   (LGModel(1, 1), olg.LGModel(olg.LGSpec(1, 1))), (SVModel(), osv.SVModel()),
                   (SEIRModel(), oseir.SEIRModel())):
        x = rng.normal(scale=2.0, size=(64, hm.d)).astype(np.float32)
        np.testing.assert_array_equal(hm.from_est(x), om.from_est(x))
        assert hm.from_est(x[0]).shape == (hm.d,)


def test_pomp_pickles_with_its_key_and_history():
    # the reference pickles a Pomp with its key and results (core/pomp.py:615-826);
    # native models need no function bytes
    import pickle
    from pypomp_amd import random as prandom
    from pypomp_amd.models.lg import LG
    P = LG(A=np.array([[0.9]]), ys=np.arange(4, dtype=np.float32)[:, None] / 4)
    P.fresh_key = prandom.key(9)
    P.results_history.append({"method": "pfilter", "logLiks": np.zeros((1, 2)), "theta": P.theta})
    Q = pickle.loads(pickle.dumps(P))
    np.testing.assert_array_equal(Q.ys.to_numpy(), P.ys.to_numpy())
    np.testing.assert_array_equal(Q.fresh_key, P.fresh_key)
    assert Q.theta == P.theta and Q.canonical_param_names == P.canonical_param_names
    assert Q.results_history[0]["method"] == "pfilter"
    assert bytes(Q.model.desc()) == bytes(P.model.desc())


def _lg_pomp(a=0.9, key=9):
    from pypomp_amd import random as prandom
    from pypomp_amd.models.lg import LG
    P = LG(A=np.array([[a]]), ys=np.arange(4, dtype=np.float32)[:, None] / 4)
    P.fresh_key = prandom.key(key)
    return P


def _fake_pfilter_entry(P, seed, n_rows=1):
    # the shape of a pfilter entry; values are arbitrary host data
    rng = np.random.default_rng(seed)
    return {"method": "pfilter", "execution_time": float(seed), "key": np.array([0, seed], np.uint32),
            "theta": [dict(t) for t in P.theta] * n_rows, "logLiks": rng.normal(size=(n_rows, 3)),
            "logLik": rng.normal(size=n_rows), "J": 100, "reps": 3, "thresh": 0.0,
            "CLL": rng.normal(size=(n_rows, 3, 4)).astype(np.float32)}


def test_pomp_equality_paths():
    # core/pomp.py:455-540 and tests/test_pomp/test_pomp_eq_merge.py: every
    # field that defines a Pomp breaks equality on its own
    import copy
    import pickle
    from pypomp_amd.models.lg import LG
    P = _lg_pomp()
    P.results_history.append(_fake_pfilter_entry(P, 1))
    assert P == pickle.loads(pickle.dumps(P)) and P == copy.deepcopy(P)
    assert P != "not a pomp"
    Q = copy.deepcopy(P); Q.theta = [dict(Q.theta[0], A11=0.5)]
    assert P != Q
    Q = copy.deepcopy(P); Q.t0 = 10.0
    assert P != Q
    Q = copy.deepcopy(P); Q.fresh_key = np.array([0, 99], np.uint32)
    assert P != Q
    Q = copy.deepcopy(P); Q.fresh_key = None
    assert P != Q
    Q = copy.deepcopy(P); Q.results_history[0]["logLiks"] = Q.results_history[0]["logLiks"] + 1
    assert P != Q
    Q = copy.deepcopy(P); Q.series.nstep = Q.series.nstep + 1
    assert P != Q
    Q = LG(A=np.array([[0.9]]), ys=np.arange(5, dtype=np.float32)[:, None] / 4)
    Q.fresh_key, Q.results_history = P.fresh_key, P.results_history
    assert P != Q                                   # different data
    R = LG(ys=np.zeros((4, 2), np.float32))
    assert P != R and P.model != R.model            # different model sizes and names
    # NaN entries compare equal to themselves (IF2 traces start with a NaN logLik)
    Q = copy.deepcopy(P); P.results_history[0]["logLiks"][0, 0] = np.nan
    Q.results_history[0]["logLiks"][0, 0] = np.nan
    assert P == Q


def test_pomp_merge_joins_parameter_sets_and_results():
    # core/pomp.py:542-613, core/results/result.py:178-213
    P, Q = _lg_pomp(0.9, 1), _lg_pomp(0.5, 2)
    P.results_history.append(_fake_pfilter_entry(P, 1))
    Q.results_history.append(_fake_pfilter_entry(Q, 2, n_rows=2))
    M = type(P).merge(P, Q)
    assert [t["A11"] for t in M.theta] == [P.theta[0]["A11"], Q.theta[0]["A11"]]
    e = M.results_history[0]
    np.testing.assert_array_equal(e["logLiks"], np.concatenate([P.results_history[0]["logLiks"],
                                                                Q.results_history[0]["logLiks"]]))
    assert e["CLL"].shape == (3, 3, 4) and len(e["theta"]) == 3 and e["execution_time"] == 2.0
    np.testing.assert_array_equal(e["key"], P.results_history[0]["key"])
    np.testing.assert_array_equal(M.fresh_key, P.fresh_key)
    assert list(M.results()["theta_idx"]) == [0, 1, 2]
    # the inputs are unchanged
    assert len(P.theta) == 1 and P.results_history[0]["logLiks"].shape == (1, 3)


def test_pomp_merge_rejects_mismatches():
    import copy
    from pypomp_amd.models.lg import LG
    P = _lg_pomp()
    with pytest.raises(ValueError, match="At least one"):
        type(P).merge()
    with pytest.raises(TypeError):
        type(P).merge(P, "x")
    R = LG(ys=np.zeros((4, 2), np.float32))
    with pytest.raises(ValueError, match="canonical_param_names"):
        type(P).merge(P, R)
    Q = LG(A=np.array([[0.9]]), ys=np.arange(4, dtype=np.float32)[:, None] / 8)
    with pytest.raises(ValueError, match="same ys"):
        type(P).merge(P, Q)
    Q = copy.deepcopy(P); Q.t0 = 3.0
    with pytest.raises(ValueError, match="same t0"):
        type(P).merge(P, Q)
    Q = copy.deepcopy(P); Q.results_history.append(_fake_pfilter_entry(Q, 3))
    with pytest.raises(ValueError, match="same number of entries"):
        type(P).merge(P, Q)
    P.results_history.append(_fake_pfilter_entry(P, 4))
    Q.results_history[0]["J"] = 7
    with pytest.raises(ValueError, match="same J"):
        type(P).merge(P, Q)


def test_pomp_unpickle_with_bad_key_payload_warns():
    # core/pomp.py:658-680: a corrupted key payload warns and leaves no key
    P = _lg_pomp()
    state = P.__getstate__()
    np.testing.assert_array_equal(state["_fresh_key_data"], P.fresh_key)
    assert "fresh_key" not in state
    state["_fresh_key_data"] = np.zeros(3, np.uint32)
    Q = object.__new__(type(P))
    with pytest.warns(UserWarning, match="Failed to reconstruct fresh_key"):
        Q.__setstate__(state)
    assert Q.fresh_key is None and Q.theta == P.theta
    state.pop("_fresh_key_data")
    R = object.__new__(type(P))
    R.__setstate__(state)
    assert R.fresh_key is None


def test_panel_parameters_merge_and_equality():
    import pandas as pd
    import pypomp_amd as pp

    def params(v):
        return pp.PanelParameters([{"shared": pd.DataFrame({"shared": [v, 1.0]}, index=["A11", "C11"]),
                                    "unit_specific": pd.DataFrame({"u1": [0.1, 0.2, v], "u2": [0.1, 0.2, 0.0]},
                                                                  index=["Q11", "R11", "X0_1"])}])
    a, b = params(0.5), params(0.7)
    assert a == params(0.5) and a != b
    m = pp.PanelParameters.merge(a, b)
    sh, un = m.arrays(["u1", "u2"])
    np.testing.assert_allclose(sh[:, 0], [0.5, 0.7])
    np.testing.assert_allclose(un[:, 0, 2], [0.5, 0.7])
    other = pp.PanelParameters([{"shared": None, "unit_specific": pd.DataFrame(
        {"u1": [0.9, 1.0, 0.1, 0.2, 0.0], "u2": [0.9, 1.0, 0.1, 0.2, 0.0]},
        index=["A11", "C11", "Q11", "R11", "X0_1"])}])
    with pytest.raises(ValueError, match="same parameter and unit names"):
        pp.PanelParameters.merge(a, other)


def test_pomp_equality_ignores_execution_time():
    import copy
    P = _lg_pomp()
    P.results_history.append(_fake_pfilter_entry(P, 1))
    Q = copy.deepcopy(P)
    Q.results_history[0]["execution_time"] = 123.0
    assert P == Q


def test_pomp_missing_plugins_raise_the_reference_errors():
    # core/estimation_mixin.py:181, 365, 554, 1168 and
    # tests/test_pomp/test_pomp_validation.py:159-209: a Pomp built without dmeas
    # cannot filter, one without rmeas cannot simulate
    import pypomp_amd as pp
    P = _lg_pomp()
    m = P.model
    Q = pp.Pomp(ys=P.ys, theta=P.theta, t0=0.0, rinit=m.rinit, rproc=m.rproc, rmeas=m.rmeas, nstep=1)
    Q.fresh_key = P.fresh_key
    assert Q.dmeas is None and Q.rmeas is not None and P.dmeas is not None
    rw = pp.RWSigma({n: 0.01 for n in Q.canonical_param_names})
    eta = pp.LearningRate({n: 0.1 for n in Q.canonical_param_names})
    for call in (lambda: Q.pfilter(J=10), lambda: Q.mif(J=10, M=1, rw_sd=rw), lambda: Q.train(J=10, M=1, eta=eta),
                 lambda: Q.mop(J=10, alpha=0.9), lambda: Q.pmcmc(J=10, M=1, proposal=None)):
        with pytest.raises(ValueError, match="self.dmeas cannot be None"):
            call()
    R = pp.Pomp(ys=P.ys, theta=P.theta, t0=0.0, rinit=m.rinit, rproc=m.rproc, dmeas=m.dmeas, nstep=1)
    R.fresh_key = P.fresh_key
    with pytest.raises(ValueError, match="self.rmeas cannot be None"):
        R.simulate()
    assert P != R and R != Q
    with pytest.raises(ValueError, match="same dmeas"):
        type(P).merge(R, Q)
    with pytest.raises(ValueError, match="same rmeas"):
        type(P).merge(P, R)


def test_pomp_slot_rules_follow_the_reference():
    # core/pomp.py:207-208 (rinit and rproc are required arguments) and
    # :319-320 (at least one of dmeas or rmeas); a component goes in its own slot
    import pypomp_amd as pp
    P = _lg_pomp()
    m = P.model
    kw = dict(ys=P.ys, theta=P.theta, t0=0.0, nstep=1)
    with pytest.raises(ValueError, match="at least one of dmeas or rmeas"):
        pp.Pomp(rinit=m.rinit, rproc=m.rproc, **kw)
    with pytest.raises(TypeError, match="rinit and rproc"):
        pp.Pomp(rproc=m.rproc, dmeas=m.dmeas, **kw)
    with pytest.raises(TypeError, match="rinit and rproc"):
        pp.Pomp(rinit=m.rinit, dmeas=m.dmeas, **kw)
    with pytest.raises(ValueError, match="passed as dmeas"):
        pp.Pomp(rinit=m.rinit, rproc=m.rproc, dmeas=m.rmeas, **kw)
    with pytest.raises(ValueError, match="passed as rinit"):
        pp.Pomp(rinit=m.rproc, rproc=m.rproc, dmeas=m.dmeas, **kw)
    ok = pp.Pomp(rinit=m.rinit, rproc=m.rproc, dmeas=m.dmeas, **kw)
    assert ok.rinit is not None and ok.rproc is not None and ok.rmeas is None


@pytest.mark.parametrize("name", ["newton", "wnewton"])
def test_newton_steps_follow_the_reference_arithmetic(name):
    # core/optimizer.py:270-330 on identical inputs (torch on the CPU): direction
    # -pinv(H_w) grad, H_w = H at step 0 and w H_prev + (1 - w) H after, with
    # w = i^log i / (i + 1)^log(i + 1); singular values below 10 d eps32 of the
    # largest dropped
    import torch
    from oracle import mop_grad as mg
    from pypomp_amd.optim import Newton, WeightedNewton
    rng = np.random.default_rng(3)
    n, d = 2, 5
    opt = Newton() if name == "newton" else WeightedNewton()
    state = opt.init_state(torch.zeros((n, d)))
    prev = [np.eye(d, dtype=np.float32)] * n
    for step in range(4):
        A = rng.normal(size=(n, d, d)).astype(np.float32)
        H = (A @ np.transpose(A, (0, 2, 1)) + np.eye(d, dtype=np.float32)).astype(np.float32)
        if step == 2:  # an indefinite one
            H[:, 0, 0] = -H[:, 0, 0]
        g = rng.normal(size=(n, d)).astype(np.float32)
        got, state = opt.step(torch.from_numpy(g), state, step, hess=torch.from_numpy(H))
        for r in range(n):
            if name == "newton" or step == 0:
                Hw = H[r]
            else:
                w = mg.newton_weight(step)
                Hw = (w * prev[r] + (np.float32(1) - w) * H[r]).astype(np.float32)
            exp = -(mg.pinv_hermitian(Hw) @ g[r].astype(np.float64))
            np.testing.assert_allclose(got[r].numpy(), exp, rtol=2e-4, atol=1e-5 * np.abs(exp).max())
        prev = list(H)
    # a singular Hessian: the null direction is dropped, not inverted
    Hs = np.diag([2.0, 1.0, 0.0, 0.0, 0.0]).astype(np.float32)[None]
    d0, _ = Newton().step(torch.ones((1, d)), (), 0, hess=torch.from_numpy(Hs))
    np.testing.assert_allclose(d0[0].numpy(), [-0.5, -1.0, 0, 0, 0], atol=1e-6)
    with pytest.raises(ValueError, match="compute_hessian_fn"):
        opt.step(torch.ones((1, d)), opt.init_state(torch.zeros((1, d))), 0)


def test_cooling_schedules_follow_the_reference():
    # core/rw_sigma.py:193-225 and the reference's tests/test_core/test_rw_sigma_cooling.py:
    # factor 1 at t = 0, the final value at t = M (cosine), 1 / (1 + s t) (hyperbolic)
    import pickle
    from oracle import filters as fl
    rw = RWSigma({"a": 0.1, "b": 0.2}, init_names=["b"])
    cos = rw.cosine_cooling(0.1, 10)
    assert np.isclose(cos.cooling_factor(0, 0, 50), 1.0) and np.isclose(cos.cooling_factor(0, 10, 50), 0.1)
    assert np.isclose(cos.cooling_factor(25, 4, 50), 0.1 + 0.9 * 0.5 * (1 + np.cos(np.pi * 4.5 / 10)), rtol=1e-6)
    hyp = rw.hyperbolic_cooling(0.3)
    assert np.isclose(hyp.cooling_factor(10, 2, 40), 1 / (1 + 0.3 * 2.25), rtol=1e-6)
    cus = rw.custom_cooling(lambda nt, m, ntimes: 0.9 ** (m + nt / ntimes))
    assert np.isclose(cus.cooling_factor(5, 1, 10), 0.9 ** 1.5, rtol=1e-6)
    for bad in (lambda: rw.cosine_cooling(1.5, 10), lambda: rw.cosine_cooling(0.5, 0),
                lambda: rw.hyperbolic_cooling(-1.0), lambda: rw.cooled(-1.0)):
        with pytest.raises(ValueError):
            bad()
    np.testing.assert_allclose(rw.cooled(0.5).sigmas_all, [0.05, 0.1])
    # the tables the kernel reads equal the oracle's restatement, row for row
    # (the geometric schedule is evaluated in the kernel, with the oracle's d_pow)
    o = dict(cosine=dict(c=0.1, M=10), hyperbolic=dict(s=0.3))
    for kind, h in (("cosine", cos), ("hyperbolic", hyp)):
        orw = fl.RWSigma(np.array([0.1, 0.2]), np.array([0.0, 1.0]), a=0.5, cooling=kind, **o.get(kind, {}))
        for m in (0, 3):
            got = h.cooling_factor(np.arange(37), m, 37)
            exp = np.array([orw.cooling_factor(i, m, 37) for i in range(37)], np.float32)
            np.testing.assert_array_equal(got, exp)
    tab = cos.cooling_table(2, [3, 5])
    np.testing.assert_array_equal(tab[:3], cos.cooling_factor(np.arange(3), 2, 3))
    np.testing.assert_array_equal(tab[3:], cos.cooling_factor(np.arange(5), 2, 5))
    assert rw.geometric_cooling(0.5).cooling_table(0, [3]) is None
    back = pickle.loads(pickle.dumps(cus))
    assert np.isclose(back.cooling_factor(5, 1, 10), 0.9 ** 1.5, rtol=1e-6)


def test_panel_train_argument_errors():
    # checked on the host before any device work (core/algorithms/train.py:
    # 196-199; panel/estimation_mixin.py:852-882)
    import pandas as pd
    import pypomp_amd as pp
    import pypomp_amd.functional as F
    from pypomp_amd.models import LGModel
    units = [Series.build(np.zeros((4, 1)), np.arange(1, 5.0), 0.0, nstep=1) for _ in range(3)]
    struct = F.PanelStruct(model=LGModel(1, 1), units=units, shared_param_names=["A11"],
                           unit_param_names=["C11", "Q11", "R11", "X0_1"])
    sh, un = np.ones((1, 1), np.float32), np.ones((1, 3, 4), np.float32)
    eta = pp.LearningRate({n: 0.01 for n in LGModel(1, 1).param_names})
    keys = np.zeros((1, 2, 3, 2), np.uint32)
    with pytest.raises(ValueError, match="not supported for panel train"):
        F.panel_train(struct, sh, un, 10, pp.BFGS(), 2, eta, 0.9, keys)
    with pytest.raises(ValueError, match="does not divide"):
        F.panel_train(struct, sh, un, 10, pp.Adam(), 2, eta, 0.9, keys, chunk_size=2)
    with pytest.raises(ValueError, match="keys must hold"):
        F.panel_train(struct, sh, un, 10, pp.Adam(), 2, eta, 0.9, keys[:, :1])
    with pytest.raises(ValueError, match="shared_array must be"):
        F.panel_train(struct, np.ones((1, 2), np.float32), un, 10, pp.Adam(), 2, eta, 0.9, keys)
    uneven = F.PanelStruct(model=LGModel(1, 1), units=units[:2] + [Series.build(np.zeros((5, 1)), np.arange(1, 6.0),
                                                                                  0.0, nstep=1)],
                           shared_param_names=["A11"], unit_param_names=["C11", "Q11", "R11", "X0_1"])
    with pytest.raises(ValueError, match="same number of observations"):
        F.panel_train(uneven, sh, un, 10, pp.Adam(), 2, eta, 0.9, keys)
    from pypomp_amd.models import LG
    A = np.array([[0.9]])
    P = pp.PanelPomp({"a": LG(A=A, ys=np.zeros((4, 1))), "b": LG(A=A, ys=np.ones((4, 1)))}, pp.PanelParameters([{
        "shared": pd.DataFrame({"shared": [0.9, 1.0]}, index=["A11", "C11"]),
        "unit_specific": pd.DataFrame({"a": [0.1, 0.3, 0.0], "b": [0.1, 0.3, 0.0]}, index=["Q11", "R11", "X0_1"])}]))
    with pytest.raises(TypeError, match="LearningRate"):
        P.train(J=10, M=2, eta=0.01, key=tf.key(1))
    with pytest.raises(ValueError, match="greater than 0"):
        P.train(J=0, M=2, eta=eta, key=tf.key(1))


def test_learning_rate_schedules_and_lookup():
    # core/learning_rate.py: subsets in any order, schedules of length M only,
    # and the four decays as closed forms
    import pypomp_amd as pp
    lr = pp.LearningRate({"a": 0.1, "b": [0.1, 0.2, 0.3]})
    np.testing.assert_array_equal(lr.to_array(["b"], 3)[:, 0], np.float32([0.1, 0.2, 0.3]))
    np.testing.assert_array_equal(lr.to_array(["b", "a"], 3)[:, 1], np.float32([0.1] * 3))
    assert lr["a"] == 0.1 and "b" in lr and len(lr) == 2 and lr.get("z", 7) == 7
    with pytest.raises(ValueError, match="has length 3, expected M=4"):
        lr.to_array(["a"], 4)
    with pytest.raises(ValueError, match="not found in learning rates"):
        lr.to_array(["z"], 3)
    with pytest.raises(ValueError, match="same length"):
        pp.LearningRate({"a": [0.1], "b": [0.1, 0.2]})
    with pytest.raises(TypeError, match="must be float or 1D sequence"):
        pp.LearningRate({"a": "x"})
    base = pp.LearningRate({"a": 0.2, "b": 0.4})
    M, m = 6, np.arange(6)
    cases = {
        "cosine": (base.cosine_decay(0.1, M), 0.1 + 0.9 * 0.5 * (1 + np.cos(np.pi * m / M))),
        "geometric": (base.geometric_decay(0.5, M), 0.5 ** m),
        "linear": (base.linear_decay(0.25, M), 1 - 0.75 * m / (M - 1)),
        "hyperbolic": (base.hyperbolic_decay(2.0, M), 1 / (1 + 2.0 * m)),
    }
    for kind, (lrd, f) in cases.items():
        np.testing.assert_allclose(lrd.to_array(["a", "b"], M), np.stack([0.2 * f, 0.4 * f], 1), rtol=1e-6,
                                   err_msg=kind)
    assert base.cosine_decay(0.1, M) == base.cosine_decay(0.1, M) != base
    with pytest.raises(ValueError, match="Cannot apply geometric decay of length 5"):
        base.cosine_decay(0.1, M).geometric_decay(0.5, 5)


def test_lg_model_dimensions_up_to_eight():
    # the native LG model takes dimensions 1..8 (the reference takes any); the
    # C ABI rejects larger descriptors
    from pypomp_amd.models import LGModel
    m = LGModel(8, 5)
    assert m.d == 64 + 40 + 36 + 15 + 8 and m.S == 8
    with pytest.raises(ValueError):
        LGModel(9, 1)
    with pytest.raises(ValueError):
        LGModel(2, 9)


def test_lg_descriptor_parameter_count_is_checked():
    # the C ABI rejects an LG descriptor whose d disagrees with its dimensions
    # (the kernels index theta rows by the dimensions)
    from pypomp_amd import _native
    from pypomp_amd.models import LGModel
    lib = _native.lib()
    for dx, dy, sep in ((2, 2, False), (6, 5, True), (8, 8, False)):
        m = LGModel(dx, dy, separate_init=sep)
        desc = m.desc()
        th = np.zeros((1, m.d), np.float32)
        assert lib.mp_host_from_est(desc, th.ctypes.data, 1) == 0
        desc.d = m.d - 1
        assert lib.mp_host_from_est(desc, th.ctypes.data, 1) != 0
        assert "parameters" in lib.mp_last_error().decode()
