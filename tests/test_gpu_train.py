"""GPU parity of the LG MOP adjoint against the oracle's autograd gradient, and of
the train loop (core/algorithms/train.py:38-180) against the reference's own
train regression values and the oracle's train restatement."""

import numpy as np
import pytest

from oracle import filters as fl, lg as olg, mop_grad as mg, threefry as tf
from tests.helpers import assert_close_tol, lg1_data, lg1_theta, workloads

pytestmark = pytest.mark.gpu


def lg1_long(T, seed):
    """LG(1, 1) observations of length T from the oracle's simulate tree."""
    model = olg.LGModel(olg.LGSpec(1, 1))
    ser = fl.make_series(np.zeros((T, 1)), np.arange(1, T + 1.0), 0.0, nstep=1)
    _, Y = fl.simulate(model, lg1_theta(), ser, tf.split(tf.key(seed), 1)[0])
    return Y[:, 0, :].astype(np.float32)


def struct_of(model, ys):
    import pypomp_amd.functional as F
    from pypomp_amd.series import Series
    return F.PompStruct(model=model, units=[Series.build(ys, np.arange(1, len(ys) + 1.0), 0.0, nstep=1)])


@pytest.mark.parametrize("K", [5, 50])
def test_lg1_gradient_matches_autograd_oracle(K):
    import pypomp_amd.functional as F
    from pypomp_amd.models import LGModel
    ys = lg1_long(60, 7)
    struct = struct_of(LGModel(1, 1), ys)
    omodel = olg.LGModel(olg.LGSpec(1, 1))
    th = omodel.to_est(lg1_theta())
    keys = tf.split(tf.key(5), 2)
    ll, g = F.mop_grad(struct, np.stack([th, th]), 200, 0.9, keys, K=K)
    for r in range(2):
        val, _, gref = mg.lg_mop_value_and_grad(omodel, th, ys, keys[r], 200, 0.9)
        assert ll[r] == val
        scale = np.abs(gref).max()
        np.testing.assert_allclose(g[r], gref, rtol=1e-3, atol=1e-4 * scale)


def test_lg2_separate_init_gradient_matches_autograd_oracle():
    import pypomp_amd.functional as F
    from pypomp_amd.models import LGModel
    w = workloads()
    ys = w["lg2_seed1_ys"].astype(np.float32)
    struct = struct_of(LGModel(2, 2, separate_init=True), ys)
    omodel = olg.LGModel(olg.LGSpec(2, 2, separate_init=True))
    th = omodel.to_est(w["lg2_seed1_theta"])
    keys = tf.split(tf.key(8), 1)
    ll, g = F.mop_grad(struct, th[None], 160, 0.95, keys, K=9)
    val, _, gref = mg.lg_mop_value_and_grad(omodel, th, ys, keys[0], 160, 0.95)
    assert ll[0] == val
    scale = np.abs(gref).max()
    np.testing.assert_allclose(g[0], gref, rtol=2e-3, atol=2e-4 * scale)


def test_unsupported_model_raises():
    import pypomp_amd.functional as F
    from pypomp_amd._native import NativeError
    from pypomp_amd.models import LGModel
    ys = np.zeros((10, 3), np.float32)
    struct = struct_of(LGModel(2, 3), ys)
    th = LGModel(2, 3).to_est(np.ones(LGModel(2, 3).d, np.float32))
    with pytest.raises(NativeError, match="no adjoint"):
        F.mop_grad(struct, th[None], 64, 0.9, tf.split(tf.key(1), 1))


def test_train_regression(golden):
    # the reference's train regression fixture: LG(1, 1), T = 4, J = 12, three
    # Adam steps with eta 0.01, alpha 0, key split(key(20260811), 1)[0]; the
    # natural theta is passed through as the estimation-scale start
    import pypomp_amd as pp
    import pypomp_amd.functional as F
    from pypomp_amd.models import LGModel
    struct = struct_of(LGModel(1, 1), lg1_data(0))
    eta = pp.LearningRate({n: 0.01 for n in struct.param_names})
    nll, tr = F.train(struct, lg1_theta()[None], 12, pp.Adam(), 3, eta, 0.0, tf.split(tf.key(20260811), 1))
    g = golden["test_train_regression"]
    assert np.isnan(nll[0, 0]) and nll.shape == (1, 4) and tr.shape == (1, 4, 5)
    assert_close_tol(nll[0, -1], g["final_neg_loglik"])
    assert_close_tol(tr[0, -1], g["final_theta"])


def test_train_matches_oracle_train():
    # 20 Adam steps with alpha cooling on a longer series against the oracle's
    # restatement driven by its own value and gradient
    import pypomp_amd as pp
    import pypomp_amd.functional as F
    from pypomp_amd.models import LGModel
    ys = lg1_long(40, 3)
    struct = struct_of(LGModel(1, 1), ys)
    omodel = olg.LGModel(olg.LGSpec(1, 1))
    th0 = omodel.to_est(lg1_theta())
    M, J = 20, 100
    rates = np.linspace(0.02, 0.005, M)
    eta = pp.LearningRate({n: rates for n in struct.param_names})
    key = tf.split(tf.key(99), 1)
    nll, tr = F.train(struct, th0[None], J, pp.Adam(), M, eta, 0.8, key, alpha_cooling=0.5)

    def vg(th, k, alpha):
        val, _, gr = mg.lg_mop_value_and_grad(omodel, th, ys, k, J, alpha)
        return val, gr

    onll, otr = mg.train(vg, th0, key[0], 40, M, eta.to_array(struct.param_names, M), 0.8, alpha_cooling=0.5)
    assert nll[0, 1] == onll[1]  # first step: same theta, same draw
    np.testing.assert_allclose(nll[0, 1:], onll[1:], rtol=1e-4)
    np.testing.assert_allclose(tr[0], otr, rtol=1e-4, atol=1e-5)


def test_train_sgd_and_monitors():
    # n_monitors > 1: the monitor is the mean pfilter neg-loglik at the trace's
    # previous theta (should_trans), keys split off after the MOP key
    import pypomp_amd as pp
    import pypomp_amd.functional as F
    from pypomp_amd import random as prandom
    from pypomp_amd.models import LGModel
    ys = lg1_long(30, 4)
    struct = struct_of(LGModel(1, 1), ys)
    th0 = LGModel(1, 1).to_est(lg1_theta())
    eta = pp.LearningRate({n: 0.01 for n in struct.param_names})
    keys = tf.split(tf.key(12), 2)
    nll, tr = F.train(struct, np.stack([th0, th0]), 64, pp.SGD(), 3, eta, 0.97, keys, n_monitors=3)
    assert np.all(np.isfinite(nll[:, 1:]))
    k = keys[1]
    for m in range(3):
        k2 = prandom.split(k, 2)
        k = k2[0]
        ks = prandom.split(k, 4)
        k, mon = ks[0], ks[1:]
        res = F.pfilter(struct, tr[1, m][None], 64, mon[None], should_trans=True)
        np.testing.assert_allclose(nll[1, m + 1], -res["logLik"].mean(), rtol=1e-6)
        k = prandom.split(k, 2)[0]
    _, g = F.mop_grad(struct, th0[None], 64, 0.97, prandom.split(keys[0], 2)[1:])
    np.testing.assert_allclose(tr[0, 1], th0 + 0.01 * (g[0] / 30), rtol=1e-5, atol=1e-7)


def test_pomp_train_object_api():
    import pypomp_amd as pp
    from pypomp_amd.models import LGModel
    model = LGModel(1, 1)
    ys = lg1_long(25, 6)
    names = model.param_names
    P = pp.Pomp(ys=ys, theta=[dict(zip(names, map(float, lg1_theta())))] * 2, t0=0.0, nstep=1, model=model,
                key=tf.key(3))
    P.train(J=50, M=4, eta=pp.LearningRate({n: 0.01 for n in names}))
    r = P.results_history[-1]
    assert r["method"] == "train" and r["traces"].shape == (2, 5, 1 + len(names))
    np.testing.assert_allclose(r["traces"][:, 0, 1:], np.stack([lg1_theta()] * 2), rtol=1e-6)
    df = P.traces()
    assert len(df) == 10 and list(df.columns[:5]) == ["theta_idx", "iteration", "method", "logLik", "se"]
    assert P.results().shape[0] == 2


def _lg1_train_case(M=8, J=100):
    import pypomp_amd as pp
    from pypomp_amd.models import LGModel
    ys = lg1_long(40, 3)
    struct = struct_of(LGModel(1, 1), ys)
    omodel = olg.LGModel(olg.LGSpec(1, 1))
    th0 = omodel.to_est(lg1_theta())
    rates = np.linspace(0.02, 0.005, M)
    eta = pp.LearningRate({n: rates for n in struct.param_names})

    def vg(th, k, alpha):
        val, _, gr = mg.lg_mop_value_and_grad(omodel, th, ys, k, J, alpha)
        return val, gr

    return struct, omodel, ys, th0, eta, vg


@pytest.mark.parametrize("name", ["fmadam", "bfgs"])
def test_train_full_matrix_adam_and_bfgs_match_oracle(name):
    # FullMatrixAdam (core/optimizer.py:157-213) and BFGS (:216-267) against the
    # oracle's restatement driven by its own value and gradient
    import pypomp_amd as pp
    import pypomp_amd.functional as F
    M, J = 8, 100
    struct, omodel, ys, th0, eta, vg = _lg1_train_case(M, J)
    opt = pp.FullMatrixAdam() if name == "fmadam" else pp.BFGS()
    key = tf.split(tf.key(41), 1)
    nll, tr = F.train(struct, th0[None], J, opt, M, eta, 0.8, key, alpha_cooling=0.5)
    onll, otr = mg.train(vg, th0, key[0], 40, M, eta.to_array(struct.param_names, M), 0.8, alpha_cooling=0.5,
                         optimizer=name)
    np.testing.assert_allclose(nll[0, 1:], onll[1:], rtol=1e-4)
    np.testing.assert_allclose(tr[0], otr, rtol=1e-4, atol=1e-5)
    assert not np.allclose(tr[0, -1], tr[0, 0])


def test_train_line_search_matches_oracle():
    # Armijo backtracking (core/algorithms/train.py:139-162, 442-499): trial
    # points scored by the bootstrap filter with the iteration's MOP key
    import pypomp_amd as pp
    import pypomp_amd.functional as F
    M, J = 5, 100
    struct, omodel, ys, th0, eta, vg = _lg1_train_case(M, J)
    eta = pp.LearningRate({n: 0.5 for n in struct.param_names})   # large: forces backtracking
    ser = fl.make_series(ys, np.arange(1, len(ys) + 1.0), 0.0, nstep=1)

    def neg_pf(th, k):
        return -fl.pfilter(omodel, th, ser, k, J, should_trans=True)["logLik"]

    key = tf.split(tf.key(43), 1)
    nll, tr = F.train(struct, th0[None], J, pp.SGD(ls=True), M, eta, 0.9, key)
    onll, otr = mg.train(vg, th0, key[0], 40, M, eta.to_array(struct.param_names, M), 0.9, optimizer="sgd",
                         ls=neg_pf)
    np.testing.assert_allclose(nll[0, 1:], onll[1:], rtol=1e-4)
    np.testing.assert_allclose(tr[0], otr, rtol=1e-4, atol=1e-5)
    # the search backtracked: fixed steps of the full rate give another path
    _, tr_fixed = F.train(struct, th0[None], J, pp.SGD(), M, eta, 0.9, key)
    assert not np.allclose(tr[0, 1:], tr_fixed[0, 1:])


@pytest.mark.parametrize("name", ["newton", "wnewton"])
def test_train_newton_matches_oracle(name):
    # Newton and WeightedNewton (core/optimizer.py:270-330) on the native
    # Hessian, drawn with each iteration's Hessian subkey (train.py:132-141),
    # against the oracle's train on its autograd gradient and double-backward
    # Hessian
    # WeightedNewton runs two steps (the first is a plain Newton step, the
    # second weights the previous Hessian): along this path the later weighted
    # Hessians are nearly singular (LG 1-D's C, Q, X0 scaling is a flat
    # direction of the likelihood; smallest eigenvalue ~1e-3 of the largest),
    # and pinv then turns float32 rounding in H into different directions.  Its
    # arithmetic on identical inputs is pinned in tests/test_host.py.
    import pypomp_amd as pp
    import pypomp_amd.functional as F
    M, J = (6, 100) if name == "newton" else (2, 100)
    struct, omodel, ys, th0, eta, vg = _lg1_train_case(M, J)
    opt = pp.Newton() if name == "newton" else pp.WeightedNewton()
    key = tf.split(tf.key(21), 1)
    nll, tr = F.train(struct, th0[None], J, opt, M, eta, 0.9, key)

    def hess(th, k, alpha):
        return mg.hessian_and_fd(mg.lg_hessian_case(omodel, th, ys, k, J, alpha), th, 40, fd=False)[2]

    onll, otr = mg.train(vg, th0, key[0], 40, M, eta.to_array(struct.param_names, M), 0.9, optimizer=name,
                         hessian=hess)
    assert nll[0, 1] == onll[1]
    np.testing.assert_allclose(nll[0, 1:], onll[1:], rtol=1e-4)
    np.testing.assert_allclose(tr[0], otr, rtol=1e-3, atol=1e-4)
    # a Newton path is not the gradient path
    _, tr_sgd = F.train(struct, th0[None], J, pp.SGD(), M, eta, 0.9, key)
    assert not np.allclose(tr[0, 1:], tr_sgd[0, 1:])


def test_train_line_search_needs_a_monitor():
    import pypomp_amd as pp
    import pypomp_amd.functional as F
    struct, _, _, th0, eta, _ = _lg1_train_case(2, 50)
    with pytest.raises(ValueError, match="monitor"):
        F.train(struct, th0[None], 50, pp.SGD(ls=True), 2, eta, 0.9, tf.split(tf.key(1), 1), n_monitors=0)


def test_train_rows_are_independent_bfgs_line_search():
    # two parameter rows in one run: each row's BFGS state, line-search rounds
    # and trace equal the oracle's single-row run with that row's key
    import pypomp_amd as pp
    import pypomp_amd.functional as F
    M, J = 5, 100
    struct, omodel, ys, th0, _, vg = _lg1_train_case(M, J)
    eta = pp.LearningRate({n: 0.3 for n in struct.param_names})
    ser = fl.make_series(ys, np.arange(1, len(ys) + 1.0), 0.0, nstep=1)

    def neg_pf(th, k):
        return -fl.pfilter(omodel, th, ser, k, J, should_trans=True)["logLik"]

    rows = np.stack([th0, th0 + np.float32(0.05)])
    keys = tf.split(tf.key(47), 2)
    nll, tr = F.train(struct, rows, J, pp.BFGS(ls=True), M, eta, 0.9, keys)
    for r in range(2):
        onll, otr = mg.train(vg, rows[r], keys[r], 40, M, eta.to_array(struct.param_names, M), 0.9,
                             optimizer="bfgs", ls=neg_pf)
        np.testing.assert_allclose(nll[r, 1:], onll[1:], rtol=1e-4)
        np.testing.assert_allclose(tr[r], otr, rtol=1e-4, atol=1e-5)
