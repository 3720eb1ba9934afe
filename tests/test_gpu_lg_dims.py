"""GPU parity of the linear-Gaussian model at the dimensions of its generic
instances (linear_gaussian.py takes any dx, dy): dimension 3 (the instance of
dimension 4) and 5 to 8 (the instance of dimension 8).  pfilter in both forms,
MOP, simulate and the IF2 perturbed filter against the oracle bit for bit, and
the particle log-likelihood against the exact Kalman value."""

import numpy as np
import pytest

from oracle import filters as fl, kalman, lg as olg, threefry as tf

pytestmark = pytest.mark.gpu
F32 = np.float32
T = 12


def lg_case(dx, dy, seed=0):
    """The GPU struct, oracle model, natural theta, observations and oracle
    series of an LG(dx, dy) model at the reference defaults, with data drawn
    from the oracle's simulator."""
    import pypomp_amd.functional as F
    from pypomp_amd.models import LGModel
    from pypomp_amd.series import Series
    spec = olg.LGSpec(dx, dy)
    th = olg.default_theta(dx, dy)
    theta = np.array([th[n] for n in spec.param_names], F32)
    omodel = olg.LGModel(spec)
    times = np.arange(1, T + 1.0)
    _, Y = fl.simulate(omodel, theta, fl.make_series(np.zeros((T, dy), F32), times, 0.0, nstep=1),
                       tf.split(tf.key(600 + seed), 1)[0])
    ys = np.ascontiguousarray(Y[:, 0, :], dtype=F32)
    model = LGModel(dx, dy)
    assert model.param_names == spec.param_names
    struct = F.PompStruct(model=model, units=[Series.build(ys, times, 0.0, nstep=1)])
    return struct, omodel, theta, ys, fl.make_series(ys, times, 0.0, nstep=1)


@pytest.mark.parametrize("form", ["batched", "plain"])
@pytest.mark.parametrize("dx,dy", [(3, 2), (6, 5), (8, 8)])
def test_lg_dims_pfilter_matches_oracle(dx, dy, form):
    import pypomp_amd.functional as F
    struct, omodel, theta, _, ser = lg_case(dx, dy)
    keys = tf.split(tf.key(610), 2)
    out = F.pfilter(struct, theta[None], 300, keys.reshape(1, 2, 2), form=form, CLL=True, prediction_mean=True)
    for r in range(2):
        o = fl.pfilter(omodel, theta, ser, keys[r], 300, diagnostics=True)
        assert float(out["logLik64"][0, r]) == o["logLik64"], (r, out["logLik64"][0, r], o["logLik64"])
        np.testing.assert_array_equal(out["CLL"][0, r], np.array(o["CLL"], dtype=F32))
        np.testing.assert_allclose(out["prediction_mean"][0, r], np.array(o["prediction_mean"]),
                                   rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("form", ["batched", "plain"])
@pytest.mark.parametrize("dx,dy", [(6, 5), (8, 8)])
def test_lg_dims_mop_matches_oracle(dx, dy, form):
    import pypomp_amd.functional as F
    struct, omodel, theta, _, ser = lg_case(dx, dy, seed=1)
    th_est = omodel.to_est(theta)
    keys = tf.split(tf.key(611), 2)
    got = F.mop(struct, np.stack([th_est, th_est]), 200, 0.9, keys, form=form, return_float64=True)
    for r in range(2):
        assert -got[r] == fl.mop(omodel, th_est, ser, keys[r], 200, 0.9)


@pytest.mark.parametrize("dx,dy", [(6, 5), (8, 8)])
def test_lg_dims_simulate_matches_oracle(dx, dy):
    import pypomp_amd.functional as F
    struct, omodel, theta, _, ser = lg_case(dx, dy, seed=2)
    keys = tf.split(tf.key(612), 2)
    X, Y = F.simulate(struct, np.stack([theta, theta]), 7, keys)
    for r in range(2):
        OX, OY = fl.simulate(omodel, theta, ser, keys[r], nsim=7)
        np.testing.assert_array_equal(X[r], OX.transpose(1, 0, 2))
        np.testing.assert_array_equal(Y[r], OY.transpose(1, 0, 2))


def test_lg_dims_perfilter_matches_oracle():
    # IF2's perturbed filter carries all 108 parameters of LG(6, 5) per particle
    import pypomp_amd as pp
    import pypomp_amd.functional as F
    struct, omodel, theta, _, ser = lg_case(6, 5, seed=3)
    names = omodel.spec.param_names if hasattr(omodel, "spec") else olg.LGSpec(6, 5).param_names
    J = 60
    rng = np.random.default_rng(9)
    th = np.repeat(theta[None], J, 0)
    th_est = (omodel.to_est(th) + F32(0.01) * rng.standard_normal(th.shape).astype(F32)).astype(F32)
    init = [n for n in names if n.startswith("X0_")]
    rw = pp.RWSigma({n: 0.01 for n in names}, init_names=init).geometric_cooling(0.5)
    k = tf.key(613)
    out, _, ll, anc = F.perfilter(struct, th_est[None], k[None], J, rw, 1, ancestry=True)
    orw = fl.RWSigma(np.full(len(names), 0.01), np.array([n in init for n in names], np.float64), a=0.5)
    eout, ell, eanc = fl.perfilter(omodel, 1, th_est, ser, k, J, orw, return_ancestry=True)
    np.testing.assert_array_equal(anc[0], eanc)
    np.testing.assert_array_equal(out[0], eout)
    assert abs(float(ll[0]) - float(ell)) <= 1e-5 * abs(float(ell))


def test_lg_dims_against_kalman():
    import pypomp_amd.functional as F
    struct, omodel, theta, ys, _ = lg_case(6, 5, seed=4)
    _, C, LQ, LR, X0, L0 = olg.LGSpec(6, 5).unpack(theta)
    A = olg.LGSpec(6, 5).unpack(theta)[0]
    d = lambda M: M.astype(float) @ M.T.astype(float)  # noqa: E731
    exact = kalman.kalman_loglik(ys, A.astype(float), C.astype(float), d(LQ), d(LR), X0.astype(float), d(L0))
    reps = 32
    for form in ("batched", "plain"):
        out = F.pfilter(struct, theta[None], 4000, tf.split(tf.key(614), reps).reshape(1, reps, 2), form=form)
        ll = out["logLik64"].ravel()
        se = ll.std(ddof=1) / np.sqrt(reps)
        assert abs(ll.mean() - exact) < 4 * se + 0.05, (form, ll.mean(), exact, se)
