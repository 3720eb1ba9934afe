"""GPU parity of PMMH (core/algorithms/pmcmc.py:30-163) against the oracle's
restatement: the same chains, log-likelihoods, priors and acceptance counts."""

import numpy as np
import pytest

from oracle import filters as fl, lg as olg, pmcmc as op, threefry as tf
from tests.helpers import lg1_data, lg1_theta

pytestmark = pytest.mark.gpu


def lg1_struct():
    import pypomp_amd.functional as F
    from pypomp_amd.models import LGModel
    from pypomp_amd.series import Series
    return F.PompStruct(model=LGModel(1, 1), units=[Series.build(lg1_data(0), np.arange(1, 5.0), 0.0, nstep=1)])


def oracle_chains(thetas, keys, proposal, M, J, dprior=None, thresh=0.0):
    model = olg.LGModel(olg.LGSpec(1, 1))
    ser = fl.make_series(lg1_data(0), np.arange(1, 5.0), 0.0, nstep=1)
    return [op.pmcmc(model, ser, thetas[c], proposal, M, J, keys[c], thresh, dprior) for c in range(len(keys))]


def compare(got, ref):
    ll, lp, th, acc = got
    for c, (oll, olp, oth, oacc) in enumerate(ref):
        np.testing.assert_array_equal(ll[c], oll)
        np.testing.assert_array_equal(th[c], oth)
        np.testing.assert_array_equal(lp[c], olp)
        assert acc[c] == oacc


def test_diag_random_walk_flat_prior():
    import pypomp_amd as pp
    import pypomp_amd.functional as F
    names = lg1_struct().param_names
    thetas = np.stack([lg1_theta(), lg1_theta() * np.float32(1.1)])
    keys = tf.split(tf.key(42), 2)
    got = F.pmcmc(lg1_struct(), thetas, pp.MVNDiagRW({n: 0.05 for n in names}), 40, 50, keys)
    ref = oracle_chains(thetas, keys, op.diag_rw([0.05] * 5), 40, 50)
    compare(got, ref)
    assert got[0].shape == (2, 41) and got[2].shape == (2, 41, 5) and np.all(got[1] == 0)
    assert 0 < got[3].sum() < 80


def test_adaptive_proposal_user_prior_and_threshold():
    # both adaptation phases, a normal prior on A11 and resampling only when
    # the weights spread
    import pypomp_amd as pp
    import pypomp_amd.functional as F
    names = lg1_struct().param_names
    thetas = lg1_theta()[None]
    keys = tf.split(tf.key(7), 1)
    prior = lambda th: -0.5 * ((float(th["A11"]) - 0.8) / 0.1) ** 2  # noqa: E731
    prop = pp.MVNRWAdaptive(rw_sd={"A11": 0.05, "R11": 0.05, "Q11": 0.02}, scale_start=5, shape_start=6)
    got = F.pmcmc(lg1_struct(), thetas, prop, 40, 60, keys, thresh=0.5, dprior=prior)
    V = np.zeros((5, 5))
    for n, s in (("A11", 0.05), ("R11", 0.05), ("Q11", 0.02)):
        V[names.index(n), names.index(n)] = s * s
    oprior = lambda th: -0.5 * ((float(th["A11"]) - 0.8) / 0.1) ** 2  # noqa: E731
    ref = oracle_chains(thetas, keys, op.adaptive_rw(V, scale_start=5, shape_start=6), 40, 60, oprior, thresh=0.5)
    compare(got, ref)
    # the parameters outside the proposal's support never move
    assert np.all(got[2][0, :, names.index("C11")] == lg1_theta()[names.index("C11")])


def test_full_covariance_proposal():
    import pypomp_amd as pp
    import pypomp_amd.functional as F
    cov = np.array([[0.004, 0.001], [0.001, 0.002]])
    thetas = lg1_theta()[None]
    keys = tf.split(tf.key(8), 1)
    got = F.pmcmc(lg1_struct(), thetas, pp.MVNRWFull(cov, ["A11", "X0_1"]), 25, 40, keys)
    names = lg1_struct().param_names
    L = np.zeros((5, 5))
    ia, ix = names.index("A11"), names.index("X0_1")
    L[np.ix_([ia, ix], [ia, ix])] = np.linalg.cholesky(cov)
    ref = oracle_chains(thetas, keys, op.full_rw(L), 25, 40)
    compare(got, ref)


def test_pomp_pmcmc_results():
    import pypomp_amd as pp
    from pypomp_amd.models.lg import LG
    P = LG(A=np.array([[0.9]]), key=pp.random.key(0))
    P.pmcmc(J=40, M=10, proposal=pp.MVNDiagRW({"A11": 0.05}), key=pp.random.key(2))
    df = P.results()
    assert list(df.columns[:4]) == ["theta_idx", "iteration", "logLik", "log_prior"] and len(df) == 11
    tr = P.traces()
    assert list(tr.columns[:6]) == ["theta_idx", "iteration", "method", "logLik", "se", "log_prior"]
    r = P.results_history[-1]
    assert r["accepts"].shape == (1,) and 0 <= r["acceptance_rate"][0] <= 1
