"""The PMMH oracle (oracle/pmcmc.py): the Metropolis-Hastings kernel against a
grid posterior under the exact Kalman likelihood, the adaptive proposal's
running covariance, and the change-of-variables term of the prior; the host
mirror's log-Jacobian against the oracle's."""

import numpy as np
import pytest

from oracle import filters as fl, kalman, lg as olg, pmcmc as op, seir as oseir, sv as osv, threefry as tf
from tests.helpers import lg1_theta

F32 = np.float32


def lg1_series(T, seed):
    model = olg.LGModel(olg.LGSpec(1, 1))
    ser = fl.make_series(np.zeros((T, 1)), np.arange(1, T + 1.0), 0.0, nstep=1)
    _, Y = fl.simulate(model, lg1_theta(), ser, tf.split(tf.key(seed), 1)[0])
    return Y[:, 0, :]


def test_mh_kernel_matches_grid_posterior_of_a_log_scale_parameter():
    # unknown R11 (the measurement sd, exp-transformed); flat prior on the
    # natural scale, so the chain's target on log R11 carries the Jacobian
    ys = lg1_series(20, 5)
    th = lg1_theta().astype(np.float64)

    def kal(sd):
        return kalman.kalman_loglik(ys, [[th[0]]], [[th[1]]], [[th[2] ** 2]], [[sd * sd]], [th[4]], [[th[2] ** 2]])

    # the Kalman log-likelihood tabulated on a fine grid and interpolated: the
    # chain and the grid posterior see the same function
    grid = np.linspace(1e-3, 3.0, 3000)
    table = np.array([kal(s) for s in grid])

    def ll_sd(sd):
        return float(np.interp(sd, grid, table, right=-np.inf))

    kinds = [op.EXP]
    chain_ll = lambda x, k: F32(ll_sd(float(np.exp(np.float64(x[0])))))  # noqa: E731
    prior = lambda x: F32(op.log_jac(kinds, x))  # noqa: E731
    ll, est, acc = op.mh_chain(chain_ll, prior, op.diag_rw([0.35]), np.array([np.log(0.3)], F32),
                               tf.key(17), 12000, 1)
    draws = np.exp(est[1000:, 0].astype(np.float64))
    lw = table
    w = np.exp(lw - lw.max())
    w /= w.sum()
    mean = float((w * grid).sum())
    sd = float(np.sqrt((w * (grid - mean) ** 2).sum()))
    assert 0.15 < acc / 12000 < 0.9
    assert abs(draws.mean() - mean) < 0.12 * sd
    assert abs(draws.std() - sd) < 0.12 * sd


def test_adaptive_running_covariance_is_the_batch_covariance():
    rng = np.random.default_rng(3)
    init, step = op.adaptive_rw(np.diag([0.1, 0.2, 0.3]), scale_start=1000, shape_start=1000)
    st = init(np.zeros(3, F32))
    xs = rng.normal(size=(50, 3)).astype(F32)
    for n, x in enumerate(xs, start=1):
        _, st = step(st, x, np.zeros(3, F32), n, 0)
    np.testing.assert_allclose(st[1], xs.astype(np.float64).mean(0), rtol=1e-5, atol=1e-6)
    # Welford divides by n after the first (seeding) state: population covariance
    np.testing.assert_allclose(st[2], np.cov(xs.astype(np.float64).T, ddof=0), rtol=1e-4, atol=1e-6)


@pytest.mark.parametrize("which", ["lg2", "sv", "seir"])
def test_log_jacobian_by_finite_differences_and_host_mirror(which):
    from pypomp_amd.models import LGModel, SEIRModel, SVModel
    rng = np.random.default_rng(1)
    if which == "lg2":
        om, hm = olg.LGModel(olg.LGSpec(2, 2, separate_init=True)), LGModel(2, 2, separate_init=True)
    elif which == "sv":
        om, hm = osv.SVModel(), SVModel()
    else:
        om, hm = oseir.SEIRModel(), SEIRModel()
    kinds = op.transform_kinds(om)
    f64 = {op.ID: lambda x: x, op.EXP: np.exp, op.TANH: np.tanh, op.LOGISTIC: lambda x: 1 / (1 + np.exp(-x))}
    for _ in range(20):
        x = rng.normal(scale=1.5, size=len(kinds))
        h = 1e-6
        fd = sum(np.log(abs((f64[k](xi + h) - f64[k](xi - h)) / (2 * h))) for k, xi in zip(kinds, x))
        assert abs(op.log_jac(kinds, x) - fd) < 1e-6 * max(1, abs(fd))
        assert abs(float(hm.log_jac_from_est(x.astype(F32))) - op.log_jac(kinds, x.astype(F32))) < 2e-5 * max(1, abs(fd))
