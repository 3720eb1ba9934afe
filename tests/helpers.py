"""Shared fixture builders: the reference's regression fixture models
(tests/test_regression/conftest.py:22-115 in the reference) restated as data,
and the benchmark workloads, for both the oracle and the product."""

import json
from functools import lru_cache
from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden"
F32 = np.float32


@lru_cache(None)
def golden():
    return json.loads((GOLDEN / "reference_regression.json").read_text())


@lru_cache(None)
def workloads():
    d = np.load(GOLDEN / "workloads.npz")
    return {k: d[k] for k in d.files}


def assert_close_tol(got, expected, atol=1e-6, rtol=1e-5):
    """Parity bar against the reference's regression values and the oracle:
    1e-5 relative (the stated float32 log-likelihood target), with 1e-6
    absolute for entries near zero."""
    got = np.asarray(got, dtype=float).ravel()
    expected = np.asarray(expected, dtype=float).ravel()
    assert got.shape == expected.shape, (got.shape, expected.shape)
    np.testing.assert_allclose(got, expected, atol=atol, rtol=rtol)


def lg1_theta():
    """1-D LG fixture theta (A=0.9, defaults otherwise) in canonical order."""
    return np.array([0.9, 1.0, 0.1, np.sqrt(0.1), 0.0], dtype=F32)


def lg1_data(seed):
    """Observations of LG(A=[[0.9]], T=4, key=key(seed)) via the oracle's simulate tree."""
    from oracle import filters as fl, lg, threefry as tf
    model = lg.LGModel(lg.LGSpec(1, 1))
    ser = fl.make_series(np.zeros((4, 1)), np.arange(1, 5, dtype=float), 0.0, nstep=1)
    k = tf.split(fl.pomp_new_key(tf.key(seed)), 1)[0]
    _, Y = fl.simulate(model, lg1_theta(), ser, k)
    return Y[:, 0, :]
