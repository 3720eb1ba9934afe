"""Write the golden fixtures of tests/golden (run in the build container).

* ``reference_regression.json`` -- the expected values of the reference's
  locked regression baselines (data files held by the reference's tests:
  tests/test_regression/*/*.csv), copied as numbers, with the fixture
  definitions of tests/test_regression/conftest.py:22-115 restated as data.
* ``workloads.npz`` -- synthetic observation series of the benchmark workloads
  (LG 2-D seeds 1 and 2, SEIR seed 3, SV seed 4, SEIR panel seed 5), simulated
  by the oracle with the reference's simulate key tree.  They are inputs to the
  filters, not expected outputs.

Usage: python tests/golden/make_fixtures.py [--reference /root/reference] [--regression-only]
       python tests/golden/make_fixtures.py --only sv[,lg,seir,panel]   (workloads only)
"""

import argparse
import csv
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))

from oracle import filters as fl  # noqa: E402
from oracle import lg, seir, sv  # noqa: E402
from oracle import threefry as tf  # noqa: E402


def read_csv(path):
    with open(path) as f:
        rows = list(csv.reader(f))
    head = rows[0][1:]
    cols = {h: [] for h in head}
    for r in rows[1:]:
        for h, v in zip(head, r[1:]):
            if v != "":
                cols[h].append(float(v))
    return cols


def regression(ref):
    base = Path(ref) / "tests" / "test_regression"
    out = {}
    for name in ["test_simulate_regression", "test_pfilter_regression", "test_pfilter_diagnostics_regression",
                 "test_mop_regression", "test_mif_regression", "test_mif_full_trace_regression",
                 "test_panel_pfilter_regression", "test_panel_mif_regression", "test_train_regression",
                 "test_panel_train_regression"]:
        d = name.replace("_diagnostics", "").replace("_full_trace", "")
        path = base / d / f"{name}.csv"
        out[name] = read_csv(path)
    out["fixtures"] = {
        "lg_struct": {"A": 0.9, "T": 4, "data_seed": 0, "key_seed": 20260811, "J": 12, "n_reps": 1},
        "lg_struct_multi": {"A": 0.9, "T": 4, "data_seed": 0, "key_seed": 20260812, "J": 12, "n_reps": 2,
                            "theta_scale_rep1": 1.05},
        "lg_panel_struct": {"data_seeds": [0, 1], "key_seed": 20260811, "J": 12,
                            "shared": ["A11", "C11"], "unit": ["Q11", "R11", "X0_1"]},
        "mif": {"M": 3, "TRACE_M": 4, "rw_sd": 0.02, "cooling": 0.5},
        "mop": {"alpha": 0.5},
        "tolerance": {"atol": 1e-6, "rtol": 1e-4},
    }
    return out


def lg_data(dx, T, seed, P0=None, Q=None, R=None):
    spec = lg.LGSpec(dx, dx, separate_init=P0 is not None)
    model = lg.LGModel(spec)
    th = lg.default_theta(dx=dx, P0=P0, Q=Q, R=R)
    theta = np.array([th[n] for n in spec.param_names], dtype=np.float32)
    ser = fl.make_series(np.zeros((T, dx)), np.arange(1, T + 1, dtype=float), 0.0, nstep=1)
    k = tf.split(fl.pomp_new_key(tf.key(seed)), 1)[0]
    _, Y = fl.simulate(model, theta, ser, k)
    return Y[:, 0, :].astype(np.float32), theta


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default="/root/reference")
    ap.add_argument("--only", default="", help="comma list of workloads to refresh (lg, seir, sv, panel); "
                    "keeps the others and the regression values")
    ap.add_argument("--regression-only", action="store_true", help="rewrite reference_regression.json only")
    args = ap.parse_args()
    if args.regression_only:
        (HERE / "reference_regression.json").write_text(json.dumps(regression(args.reference), indent=1))
        print("wrote", HERE / "reference_regression.json")
        return
    only = set(filter(None, args.only.split(",")))
    arrays = {}
    if only:
        old = np.load(HERE / "workloads.npz")
        arrays = {k: old[k] for k in old.files}
    else:
        (HERE / "reference_regression.json").write_text(json.dumps(regression(args.reference), indent=1))
    want = (lambda name: not only or name in only)
    if want("lg"):
        # LG 2-D, Q = R = 0.01 I, x0 ~ N(0, I): seed 1 T=50 (correctness
        # workload), seed 2 T=200 (throughput workload)
        for seed, T in ((1, 50), (2, 200)):
            ys, theta = lg_data(2, T, seed, P0=np.eye(2), Q=0.01 * np.eye(2), R=0.01 * np.eye(2))
            arrays[f"lg2_seed{seed}_ys"] = ys
            arrays[f"lg2_seed{seed}_theta"] = theta
    if want("seir"):
        # SEIR, seed 3, T = 260 weeks, 7 Euler steps per week
        ys, theta = seir.simulate_data(seed=3, T=260)
        arrays["seir_seed3_ys"], arrays["seir_seed3_theta"] = ys, theta
    if want("sv"):
        # SV, seed 4, T = 1000
        ys, theta = sv.simulate_data(seed=4, T=1000)
        arrays["sv_seed4_ys"], arrays["sv_seed4_theta"] = ys, theta
    if want("panel"):
        # SEIR panel, seed 5: 100 units, beta_u ~ LogNormal(log 1.5, 0.2), T_u ~ U{52..520}
        panel = seir.simulate_panel(seed=5, U=100)
        for k, v in panel.items():
            arrays[f"panel_{k}"] = v
    np.savez_compressed(HERE / "workloads.npz", **arrays)
    # exact FP64 Kalman log-likelihoods of the LG series (the accuracy yardstick
    # the benchmark reports against, without running the oracle itself)
    from oracle import kalman
    kal = {}
    for seed in (1, 2):
        spec = lg.LGSpec(2, 2, separate_init=True)
        A, C, LQ, LR, X0, L0 = (np.asarray(m, np.float64) for m in spec.unpack(arrays[f"lg2_seed{seed}_theta"]))
        kal[f"lg2_seed{seed}"] = kalman.kalman_loglik(arrays[f"lg2_seed{seed}_ys"], A, C, LQ @ LQ.T, LR @ LR.T, X0,
                                                      L0 @ L0.T)
    (HERE / "kalman_loglik.json").write_text(json.dumps(kal, indent=1))
    print("wrote", HERE / "workloads.npz", HERE / "kalman_loglik.json")


if __name__ == "__main__":
    main()
