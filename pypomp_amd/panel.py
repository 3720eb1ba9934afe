"""``PanelPomp``: the object API of panels (pypomp/panel/panel.py,
panel/estimation_mixin.py:384-773, panel/analysis_mixin.py:74-200).

A panel holds unit ``Pomp`` objects of one native model with shared and
unit-specific parameters (``PanelParameters``, one or more replicate sets).
``pfilter`` runs every (set, replicate, unit) filter in one launch; ``mif`` is
panel IF2 (functional.panel_mif).  Keys follow the reference:
``fresh_key``-derived key split into n_sets * reps * U filter keys for pfilter
and into n_sets chain keys for mif.
"""

import time

import numpy as np
import pandas as pd

from . import functional as F
from . import random as prandom
from .maths import logmeanexp, logmeanexp_se


class PanelParameters:
    """Replicate sets of panel parameters (core/parameters.py PanelParameters):
    ``sets`` is a list of {"shared": DataFrame(index = names, one column) or
    None, "unit_specific": DataFrame(index = names, columns = unit names) or
    None}, natural scale."""

    def __init__(self, sets):
        if isinstance(sets, dict):
            sets = [sets]
        if not sets:
            raise ValueError("at least one parameter set is required")
        self.sets = [dict(s) for s in sets]
        first = self.sets[0]
        sh, us = first.get("shared"), first.get("unit_specific")
        self.shared_names = [] if sh is None else list(sh.index)
        self.unit_param_names = [] if us is None else list(us.index)
        self.unit_names = None if us is None else list(us.columns)
        for s in self.sets:
            if (s.get("shared") is None and self.shared_names) or \
                    (s.get("shared") is not None and list(s["shared"].index) != self.shared_names):
                raise ValueError("every set must share the same shared-parameter names")
            if (s.get("unit_specific") is None and self.unit_param_names) or \
                    (s.get("unit_specific") is not None and list(s["unit_specific"].index) != self.unit_param_names):
                raise ValueError("every set must share the same unit-specific parameter names")
        self.logLik_unit = None

    def num_replicates(self):
        return len(self.sets)

    def arrays(self, unit_names):
        """shared (n, n_shared), unit (n, U, n_spec) float32."""
        n, U = len(self.sets), len(unit_names)
        sh = np.zeros((n, len(self.shared_names)), np.float32)
        un = np.zeros((n, U, len(self.unit_param_names)), np.float32)
        for i, s in enumerate(self.sets):
            if self.shared_names:
                sh[i] = s["shared"].iloc[:, 0].to_numpy(np.float32)
            if self.unit_param_names:
                un[i] = s["unit_specific"][list(unit_names)].to_numpy(np.float32).T
        return sh, un

    @classmethod
    def from_arrays(cls, shared, unit, shared_names, unit_param_names, unit_names):
        sets = []
        for i in range(shared.shape[0] if len(shared_names) else unit.shape[0]):
            s = {"shared": pd.DataFrame({"shared": shared[i]}, index=shared_names) if shared_names else None,
                 "unit_specific": pd.DataFrame(unit[i].T, index=unit_param_names, columns=unit_names)
                 if unit_param_names else None}
            sets.append(s)
        return cls(sets)


class PanelPomp:
    def __init__(self, Pomp_dict, theta):
        if not isinstance(theta, PanelParameters):
            raise TypeError("theta must be a PanelParameters instance")
        if not Pomp_dict:
            raise ValueError("Pomp_dict must hold at least one unit")
        self.unit_objects = dict(Pomp_dict)
        models = {type(p.model) for p in self.unit_objects.values()}
        self.model = next(iter(self.unit_objects.values())).model
        if len(models) != 1 or any(p.model.param_names != self.model.param_names
                                   for p in self.unit_objects.values()):
            raise ValueError("all units must use the same native model")
        self.theta = theta
        self.canonical_shared_param_names = list(theta.shared_names)
        self.canonical_unit_param_names = list(theta.unit_param_names)
        self.canonical_param_names = self.canonical_shared_param_names + self.canonical_unit_param_names
        if sorted(self.canonical_param_names) != sorted(self.model.param_names):
            raise ValueError("shared + unit-specific names must be the model's parameter names")
        if theta.unit_names is not None and sorted(theta.unit_names) != sorted(self.unit_objects):
            raise ValueError("unit-specific parameter columns must be the panel's unit names")
        self.fresh_key = None
        self.results_history = []

    def get_unit_names(self):
        return list(self.unit_objects)

    def to_struct(self):
        return F.PanelStruct(model=self.model, units=[p.series for p in self.unit_objects.values()],
                             shared_param_names=self.canonical_shared_param_names,
                             unit_param_names=self.canonical_unit_param_names,
                             unit_names=self.get_unit_names(),
                             par_trans=getattr(next(iter(self.unit_objects.values())), "par_trans", None))

    def _update_fresh_key(self, key):
        key = self.fresh_key if key is None else np.asarray(key, dtype=np.uint32)
        if key is None:
            raise ValueError("Both the key argument and the fresh_key attribute are None. "
                             "At least one key must be given.")
        self.fresh_key, new = prandom.split(key, 2)
        return new, key

    def pfilter(self, J, key=None, theta=None, thresh=0.0, reps=1, chunk_size=1, CLL=False, ESS=False,
                filter_mean=False, prediction_mean=False):
        start = time.time()
        if J < 1:
            raise ValueError("J should be greater than 0.")
        theta = self.theta if theta is None else theta
        new_key, old_key = self._update_fresh_key(key)
        units = self.get_unit_names()
        U, n = len(units), theta.num_replicates()
        sh, un = theta.arrays(units)
        panel = np.concatenate([np.broadcast_to(sh[:, None, :], (n, U, sh.shape[1])), un], axis=2)
        panel = np.repeat(panel, reps, axis=0)
        keys = prandom.split(new_key, n * reps * U).reshape(n * reps, U, 2)
        out = F.panel_pfilter(self.to_struct(), panel, J, keys, thresh=thresh, chunk_size=chunk_size, CLL=CLL,
                              ESS=ESS, filter_mean=filter_mean, prediction_mean=prediction_mean)
        ll = out["logLik64"].reshape(n, reps, U)
        theta.logLik_unit = np.stack([[logmeanexp(ll[i, :, u]) for u in range(U)] for i in range(n)])
        self.theta = theta
        self.results_history.append(dict(method="pfilter", theta=theta, logLik=ll, units=units, J=J, reps=reps,
                                          key=old_key, out=out, time=time.time() - start))

    def mif(self, J, M, rw_sd, key=None, theta=None, thresh=0.0, n_monitors=0, block=True, graph=False):
        start = time.time()
        if J < 1 or M < 1:
            raise ValueError("J and M must be greater than 0.")
        theta = self.theta if theta is None else theta
        units = self.get_unit_names()
        U, n = len(units), theta.num_replicates()
        sh, un = theta.arrays(units)
        sh_a = np.repeat(sh[:, None, :], J, axis=1)
        un_a = np.repeat(un[:, None, :, :], J, axis=1)
        new_key, old_key = self._update_fresh_key(key)
        keys = prandom.split(new_key, n)
        st, ut, fs, fu = F.panel_mif(self.to_struct(), sh_a, un_a, rw_sd, M, J, keys, thresh=thresh,
                                     n_monitors=n_monitors, block=block, graph=graph)
        self.theta = PanelParameters.from_arrays(st[:, -1, 1:], ut[:, -1, :, 1:], self.canonical_shared_param_names,
                                                 self.canonical_unit_param_names, units)
        self.theta.logLik_unit = ut[:, -1, :, 0].astype(float)
        self.results_history.append(dict(method="mif", theta=theta, shared_traces=st, unit_traces=ut, units=units,
                                          J=J, M=M, key=old_key, time=time.time() - start))

    def results(self, index=-1):
        """Tidy frame: one row per (parameter set, unit) plus a 'shared' row with
        the panel total; columns theta_idx, unit, logLik, se, then parameters."""
        r = self.results_history[index]
        rows = []
        theta = r["theta"]
        sh, un = theta.arrays(r["units"])
        for i in range(theta.num_replicates()):
            if r["method"] == "pfilter":
                per = [(logmeanexp(r["logLik"][i, :, u]), logmeanexp_se(r["logLik"][i, :, u]))
                       for u in range(len(r["units"]))]
            else:
                per = [(float(r["unit_traces"][i, -1, u, 0]), np.nan) for u in range(len(r["units"]))]
            total = float(np.sum([p[0] for p in per]))
            se = float(np.sqrt(np.sum([p[1] ** 2 for p in per]))) if r["method"] == "pfilter" else np.nan
            row = {"theta_idx": i, "unit": "shared", "logLik": total, "se": se}
            row.update({nm: float(sh[i, k]) for k, nm in enumerate(self.canonical_shared_param_names)})
            rows.append(row)
            for u, name in enumerate(r["units"]):
                row = {"theta_idx": i, "unit": name, "logLik": per[u][0], "se": per[u][1]}
                row.update({nm: float(un[i, u, k]) for k, nm in enumerate(self.canonical_unit_param_names)})
                rows.append(row)
        return pd.DataFrame(rows)

    def traces(self):
        """Long-format traces of every mif in the history: theta_idx, unit
        ('shared' for the panel total and shared parameters), iteration, method,
        logLik, se, parameters."""
        rows = []
        it0 = 0
        for r in self.results_history:
            if r["method"] != "mif":
                continue
            st, ut = r["shared_traces"], r["unit_traces"]
            for i in range(st.shape[0]):
                for m in range(st.shape[1]):
                    row = {"theta_idx": i, "unit": "shared", "iteration": it0 + m, "method": "mif",
                           "logLik": float(st[i, m, 0]), "se": np.nan}
                    row.update({nm: float(st[i, m, 1 + k]) for k, nm in enumerate(self.canonical_shared_param_names)})
                    rows.append(row)
                    for u, name in enumerate(r["units"]):
                        row = {"theta_idx": i, "unit": name, "iteration": it0 + m, "method": "mif",
                               "logLik": float(ut[i, m, u, 0]), "se": np.nan}
                        row.update({nm: float(ut[i, m, u, 1 + k])
                                    for k, nm in enumerate(self.canonical_unit_param_names)})
                        rows.append(row)
            it0 += st.shape[1] - 1
        return pd.DataFrame(rows)
