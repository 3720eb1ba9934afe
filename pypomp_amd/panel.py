"""``PanelPomp``: the object API of panels (pypomp/panel/panel.py,
panel/estimation_mixin.py:384-773, panel/analysis_mixin.py:74-200).

A panel holds unit ``Pomp`` objects of one native model with shared and
unit-specific parameters (``PanelParameters``, one or more replicate sets).
``pfilter`` runs every (set, replicate, unit) filter in one launch; ``mif`` is
panel IF2 (functional.panel_mif); ``train`` is panel MOP gradient ascent
(functional.panel_train).  Keys follow the reference: ``fresh_key``-derived
key split into n_sets * reps * U filter keys for pfilter, into n_sets chain
keys for mif and into n_sets * M * U MOP keys for train.
"""

import copy
import time
import warnings

import numpy as np
import pandas as pd

from . import functional as F
from . import random as prandom
from .maths import logmeanexp, logmeanexp_se
from .pomp import _same, _same_history


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

    def __eq__(self, other):
        return (isinstance(other, PanelParameters) and self.shared_names == other.shared_names
                and self.unit_param_names == other.unit_param_names and self.unit_names == other.unit_names
                and _same(self.sets, other.sets) and _same(self.logLik_unit, other.logLik_unit))

    __hash__ = None

    @classmethod
    def merge(cls, *param_objs):
        """The replicate sets of several parameter objects, in order
        (core/parameters/base.py:205-240)."""
        if len(param_objs) == 0:
            raise ValueError("At least one PanelParameters object must be provided.")
        first = param_objs[0]
        for p in param_objs:
            if not isinstance(p, cls):
                raise TypeError("All merged objects must be of type PanelParameters.")
            if (p.shared_names, p.unit_param_names, p.unit_names) != \
                    (first.shared_names, first.unit_param_names, first.unit_names):
                raise ValueError("All PanelParameters objects must have the same parameter and unit names.")
        out = cls([dict(s) for p in param_objs for s in p.sets])
        if all(p.logLik_unit is not None for p in param_objs):
            out.logLik_unit = np.concatenate([np.asarray(p.logLik_unit) for p in param_objs], axis=0)
        return out

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


def _merge_panel_results(entries):
    """Same-index panel results entries joined along the parameter-set axis;
    the configuration must agree (core/results/result.py:178-213)."""
    first = entries[0]
    per_set = ("theta", "logLik", "out", "shared_traces", "unit_traces")
    for e in entries:
        if e["method"] != first["method"]:
            raise ValueError("All merged results must share method/kind/panel.")
        if e.keys() != first.keys():
            raise ValueError("All merged results must have the same fields.")
        for k in first:
            if k not in per_set and k not in ("key", "time") and not _same(e[k], first[k]):
                raise ValueError(f"All merged results must have the same {k}.")
    out = dict(first)
    for k in first:
        if k == "theta":
            out[k] = PanelParameters.merge(*[e[k] for e in entries])
        elif k == "out":
            out[k] = {n: np.concatenate([np.asarray(e[k][n]) for e in entries], axis=0) for n in first[k]}
        elif k in per_set:
            out[k] = np.concatenate([np.asarray(e[k]) for e in entries], axis=0)
    if "time" in first:
        out["time"] = max(e["time"] for e in entries)
    return out


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

    # -- equality, merge and pickling (panel/panel.py:171-300) --------------
    def __eq__(self, other):
        if not isinstance(other, type(self)):
            return False
        if (self.canonical_param_names, self.canonical_shared_param_names, self.canonical_unit_param_names) != \
                (other.canonical_param_names, other.canonical_shared_param_names, other.canonical_unit_param_names):
            return False
        if self.theta != other.theta or self.get_unit_names() != other.get_unit_names():
            return False
        if any(self.unit_objects[u] != other.unit_objects[u] for u in self.unit_objects):
            return False
        if not _same_history(self.results_history, other.results_history):
            return False
        if (self.fresh_key is None) != (other.fresh_key is None):
            return False
        return self.fresh_key is None or _same(np.asarray(self.fresh_key, np.uint32),
                                               np.asarray(other.fresh_key, np.uint32))

    __hash__ = None

    @staticmethod
    def merge(*panel_pomp_objs):
        """One panel holding the parameter sets and results of several panels
        with the same units and parameter names (panel/panel.py:230-268)."""
        if len(panel_pomp_objs) == 0:
            raise ValueError("At least one PanelPomp object must be provided.")
        first = panel_pomp_objs[0]
        for obj in panel_pomp_objs:
            if not isinstance(obj, type(first)):
                raise TypeError("All merged objects must be of type PanelPomp.")
            if obj.canonical_param_names != first.canonical_param_names:
                raise ValueError("All PanelPomp objects must have the same canonical_param_names.")
            if obj.canonical_shared_param_names != first.canonical_shared_param_names:
                raise ValueError("All PanelPomp objects must have the same canonical_shared_param_names.")
            if obj.canonical_unit_param_names != first.canonical_unit_param_names:
                raise ValueError("All PanelPomp objects must have the same canonical_unit_param_names.")
            if obj.get_unit_names() != first.get_unit_names():
                raise ValueError("All PanelPomp objects must have the same unit names.")
            if len(obj.results_history) != len(first.results_history):
                raise ValueError("All histories must have the same number of entries to be merged.")
        merged = copy.deepcopy(first)
        merged.theta = PanelParameters.merge(*[obj.theta for obj in panel_pomp_objs])
        merged.results_history = [_merge_panel_results([obj.results_history[i] for obj in panel_pomp_objs])
                                  for i in range(len(first.results_history))]
        merged.fresh_key = None if first.fresh_key is None else np.array(first.fresh_key, np.uint32)
        return merged

    def __getstate__(self):
        state = self.__dict__.copy()
        key = state.pop("fresh_key", None)
        if key is not None:
            state["_fresh_key_data"] = np.array(key, dtype=np.uint32)
        return state

    def __setstate__(self, state):
        state = dict(state)
        data = state.pop("_fresh_key_data", None)
        self.__dict__.update(state)
        if data is not None:
            arr = np.asarray(data)
            if arr.shape == (2,) and arr.dtype.kind in "ui" and not np.any(arr < 0) and not np.any(arr > 0xFFFFFFFF):
                self.fresh_key = arr.astype(np.uint32)
            else:
                warnings.warn(f"Failed to reconstruct fresh_key: shape {arr.shape} dtype {arr.dtype}",
                              UserWarning, stacklevel=2)
                self.fresh_key = None
        elif "fresh_key" not in self.__dict__:
            self.fresh_key = None

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

    def train(self, J, M, eta, chunk_size=1, optimizer=None, alpha=0.97, key=None, theta=None, alpha_cooling=1.0):
        """Panel MOP gradient ascent (panel/estimation_mixin.py:775-990): keys
        split(key, n_sets M U) as (n_sets, M, U), F.panel_train, traces with the
        panel objective's log-likelihood in the shared traces' column 0 and NaN
        in the unit traces' column 0; theta becomes the final values."""
        from .optim import Adam, LearningRate
        start = time.time()
        optimizer = optimizer or Adam()
        theta = self.theta if theta is None else theta
        n = theta.num_replicates()
        new_key, old_key = self._update_fresh_key(key)
        if J < 1 or M < 1:
            raise ValueError("J and M must be greater than 0.")
        if not isinstance(eta, LearningRate):
            raise TypeError("eta must be a LearningRate object")
        units = self.get_unit_names()
        U = len(units)
        chunk_size = max(1, int(chunk_size))
        sh, un = theta.arrays(units)
        keys = prandom.split(new_key, n * M * U).reshape(n, M, U, 2)
        nll, sh_h, un_h = F.panel_train(self.to_struct(), sh, un, J, optimizer, M, eta, alpha, keys,
                                        alpha_cooling=alpha_cooling, chunk_size=chunk_size)
        st = np.concatenate([-nll[:, :, None], sh_h], axis=-1)
        ut = np.concatenate([np.full((n, M + 1, U, 1), np.nan, np.float32), un_h], axis=-1)
        self.theta = PanelParameters.from_arrays(sh_h[:, -1], un_h[:, -1], self.canonical_shared_param_names,
                                                 self.canonical_unit_param_names, units)
        self.theta.logLik_unit = np.full((n, U), np.nan)
        self.results_history.append(dict(method="train", theta=theta, shared_traces=st, unit_traces=ut, units=units,
                                          J=J, M=M, eta=eta, optimizer=optimizer, alpha=alpha,
                                          alpha_cooling=alpha_cooling, chunk_size=chunk_size, key=old_key,
                                          time=time.time() - start))

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
        """Long-format traces of every mif and train in the history: theta_idx,
        unit ('shared' for the panel total and shared parameters), iteration,
        method, logLik, se, parameters."""
        rows = []
        it0 = 0
        for r in self.results_history:
            if r["method"] not in ("mif", "train"):
                continue
            st, ut = r["shared_traces"], r["unit_traces"]
            for i in range(st.shape[0]):
                for m in range(st.shape[1]):
                    row = {"theta_idx": i, "unit": "shared", "iteration": it0 + m, "method": r["method"],
                           "logLik": float(st[i, m, 0]), "se": np.nan}
                    row.update({nm: float(st[i, m, 1 + k]) for k, nm in enumerate(self.canonical_shared_param_names)})
                    rows.append(row)
                    for u, name in enumerate(r["units"]):
                        row = {"theta_idx": i, "unit": name, "iteration": it0 + m, "method": r["method"],
                               "logLik": float(ut[i, m, u, 0]), "se": np.nan}
                        row.update({nm: float(ut[i, m, u, 1 + k])
                                    for k, nm in enumerate(self.canonical_unit_param_names)})
                        rows.append(row)
            it0 += st.shape[1] - 1
        return pd.DataFrame(rows)
