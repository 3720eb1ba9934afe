"""MCMC proposals for ``pmcmc`` (pypomp/proposals.py): diagonal and full
covariance random walks and the Roberts & Rosenthal (2009) adaptive random
walk.  A proposal acts on the estimation-scale parameter vector of one chain
in float32 on the host; its normal draw ``z`` (d,) is ``normal(key, (d,))``,
drawn on the device for every iteration of every chain in one launch before
the chain runs (the key tree does not depend on the chain's decisions).

Interface: ``init_state(theta)``, ``step(state, theta, z, n, accepts) ->
(theta_proposed, new_state)`` and ``canonicalize(names)`` (expand and reorder
to the model's canonical parameter vector).
"""

import numpy as np

F32 = np.float32


def _matvec(L, z):
    """L (d, d) @ z (d,) in float32, sequential accumulation over columns."""
    d = z.shape[0]
    out = np.zeros(d, dtype=F32)
    for i in range(d):
        acc = F32(0)
        for j in range(d):
            acc = F32(acc + F32(L[i, j] * z[j]))
        out[i] = acc
    return out


def _expand(names, local_names, mat):
    idx = {n: i for i, n in enumerate(names)}
    for p in local_names:
        if p not in idx:
            raise ValueError(f"Proposal parameter {p!r} not in model.")
    pos = [idx[p] for p in local_names]
    full = np.zeros((len(names), len(names)))
    full[np.ix_(pos, pos)] = mat
    return full


class MVNDiagRW:
    """Diagonal random walk: theta + z * sd (proposals.py MVNDiagRW).
    Parameters with sd <= 0 are dropped."""

    def __init__(self, rw_sd):
        rw_sd = {k: float(v) for k, v in rw_sd.items() if v > 0}
        if not rw_sd:
            raise ValueError("rw_sd must contain at least one positive entry.")
        self.param_names = tuple(rw_sd)
        self.sd_arr = np.asarray([rw_sd[p] for p in self.param_names])

    def init_state(self, theta):
        return ()

    def step(self, state, theta, z, n, accepts):
        return (np.asarray(theta, F32) + (z * self.sd_arr.astype(F32)).astype(F32)).astype(F32), state

    def canonicalize(self, names):
        idx = {n: i for i, n in enumerate(names)}
        sd = np.zeros(len(names))
        for i, p in enumerate(self.param_names):
            if p not in idx:
                raise ValueError(f"Proposal parameter {p!r} not in model.")
            sd[idx[p]] = self.sd_arr[i]
        out = object.__new__(MVNDiagRW)
        out.param_names, out.sd_arr = tuple(names), sd
        return out


class MVNRWFull:
    """Full-covariance random walk: theta + chol(rw_var) z (proposals.py MVNRWFull)."""

    def __init__(self, rw_var, param_names):
        rw_var = np.asarray(rw_var, dtype=float)
        if rw_var.ndim != 2 or rw_var.shape[0] != rw_var.shape[1]:
            raise ValueError("rw_var must be a square matrix.")
        if rw_var.shape[0] != len(param_names):
            raise ValueError("rw_var dimensions must match len(param_names).")
        self.chol = np.linalg.cholesky(rw_var)
        self.param_names = tuple(param_names)

    def init_state(self, theta):
        return ()

    def step(self, state, theta, z, n, accepts):
        return (np.asarray(theta, F32) + _matvec(self.chol.astype(F32), z)).astype(F32), state

    def canonicalize(self, names):
        out = object.__new__(MVNRWFull)
        out.param_names, out.chol = tuple(names), _expand(names, self.param_names, self.chol)
        return out


class MVNRWAdaptive:
    """Adaptive random walk (Roberts & Rosenthal 2009; proposals.py
    MVNRWAdaptive).  Phase 1 (n >= scale_start, accepts < shape_start) scales
    the initial covariance toward the target acceptance rate; phase 2
    (accepts >= shape_start) uses (2.38^2 / d) times the running empirical
    covariance (Welford).  Parameters outside the initial covariance's support
    are not moved."""

    def __init__(self, rw_sd=None, rw_var=None, param_names=None, scale_start=200, scale_cooling=0.999,
                 shape_start=200, target=0.234, max_scaling=50.0):
        if (rw_sd is None) == (rw_var is None):
            raise ValueError("Exactly one of rw_sd and rw_var must be given.")
        if rw_sd is not None:
            rw_sd = {k: float(v) for k, v in rw_sd.items() if v > 0}
            names = tuple(rw_sd)
            init_var = np.diag([rw_sd[p] ** 2 for p in names])
        else:
            if param_names is None:
                raise ValueError("param_names required when rw_var is given.")
            init_var = np.asarray(rw_var, dtype=float)
            if init_var.shape != (len(param_names), len(param_names)):
                raise ValueError("rw_var shape must match param_names.")
            names = tuple(param_names)
        if scale_start < 1:
            raise ValueError("scale_start must be a positive integer.")
        if not (0 < scale_cooling <= 1):
            raise ValueError("scale_cooling must be in (0, 1].")
        if shape_start < 1:
            raise ValueError("shape_start must be a positive integer.")
        if not (0 < target < 1):
            raise ValueError("target must be in (0, 1).")
        self.init_rw_var = init_var
        self.param_names = names
        self.scale_start = int(scale_start)
        self.scale_cooling = float(scale_cooling)
        self.shape_start = int(shape_start)
        self.target = float(target)
        self.max_scaling = float(max_scaling)

    def init_state(self, theta):
        d = np.asarray(theta).shape[-1]
        return {"scaling": F32(1), "theta_mean": np.zeros(d, F32), "covmat_emp": np.zeros((d, d), F32),
                "initialized": False}

    def step(self, state, theta, z, n, accepts):
        theta = np.asarray(theta, F32)
        d = theta.shape[-1]
        nf = F32(max(n, 1))
        af = F32(accepts)
        mean0 = state["theta_mean"] if state["initialized"] else theta
        in1 = (n >= self.scale_start) and (accepts < self.shape_start)
        in2 = accepts >= self.shape_start
        scaling = state["scaling"]
        if in1:
            cool = np.power(F32(self.scale_cooling), F32(max(n - self.scale_start, 0)), dtype=F32)
            factor = np.exp(F32(cool * F32(F32(af / nf) - F32(self.target))), dtype=F32)
            scaling = F32(min(F32(scaling * factor), F32(self.max_scaling)))
        init_var = self.init_rw_var.astype(F32)
        if in2:
            cov = (F32(2.38 ** 2 / d) * state["covmat_emp"]).astype(F32)
        elif in1:
            cov = (F32(scaling * scaling) * init_var).astype(F32)
        else:
            cov = init_var
        new_mean = (mean0 + ((theta - mean0) / nf).astype(F32)).astype(F32)
        diff_old = (theta - mean0).astype(F32)
        diff_new = (theta - new_mean).astype(F32)
        new_cov = (((nf - F32(1)) * state["covmat_emp"]).astype(F32)
                   + np.outer(diff_old, diff_new).astype(F32)).astype(F32) / nf
        active = (np.any(self.init_rw_var != 0, axis=0) | np.any(self.init_rw_var != 0, axis=1)).astype(F32)
        chol = np.linalg.cholesky((cov + F32(1e-10) * np.eye(d, dtype=F32)).astype(F32)).astype(F32)
        prop = (theta + (_matvec(chol, z) * active).astype(F32)).astype(F32)
        return prop, {"scaling": scaling, "theta_mean": new_mean, "covmat_emp": new_cov.astype(F32),
                      "initialized": True}

    def canonicalize(self, names):
        out = object.__new__(MVNRWAdaptive)
        out.__dict__.update(self.__dict__)
        out.param_names = tuple(names)
        out.init_rw_var = _expand(names, self.param_names, self.init_rw_var)
        return out
