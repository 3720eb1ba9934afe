"""ctypes binding of the C restatement (oracle/cpu_pf.c) -- TEST INFRASTRUCTURE ONLY."""

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_lib" / "libcpupf.so"
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        vp = C.c_void_p
        L.cpu_lg_pfilter.argtypes = [vp, C.c_int, vp, C.c_int, C.c_int64, vp, C.c_int, vp]
        L.cpu_sv_mop.argtypes = [vp, vp, C.c_int, C.c_int64, C.c_float, vp, C.c_int, vp]
        L.cpu_sv_mop_grad.argtypes = [vp, vp, C.c_int, C.c_int64, C.c_float, vp, C.c_int, vp, vp]
        L.cpu_num_threads.restype = C.c_int
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def lg_pfilter(theta, dx, ys, J, keys):
    theta = np.ascontiguousarray(theta, np.float32)
    ys = np.ascontiguousarray(ys, np.float32)
    keys = np.ascontiguousarray(np.asarray(keys, np.uint32).reshape(-1, 2))
    out = np.zeros(keys.shape[0], np.float64)
    rc = lib().cpu_lg_pfilter(_p(theta), dx, _p(ys), ys.shape[0], J, _p(keys), keys.shape[0], _p(out))
    assert rc == 0
    return out


def sv_mop(theta_est, ys, J, alpha, keys):
    theta_est = np.ascontiguousarray(theta_est, np.float32)
    ys = np.ascontiguousarray(np.asarray(ys, np.float32).ravel())
    keys = np.ascontiguousarray(np.asarray(keys, np.uint32).reshape(-1, 2))
    out = np.zeros(keys.shape[0], np.float64)
    rc = lib().cpu_sv_mop(_p(theta_est), _p(ys), ys.shape[0], J, float(alpha), _p(keys), keys.shape[0], _p(out))
    assert rc == 0
    return out


def sv_mop_grad(theta_est, ys, J, alpha, keys):
    """(loglik (R,), grad (R, 6)) of the SV MOP by the C hand adjoint."""
    theta_est = np.ascontiguousarray(theta_est, np.float32)
    ys = np.ascontiguousarray(np.asarray(ys, np.float32).ravel())
    keys = np.ascontiguousarray(np.asarray(keys, np.uint32).reshape(-1, 2))
    out = np.zeros(keys.shape[0], np.float64)
    g = np.zeros((keys.shape[0], 6), np.float64)
    rc = lib().cpu_sv_mop_grad(_p(theta_est), _p(ys), ys.shape[0], J, float(alpha), _p(keys), keys.shape[0],
                               _p(out), _p(g))
    assert rc == 0
    return out, g


def num_threads():
    return int(lib().cpu_num_threads())
