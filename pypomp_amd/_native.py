"""ctypes binding of libmipomp.so (the C ABI declared in include/mipomp.h).

The library is loaded from ``pypomp_amd/_lib/libmipomp.so``; if it is missing
the import of any compute entry point raises -- there is no CPU fallback.
Device buffers are torch tensors on the ROCm device; their data pointers and
the current HIP stream are handed to the library as plain integers.
"""

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libmipomp.so"

MP_OK = 0
MP_MODEL_LG, MP_MODEL_SEIR, MP_MODEL_SV, MP_MODEL_SPX = 1, 2, 3, 4
MP_ALGO_PF, MP_ALGO_MOP, MP_ALGO_PER = 0, 1, 2
MP_FORM_BATCHED, MP_FORM_PLAIN = 0, 1
MP_CLL, MP_ESS, MP_FILTER_MEAN, MP_PREDICTION_MEAN, MP_SHOULD_TRANS = 1, 2, 4, 8, 16

EXPORTS = [
    "mp_last_error", "mp_abi_version", "mp_host_threefry2x32", "mp_host_split",
    "mp_random_bits", "mp_uniform", "mp_normal", "mp_split",
    "mp_resample_workspace_size", "mp_resample_systematic",
    "mp_filter_workspace_size", "mp_pfilter", "mp_pfilter_noalloc", "mp_mop", "mp_perfilter",
    "mp_mop_grad_workspace_size", "mp_mop_grad", "mp_mop_hess_workspace_size", "mp_mop_hess", "mp_simulate", "mp_host_from_est",
    "mp_profile_enable", "mp_profile_read", "mp_tune_set", "mp_tune_get", "mp_valu_probe",
]
MP_TUNE_PLAIN_SCAN_ROWS, MP_TUNE_PLAIN_PROP_ROWS, MP_TUNE_BATCHED_BLOCK, MP_TUNE_GRAD_BLOCK = 1, 2, 3, 4
MP_TUNE_PLAIN_SCAN_THREADS, MP_TUNE_PLAIN_PROP_THREADS, MP_TUNE_PLAIN_LDS_TILES = 5, 6, 7
MP_TUNE_PLAIN_XCD_MAP = 8
MP_TUNE_PLAIN_LANES = 9
MP_NOALLOC_SCRATCH_BYTES = 64 * 1024 * 1024


class ModelDesc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("S", C.c_int32), ("d", C.c_int32), ("dy", C.c_int32),
                ("dx", C.c_int32), ("lg_sep_init", C.c_int32), ("accum_mask", C.c_uint32),
                ("identity_trans", C.c_int32), ("vectorized", C.c_int32)]


class SeriesDesc(C.Structure):
    _fields_ = [("ys", C.c_void_p), ("times", C.c_void_p), ("nstep", C.c_void_p),
                ("tidx0", C.c_void_p), ("dt_ext", C.c_void_p), ("obs_off", C.c_void_p),
                ("dt_off", C.c_void_p), ("t0", C.c_float), ("U", C.c_int32),
                ("Tmax", C.c_int32), ("max_euler", C.c_int64), ("covars", C.c_void_p),
                ("cv_off", C.c_void_p), ("dc", C.c_int32), ("max_nstep", C.c_int32)]


class NativeError(RuntimeError):
    pass


_lib = None


def lib():
    """Load libmipomp.so once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise NativeError(
            f"{LIB_PATH} is missing: build it with `python -m pypomp_amd.build` "
            "(there is no CPU fallback for the particle-filter engine)")
    # torch first: its bundled HIP runtime then satisfies the library's
    # libamdhip64.so.7 dependency.  Loaded the other way round, the process
    # would map two HIP/HSA runtimes and the library's would see no device.
    import torch  # noqa: F401
    L = C.CDLL(str(LIB_PATH))
    vp, i32, i64, u32p, f32p, f64p, sz = (C.c_void_p, C.c_int32, C.c_int64, C.c_void_p,
                                           C.c_void_p, C.c_void_p, C.c_size_t)
    L.mp_last_error.restype = C.c_char_p
    L.mp_abi_version.restype = i32
    L.mp_host_threefry2x32.argtypes = [vp, vp, vp, i64]
    L.mp_host_split.argtypes = [vp, i64, vp]
    for n in ("mp_random_bits", "mp_uniform", "mp_normal", "mp_split"):
        getattr(L, n).argtypes = [u32p, i64, i64, vp, vp]
    L.mp_resample_workspace_size.argtypes = [i64, i64]
    L.mp_resample_workspace_size.restype = sz
    L.mp_resample_systematic.argtypes = [f32p, u32p, i64, i64, vp, vp, sz, vp]
    L.mp_filter_workspace_size.argtypes = [i32, C.POINTER(ModelDesc), C.POINTER(SeriesDesc), i32, i32, i32]
    L.mp_filter_workspace_size.restype = sz
    L.mp_pfilter.argtypes = [C.POINTER(ModelDesc), C.POINTER(SeriesDesc), f32p, i64, u32p, vp, i32, i32,
                             C.c_float, i32, i32, f64p, f32p, f32p, f32p, f32p, vp, sz, vp]
    L.mp_pfilter_noalloc.argtypes = [C.POINTER(ModelDesc), C.POINTER(SeriesDesc), f32p, i64, u32p, vp, i32, i32,
                                     C.c_float, i32, i32, f64p, f32p, f32p, f32p, f32p, vp]
    L.mp_mop.argtypes = [C.POINTER(ModelDesc), C.POINTER(SeriesDesc), f32p, i64, u32p, vp, i32, i32,
                         C.c_float, i32, f64p, vp, sz, vp]
    L.mp_perfilter.argtypes = [C.POINTER(ModelDesc), C.POINTER(SeriesDesc), f32p, u32p, vp, i32, i32,
                               C.c_float, f32p, f32p, C.c_float, f32p, i32, f32p, f32p, f64p, vp, vp, sz, vp]
    L.mp_mop_grad_workspace_size.argtypes = [C.POINTER(ModelDesc), C.POINTER(SeriesDesc), i32, i32, i32]
    L.mp_mop_grad_workspace_size.restype = sz
    L.mp_mop_grad.argtypes = [C.POINTER(ModelDesc), C.POINTER(SeriesDesc), f32p, i64, u32p, vp, i32, i32,
                              C.c_float, i32, i32, f64p, f32p, vp, sz, vp]
    L.mp_mop_hess_workspace_size.argtypes = [C.POINTER(ModelDesc), C.POINTER(SeriesDesc), i32, i32]
    L.mp_mop_hess_workspace_size.restype = sz
    L.mp_mop_hess.argtypes = [C.POINTER(ModelDesc), C.POINTER(SeriesDesc), f32p, i64, u32p, i32, i32, C.c_float,
                              f64p, f32p, f32p, vp, sz, vp]
    L.mp_simulate.argtypes = [C.POINTER(ModelDesc), C.POINTER(SeriesDesc), f32p, i64, u32p, vp, i32, i32, i32,
                              f32p, f32p, vp]
    L.mp_host_from_est.argtypes = [C.POINTER(ModelDesc), vp, i64]
    L.mp_tune_set.argtypes = [i32, i32]
    L.mp_tune_get.argtypes = [i32]
    L.mp_profile_enable.argtypes = [i32]
    L.mp_profile_read.argtypes = [vp, vp, i32]
    L.mp_valu_probe.argtypes = [i32, i64, i32, vp, vp]
    for n in EXPORTS:
        fn = getattr(L, n)
        if fn.restype is C.c_int or fn.restype is None:
            if n not in ("mp_last_error",):
                fn.restype = i32
    L.mp_last_error.restype = C.c_char_p
    L.mp_resample_workspace_size.restype = sz
    L.mp_filter_workspace_size.restype = sz
    L.mp_mop_grad_workspace_size.restype = sz
    L.mp_mop_hess_workspace_size.restype = sz
    _lib = L
    return L


def check(rc):
    if rc != MP_OK:
        msg = lib().mp_last_error().decode(errors="replace")
        raise NativeError(f"libmipomp error {rc}: {msg}")


def ptr(t):
    """Raw device (or host) pointer of a torch tensor, or None."""
    return None if t is None else C.c_void_p(t.data_ptr())


def stream_ptr():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def require_device():
    """The engine runs on the GPU only; fail loudly otherwise."""
    import torch
    if not torch.cuda.is_available():
        raise NativeError("no ROCm GPU is visible: the particle-filter engine runs on the GPU only")
    lib()
    return torch.device("cuda", torch.cuda.current_device())


PROFILE_KINDS = ("plain_prop", "plain_scan", "batched", "sched", "valu_probe")


def profile_enable(on=True):
    check(lib().mp_profile_enable(1 if on else 0))


def profile_read():
    """{kind: (total_ms, launches)} of the launches recorded since the last read."""
    import numpy as np
    ms = np.zeros(len(PROFILE_KINDS), dtype=np.float64)
    cnt = np.zeros(len(PROFILE_KINDS), dtype=np.int64)
    check(lib().mp_profile_read(ms.ctypes.data_as(C.c_void_p), cnt.ctypes.data_as(C.c_void_p),
                                len(PROFILE_KINDS)))
    return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(PROFILE_KINDS)}


def env_flag(name):
    return os.environ.get(name, "") not in ("", "0")
