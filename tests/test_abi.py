"""The C ABI library loads, exports every symbol include/mipomp.h declares,
and its host-side functions and argument checks work without a GPU."""

import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from oracle import threefry as tf

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "mipomp.h").read_text()
    return sorted(set(re.findall(r"^[A-Za-z_][\w\s\*]*?\b(mp_\w+)\s*\(", text, flags=re.M)))


@pytest.fixture(scope="module")
def lib():
    from pypomp_amd import _native
    return _native.lib()


def test_header_declares_the_expected_entry_points():
    syms = declared_symbols()
    for s in ("mp_pfilter", "mp_mop", "mp_perfilter", "mp_resample_systematic", "mp_normal", "mp_host_split"):
        assert s in syms


def test_library_exports_every_declared_symbol(lib):
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_python_binding_lists_every_declared_symbol():
    from pypomp_amd import _native
    assert sorted(_native.EXPORTS) == declared_symbols()


def test_abi_version(lib):
    assert lib.mp_abi_version() == 4


def test_host_split_matches_oracle():
    import pypomp_amd.random as prandom
    for seed in (0, 1, 20260811, 2 ** 40 + 3):
        assert np.array_equal(prandom.split(prandom.key(seed), 17), tf.split(tf.key(seed), 17))


def test_host_threefry_known_answer():
    import pypomp_amd.random as prandom
    out = prandom.threefry2x32([[0x13198A2E, 0x03707344]], [[0x243F6A88, 0x85A308D3]])
    assert out.tolist() == [[0xC4923A9C, 0x483DF7A0]]


def test_argument_errors_are_reported_without_touching_the_device(lib):
    from pypomp_amd import _native
    md = _native.ModelDesc(kind=99, S=1, d=1, dy=1, dx=1, lg_sep_init=0, accum_mask=0)
    sd = _native.SeriesDesc()
    rc = lib.mp_pfilter(md, sd, None, 0, None, None, 1, 10, 0.0, 0, 0, None, None, None, None, None,
                        None, 0, None)
    assert rc == 1
    assert b"unknown model kind" in lib.mp_last_error()
    md = _native.ModelDesc(kind=1, S=2, d=19, dy=2, dx=2, lg_sep_init=1, accum_mask=0)
    rc = lib.mp_mop(md, None, None, 0, None, None, 1, 10, 0.5, 0, None, None, 0, None)
    assert rc == 1 and b"series" in lib.mp_last_error()
    rc = lib.mp_resample_systematic(None, None, 1, 0, None, None, 0, None)
    assert rc == 1


def test_tuning_settings_round_trip_and_reject_bad_values(lib):
    from pypomp_amd import autotune
    for name, (pid, cands) in autotune.PARAMS.items():
        old = autotune.get(name)
        for v in cands:
            autotune.set(name, v)
            assert autotune.get(name) == v
        autotune.set(name, old)
    assert lib.mp_tune_set(3, 300) != 0
    # the plain-form shapes: 0 is automatic (the default), other values are checked
    for knob in (1, 2, 5, 6):
        old = lib.mp_tune_get(knob)
        assert old == 0
        assert lib.mp_tune_set(knob, 0) == 0 and lib.mp_tune_get(knob) == 0
    assert lib.mp_tune_set(1, 3) != 0 and lib.mp_tune_set(2, 17) != 0
    assert lib.mp_tune_set(5, 128) != 0 and lib.mp_tune_set(6, 768) != 0
    assert lib.mp_tune_set(99, 1) != 0
    assert lib.mp_tune_set(7, -1) != 0
    old = lib.mp_tune_get(7)
    assert lib.mp_tune_set(7, 0) == 0 and lib.mp_tune_get(7) == 0
    assert lib.mp_tune_set(7, old) == 0
    assert lib.mp_tune_get(8) == 1 and lib.mp_tune_set(8, 2) != 0
    assert lib.mp_tune_set(8, 0) == 0 and lib.mp_tune_get(8) == 0
    assert lib.mp_tune_set(8, 1) == 0
    # plain-form lanes: 1..4, default 4
    assert lib.mp_tune_get(9) == 4
    assert lib.mp_tune_set(9, 0) != 0 and lib.mp_tune_set(9, 5) != 0 and lib.mp_tune_get(9) == 4


def test_library_shares_the_one_hip_runtime_in_the_process():
    # a host call first (key splitting) loads the library; torch's HIP runtime
    # must then be the only one mapped, or the library would see no device
    import subprocess
    import sys
    from pathlib import Path
    code = ("from pypomp_amd import random as r; r.split(r.key(1), 2); import torch; "
            "m = open('/proc/self/maps').read().splitlines(); "
            "print(len({l.split()[-1] for l in m if 'libamdhip64' in l}))")
    out = subprocess.run([sys.executable, "-c", code], cwd=Path(__file__).resolve().parents[1],
                         capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip() == "1"


def test_mop_grad_rejects_a_grid_with_several_steps_in_one_interval(lib):
    # nstep [0, 2, 1] has as many Euler steps as intervals, but the adjoint
    # needs exactly one per interval: the per-interval maximum catches it
    # before anything is launched (the pointers are never dereferenced)
    from pypomp_amd import _native
    fake = C.c_void_p(256)
    md = _native.ModelDesc(kind=_native.MP_MODEL_SV, S=2, d=6, dy=1)
    sd = _native.SeriesDesc(ys=fake, times=fake, nstep=fake, tidx0=fake, dt_ext=fake, obs_off=fake, dt_off=fake,
                            t0=0.0, U=1, Tmax=3, max_euler=3, max_nstep=2)
    rc = lib.mp_mop_grad(md, sd, fake, 6, fake, None, 4, 100, 0.9, 10, 0, fake, fake, fake, 1 << 20, None)
    assert rc == 1
    assert b"one Euler step per observation interval" in lib.mp_last_error()
    md_spx = _native.ModelDesc(kind=_native.MP_MODEL_SPX, S=2, d=6, dy=1, vectorized=1)
    sd.max_nstep = 1
    rc = lib.mp_pfilter(md_spx, sd, fake, 6, fake, None, 1, 10, 0.0, 0, 0, fake, None, None, None, None, fake,
                        1 << 20, None)
    assert rc == 1 and b"reads 1 covariates per row, the series has 0" in lib.mp_last_error()
    md_sv = _native.ModelDesc(kind=_native.MP_MODEL_SV, S=2, d=6, dy=1, vectorized=1)
    rc = lib.mp_pfilter(md_sv, sd, fake, 6, fake, None, 1, 10, 0.0, 0, 0, fake, None, None, None, None, fake,
                        1 << 20, None)
    assert rc == 1 and b"no @vectorized rproc" in lib.mp_last_error()
