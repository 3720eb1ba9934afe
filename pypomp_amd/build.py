"""Build libmipomp.so (all HIP sources, gfx950) in-tree with hipcc.

Run as ``python -m pypomp_amd.build``; ``__graft_entry__.build()`` calls
``build()``.  The library lands in ``pypomp_amd/_lib/libmipomp.so``.
"""

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OUT_DIR = HERE / "_lib"
LIB = OUT_DIR / "libmipomp.so"
SOURCES = ["capi.hip", "prims.hip", "batched.hip", "plain.hip", "mopgrad.hip", "mophess.hip", "simulate.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    "-munsafe-fp-atomics",
    # no fused multiply-add contraction: the state/weight path is a fixed
    # sequence of correctly rounded operations (reproducible by the oracle)
    "-ffp-contract=off",
    # packed FP32 ops stay off: the compiler must not form v_pk_* f32 ops
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src, obj):
    cmd = [hipcc(), *HIPCC_FLAGS, "-c", str(CSRC / src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return r.stderr


def build(verbose=False):
    OUT_DIR.mkdir(exist_ok=True)
    objs = [OUT_DIR / (Path(s).stem + ".o") for s in SOURCES]
    with ThreadPoolExecutor(max_workers=min(7, len(SOURCES))) as ex:
        logs = list(ex.map(_compile, SOURCES, objs))
    if verbose:
        for s, log in zip(SOURCES, logs):
            if log.strip():
                print(f"[{s}]\n{log}", file=sys.stderr)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    for o in objs:
        o.unlink(missing_ok=True)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
