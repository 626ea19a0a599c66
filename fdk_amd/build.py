"""Build recipe for the fdk_amd native library (gfx950) and its C++ drivers.

``python -m fdk_amd.build`` (or ``__graft_entry__.build()``) compiles every
source under ``fdk_amd/csrc`` with hipcc into ``fdk_amd/libfdk_amd.so`` and the
C++ drop-in driver ``fdk_amd/bin/fdk_drivers``.  It always recompiles.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
OBJ = os.path.join(HERE, "_obj")
LIB = os.path.join(HERE, "libfdk_amd.so")
BIN = os.path.join(HERE, "bin")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("FDK_OFFLOAD_ARCH", "gfx950")
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
          f"-I{INCLUDE}", f"-I{CSRC}"]
# Packed FP32 math is switched off for all device code (a pool-wide rule for
# these boxes: hipcc -O3 forms v_pk_* FP32 ops on its own otherwise).
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
          "-packed-fp32-ops", "-munsafe-fp-atomics"]

SOURCES = ["capi.cpp", "geometry.cpp", "bp.hip", "filter.hip", "fp.hip", "nn.hip", "pipeline.cpp",
           "unet.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return r


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON + DEVICE + ["-x", "hip" if src.endswith(".hip") else "c++",
                                       "-c", path, "-o", obj]
    if not src.endswith(".hip"):
        cmd = [HIPCC] + COMMON + ["-c", path, "-o", obj]
    _run(cmd)
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(BIN, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-lz"])
    os.replace(tmp, LIB)
    drv_src = os.path.join(CSRC, "drivers.cpp")
    if os.path.exists(drv_src):
        _run([HIPCC] + COMMON + [drv_src, "-o", os.path.join(BIN, "fdk_drivers"),
                                 f"-L{HERE}", "-lfdk_amd", f"-Wl,-rpath,{HERE}"])
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
