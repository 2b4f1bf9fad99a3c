"""Build libdbsr3_amd.so in-tree with hipcc for gfx950.

Usage: python -m dbsr3_amd.build   (or dbsr3_amd.build.build())
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
LIB = os.path.join(HERE, "libdbsr3_amd.so")
OBJDIR = os.path.join(HERE, "_build")

SOURCES = ["dbs_context.hip", "dbs_plain.hip", "dbs_engine.hip", "dbs_accum.hip", "dbs_updates.hip", "dbs_mat3.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # matrix-core accumulators stay in VGPRs: no AGPR copies and result
    # waits around every v_mfma in the operand loops
    "-mllvm", "-amdgpu-mfma-vgpr-form",
    "-I" + INCLUDE, "-I" + CSRC, "-Wall", "-Wno-unused-function",
]


def _compile(src):
    obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    return obj


def build(verbose=False):
    os.makedirs(OBJDIR, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=min(4, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr[-4000:])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_fortran(verbose)
    return LIB


FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
FSRC = os.path.join(HERE, "fortran", "dbs_gpu.f90")
FLIB = os.path.join(HERE, "libdbsr3_amd_fortran.so")
DRIVER_SRC = os.path.join(ROOT, "tests", "fortran", "drv_dropin.f90")
DRIVER = os.path.join(ROOT, "tests", "fortran", "drv_dropin")


def build_fortran(verbose=False):
    """the ISO_C_BINDING drop-in layer (libdbsr3_amd_fortran.so) and the
    Fortran host program used by the drop-in test; skipped without flang"""
    if not os.path.exists(FLANG):
        if verbose:
            print("flang not found: Fortran drop-in layer not built")
        return None
    mod = os.path.join(OBJDIR, "fmod")
    os.makedirs(mod, exist_ok=True)
    obj = os.path.join(OBJDIR, "dbs_gpu.o")
    steps = [
        [FLANG, "-O2", "-fPIC", "-module-dir", mod, "-c", FSRC, "-o", obj],
        [FLANG, "-shared", "-fPIC", obj, "-L" + HERE, "-ldbsr3_amd",
         "-Wl,-rpath,$ORIGIN", "-o", FLIB],
        [FLANG, "-O2", DRIVER_SRC, "-I" + mod, "-L" + HERE, "-ldbsr3_amd_fortran", "-ldbsr3_amd",
         "-Wl,-rpath,$ORIGIN/../../dbsr3_amd", "-o", DRIVER],
    ]
    for cmd in steps:
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("flang failed: " + " ".join(cmd) + "\n" + r.stderr[-3000:])
    if verbose:
        print("built", FLIB, "and", DRIVER)
    return FLIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
