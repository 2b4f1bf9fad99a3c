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

SOURCES = ["dbs_context.hip", "dbs_plain.hip", "dbs_engine.hip", "dbs_accum.hip", "dbs_updates.hip", "dbs_mat3.hip", "dbs_channels.hip"]

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
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-L/opt/rocm/lib", "-lrocsolver", "-lrocblas",
           "-Wl,-rpath,/opt/rocm/lib", "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr[-4000:])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_fortran(verbose)
    return LIB


FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
FDIR = os.path.join(HERE, "fortran")
FLIB = os.path.join(HERE, "libdbsr3_amd_fortran.so")
TDIR = os.path.join(ROOT, "tests", "fortran")
# drop-in layers with the reference's names for one host program each
# (DBSR_CI3 and DBSR_MAT3 define Add_integral / Add_matrix / Gen_matrix
# differently): static archives linked into the host program
ARCHIVES = {"ci3": ["dbs_gpu_ci3.f90"], "mat3": ["dbs_gpu_mat3.f90", "dbs_gpu_mat3_post.f90"]}
# test host programs: source, archive it links (or None)
DRIVERS = {"drv_dropin": None, "drv_ci3": "ci3", "drv_mat3": "mat3", "drv_chan": "mat3"}


def build_fortran(verbose=False):
    """the ISO_C_BINDING drop-in layers (libdbsr3_amd_fortran.so with the DBS
    routines, hf_matrix and the engine; libdbsr3_amd_ci3.a and
    libdbsr3_amd_mat3.a with the accumulation routines of each host program)
    and the Fortran host programs of the drop-in tests; skipped without flang"""
    if not os.path.exists(FLANG):
        if verbose:
            print("flang not found: Fortran drop-in layer not built")
        return None
    mod = os.path.join(OBJDIR, "fmod")
    os.makedirs(mod, exist_ok=True)
    obj = os.path.join(OBJDIR, "dbs_gpu.o")
    steps = [
        [FLANG, "-O2", "-fPIC", "-module-dir", mod, "-c", os.path.join(FDIR, "dbs_gpu.f90"), "-o", obj],
        [FLANG, "-shared", "-fPIC", obj, "-L" + HERE, "-ldbsr3_amd", "-Wl,-rpath,$ORIGIN", "-o", FLIB],
    ]
    for name, srcs in ARCHIVES.items():
        a = os.path.join(HERE, f"libdbsr3_amd_{name}.a")
        objs = []
        for src in srcs:
            o = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
            steps.append([FLANG, "-O2", "-fPIC", "-module-dir", mod, "-I" + mod, "-c", os.path.join(FDIR, src),
                          "-o", o])
            objs.append(o)
        steps.append(["rm", "-f", a])
        steps.append(["ar", "rcs", a, *objs])
    for drv, arch in DRIVERS.items():
        libs = ([os.path.join(HERE, f"libdbsr3_amd_{arch}.a")] if arch else [])
        steps.append([FLANG, "-O2", os.path.join(TDIR, drv + ".f90"), "-I" + mod, *libs, "-L" + HERE,
                      "-ldbsr3_amd_fortran", "-ldbsr3_amd", "-Wl,-rpath,$ORIGIN/../../dbsr3_amd",
                      "-o", os.path.join(TDIR, drv)])
    for cmd in steps:
        r = subprocess.run(cmd, capture_output=True, text=True, cwd=OBJDIR)
        if r.returncode != 0:
            raise RuntimeError("flang failed: " + " ".join(cmd) + "\n" + r.stderr[-3000:])
    if verbose:
        print("built", FLIB, "the ci3/mat3 archives and the test drivers")
    return FLIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
