"""Builds libcism_gpu.so (HIP, gfx950) and the Fortran drop-in modules in-tree.

    python -m cism_amd.build            # library + Fortran shim
    python -m cism_amd.build --lib      # library only

Compiler flags:
  --offload-arch=gfx950                 CDNA4 code objects only
  -ffp-contract=off                     every multiply and add rounded separately, as
                                        the reference's expressions are written (bitwise
                                        comparable with the CPU restatement)
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                        no packed FP32 instructions (their results have
                                        been unreliable on this hardware pool)
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
FSRC = os.path.join(PKG, "fortran")
LIB = os.path.join(PKG, "libcism_gpu.so")
FLIB = os.path.join(PKG, "libcism_fortran.so")
OBJ = os.path.join(PKG, "_build")

SOURCES = ["api.hip", "solver.hip", "assemble.hip", "surrogate.hip", "upstream.hip", "ssa.hip",
           "noalloc.hip", "coarse.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("FC", "/opt/rocm/bin/amdflang")

HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument", "-Wno-unused-result"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def _stale(target, deps):
    if not os.path.exists(target):
        return True
    t = os.path.getmtime(target)
    return any(os.path.getmtime(d) > t for d in deps)


def build_lib(force=True, jobs=4):
    os.makedirs(OBJ, exist_ok=True)
    hdrs = [os.path.join(CSRC, "internal.h"), os.path.join(ROOT, "include", "cism_gpu.h")]
    objs, procs = [], []
    for s in SOURCES:
        src = os.path.join(CSRC, s)
        if not os.path.exists(src):
            continue
        o = os.path.join(OBJ, s.replace(".hip", ".o"))
        objs.append(o)
        if force or _stale(o, [src] + hdrs):
            cmd = [HIPCC] + HIP_FLAGS + ["-c", src, "-o", o]
            print(" ".join(cmd), flush=True)
            procs.append(subprocess.Popen(cmd))
            if len(procs) >= jobs:
                if procs.pop(0).wait() != 0:
                    raise RuntimeError("hipcc failed")
    for p in procs:
        if p.wait() != 0:
            raise RuntimeError("hipcc failed")
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs +
         ["-L/opt/rocm/lib", "-lrocprofiler-sdk-roctx", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, LIB)
    return LIB


def build_fortran():
    if not os.path.isdir(FSRC) or not shutil.which(FLANG) and not os.path.exists(FLANG):
        return None
    srcs = [os.path.join(FSRC, f) for f in sorted(os.listdir(FSRC)) if f.endswith(".F90")]
    if not srcs:
        return None
    mod = os.path.join(OBJ, "mod")
    os.makedirs(mod, exist_ok=True)
    objs = []
    for s in srcs:
        o = os.path.join(OBJ, os.path.basename(s).replace(".F90", ".o"))
        _run([FLANG, "-O2", "-fPIC", "-c", s, "-o", o, "-module-dir", mod])
        objs.append(o)
    tmp = FLIB + ".tmp"
    _run([FLANG, "-shared", "-fPIC", "-o", tmp] + objs +
         ["-L" + PKG, "-lcism_gpu", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, FLIB)
    return FLIB


def build_fortran_driver():
    """Test driver of the Fortran module (tests/fortran); built beside its source."""
    src = os.path.join(ROOT, "tests", "fortran", "fortran_driver.F90")
    if not os.path.exists(src) or not os.path.exists(FLIB):
        return None
    exe = os.path.join(ROOT, "tests", "fortran", "fortran_driver")
    _run([FLANG, "-O2", src, "-o", exe, "-I", os.path.join(OBJ, "mod"), "-L" + PKG,
          "-lcism_fortran", "-lcism_gpu", "-Wl,-rpath,$ORIGIN/../../cism_amd"])
    return exe


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib", action="store_true", help="HIP library only")
    ap.add_argument("--incremental", action="store_true")
    a = ap.parse_args(argv)
    build_lib(force=not a.incremental)
    if not a.lib:
        build_fortran()
        build_fortran_driver()


if __name__ == "__main__":
    sys.exit(main())
