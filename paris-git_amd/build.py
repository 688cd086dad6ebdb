"""Build recipe for libpps.so (HIP, gfx950) and the Fortran binding module.

Every HIP source is compiled for gfx950 only.  Packed FP32 instructions are
disabled (``-packed-fp32-ops``) for all device code.  Objects go to
``_build/`` next to this file and the shared library to ``libpps.so``.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpps.so")
SOURCES = ["pps_api.cpp", "pps_assemble.hip", "pps_ref.hip", "pps_fast.hip", "pps_adjoint.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("command failed: " + " ".join(cmd))
    return res


def build(verbose=False, jobs=None):
    """Compile every source and link libpps.so; returns the library path."""
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    cmds = []
    for src in SOURCES:
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        cmds.append([HIPCC] + FLAGS + lang + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    jobs = jobs or min(len(cmds), os.cpu_count() or 1, 8)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(lambda c: _run(c, verbose), cmds))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs, verbose)
    os.replace(tmp, LIB)
    return LIB


FORTRAN_SRC = os.path.join(HERE, "fortran", "pps_paris.f90")
FORTRAN_MOD_DIR = os.path.join(BUILD, "fortran")


DRIVER_SRC = os.path.join(ROOT, "tests", "fortran", "newsolver_driver.f90")
DRIVER_BIN = os.path.join(FORTRAN_MOD_DIR, "newsolver_driver")


def build_fortran(verbose=False):
    """Compile the iso_c_binding module with amdflang (if present) and link
    the Fortran test caller against libpps.so."""
    flang = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    if not os.path.exists(flang):
        return None
    os.makedirs(FORTRAN_MOD_DIR, exist_ok=True)
    obj = os.path.join(FORTRAN_MOD_DIR, "pps_paris.o")
    _run([flang, "-O2", "-fPIC", "-c", FORTRAN_SRC, "-J", FORTRAN_MOD_DIR, "-o", obj], verbose)
    if os.path.exists(DRIVER_SRC) and os.path.exists(LIB):
        _run([flang, "-O2", "-I", FORTRAN_MOD_DIR, DRIVER_SRC, obj, "-L" + HERE, "-lpps",
              "-Wl,-rpath," + HERE, "-Wl,-rpath,$ORIGIN/../..", "-o", DRIVER_BIN], verbose)
    return obj


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
