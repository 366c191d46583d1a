"""Builds the in-tree native libraries of the package.

* ``libopensc_hip.so``  -- the HIP library (gfx950) behind include/opensc_hip.h
* ``libopensc_f.so``    -- the Fortran drop-in module open_spacecharge_mod
  (amdflang, ISO_C_BINDING over the C-ABI), with its ``.mod`` file in
  ``openspacecharge_amd/fortran/include``

``python -m openspacecharge_amd.build`` builds both.  Object files go to
``openspacecharge_amd/_build`` (git-ignored); every build recompiles.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
FSRC = os.path.join(PKG, "fortran")
BUILD = os.path.join(PKG, "_build")
HIP_LIB = os.path.join(PKG, "libopensc_hip.so")
F_LIB = os.path.join(PKG, "libopensc_f.so")
F_MODDIR = os.path.join(FSRC, "include")

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
FLANG = os.path.join(ROCM, "bin", "amdflang")
ARCH = os.environ.get("OSC_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["context.hip", "fft_f64.hip", "fft_f32.hip", "green.hip", "solve.hip",
               "solvers.hip", "deposit.hip", "gather.hip", "capi.hip", "server.hip",
               "noalloc.hip"]

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # FP64 index and Green arithmetic must not be contracted into FMAs
    "-ffp-contract=off",
    # no packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fvisibility=hidden", "-Wall", "-Wno-unused-function", "-Wno-unused-result",
    "-I" + os.path.join(ROOT, "include"),
]


def _run(cmd, quiet):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"command failed ({p.returncode}): {' '.join(cmd)}\n{p.stdout}")
    if not quiet and p.stdout.strip():
        lines = [l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l]
        if lines:
            print("\n".join(lines))
    return p.stdout


def build_hip(jobs: int | None = None, quiet: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 4)
    objs = []
    cmds = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        objs.append(obj)
        cmds.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(lambda c: _run(c, quiet), cmds))
    tmp = HIP_LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs], quiet)
    os.replace(tmp, HIP_LIB)
    return HIP_LIB


F_SOURCES = ["opensc_hip.f90", "open_spacecharge_mod.f90"]
F_DRIVER_SRC = os.path.join(ROOT, "tests", "fortran", "drive_shim.f90")
F_DRIVER = os.path.join(BUILD, "drive_shim")


def build_fortran(quiet: bool = True, driver: bool = True) -> str | None:
    """The Fortran drop-in (libopensc_f.so, .mod files in fortran/include) and,
    with driver, the test program tests/fortran/drive_shim.f90 linked to it."""
    if not os.path.exists(FLANG):
        return None
    os.makedirs(F_MODDIR, exist_ok=True)
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src in F_SOURCES:
        obj = os.path.join(BUILD, src.replace(".f90", "_f.o"))
        _run([FLANG, "-O2", "-ffp-contract=off", "-fPIC", "-cpp", "-ffree-form", "-module-dir",
              F_MODDIR, "-c",
              os.path.join(FSRC, src), "-o", obj], quiet)
        objs.append(obj)
    tmp = F_LIB + ".tmp"
    _run([FLANG, "-shared", "-fPIC", "-o", tmp, *objs, f"-L{PKG}", "-lopensc_hip",
          f"-Wl,-rpath,{PKG}", "-Wl,-rpath,$ORIGIN"], quiet)
    os.replace(tmp, F_LIB)
    if driver and os.path.exists(F_DRIVER_SRC):
        _run([FLANG, "-O2", "-cpp", "-ffree-form", "-I", F_MODDIR, "-module-dir", BUILD,
              F_DRIVER_SRC, "-o", F_DRIVER + ".tmp", f"-L{PKG}", "-lopensc_f", "-lopensc_hip",
              f"-Wl,-rpath,{PKG}"], quiet)
        os.replace(F_DRIVER + ".tmp", F_DRIVER)
    return F_LIB


def build_all(quiet: bool = True) -> None:
    build_hip(quiet=quiet)
    build_fortran(quiet=quiet)


def clean() -> None:
    shutil.rmtree(BUILD, ignore_errors=True)
    for f in (HIP_LIB, F_LIB):
        if os.path.exists(f):
            os.remove(f)


if __name__ == "__main__":
    build_all(quiet="-v" not in sys.argv)
    print(HIP_LIB)
