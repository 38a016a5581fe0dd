"""Builds the HIP shared library ``coral_amd/lib/libcoral_amd.so`` for gfx950.

Run ``python -m coral_amd.build`` (or ``__graft_entry__.build()``).  Sources
live in ``coral_amd/csrc``; objects go to ``coral_amd/_build``.  The library is
compiled in-tree so that it travels with the repository snapshot.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libcoral_amd.so")

ARCH = os.environ.get("CORAL_AMD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX_HOST", shutil.which("g++") or "g++")

HIP_FLAGS = [
    "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
    # no packed-FP32 instructions anywhere in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]
HOST_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-Wall"]

HIP_SOURCES = ["kernels.hip", "xpass.hip", "ypass.hip", "ztransform.hip", "capi.hip", "ztransform32.hip",
               "noalloc.hip", "nn.hip"]
HOST_SOURCES = ["assembly.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(ln for ln in r.stdout.splitlines() if "packed-fp32-ops" not in ln)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    return out


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    jobs = []
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        objs.append(obj)
        jobs.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    for src in HOST_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        objs.append(obj)
        jobs.append([CXX, *HOST_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for out in ex.map(_run, jobs):
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


REPO = os.path.dirname(HERE)
FORTRAN_DIR = os.path.join(REPO, "tests", "fortran")
FORTRAN_BUILD = os.path.join(FORTRAN_DIR, "_build")
FC = os.environ.get("FC_AMD", shutil.which("amdflang") or "/opt/rocm/bin/amdflang")
CC = os.environ.get("CC_HOST", shutil.which("gcc") or "gcc")


def build_fortran_host() -> str:
    """Compiles the Fortran interface module ``include/coral_amd_binding.f90``
    with amdflang and links the Fortran test programs under ``tests/fortran``
    against the library (run ``build`` first): ``binding_layout`` (layout of
    the bind(C) types), ``banded_kat`` (a Fortran host calling the banded
    solve), ``step_driver`` (a Fortran host running whole time steps), and the
    C layout probe ``c_layout``."""
    os.makedirs(FORTRAN_BUILD, exist_ok=True)
    mod_obj = os.path.join(FORTRAN_BUILD, "coral_amd_binding.o")
    _run([FC, "-O2", "-c", os.path.join(REPO, "include", "coral_amd_binding.f90"), "-J", FORTRAN_BUILD,
          "-o", mod_obj])
    link = ["-L", LIBDIR, "-lcoral_amd", "-Wl,-rpath,$ORIGIN/../../../coral_amd/lib"]
    for prog in ("binding_layout", "banded_kat", "step_driver"):
        _run([FC, "-O2", "-I", FORTRAN_BUILD, os.path.join(FORTRAN_DIR, prog + ".f90"), mod_obj, *link,
              "-o", os.path.join(FORTRAN_BUILD, prog)])
    _run([CC, "-std=c11", "-Wall", "-I", os.path.join(REPO, "include"), os.path.join(FORTRAN_DIR, "c_layout.c"),
          "-o", os.path.join(FORTRAN_BUILD, "c_layout")])
    return FORTRAN_BUILD


if __name__ == "__main__":
    path = build(verbose="-v" in sys.argv)
    build_fortran_host()
    print(path)
