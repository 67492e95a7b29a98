"""Ahead-of-time build of libmlbvh.so (all HIP sources, gfx950 only).

Run as ``python -m minimumlbvh_amd._build`` or through
``__graft_entry__.build()``.  Objects go to ``minimumlbvh_amd/_build/`` and
the shared library to ``minimumlbvh_amd/_lib/libmlbvh.so`` (both git-ignored;
the library travels with the source tree).
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
OBJ = os.path.join(PKG, "_build")
LIBDIR = os.path.join(PKG, "_lib")
LIB = os.path.join(LIBDIR, "libmlbvh.so")
ARCH = "gfx950"

# -ffp-contract=off keeps every a*b+c as two rounded operations (bit parity
# with the reference's host arithmetic); HIP's fp32 divide and sqrt stay
# correctly rounded by default.  Packed fp32 instructions are disabled.
HIPFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-fvisibility=hidden",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unknown-pragmas",
    f"-I{os.path.join(ROOT, 'include')}",
]


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found; the ROCm toolchain is required to build libmlbvh")
    return path


def sources() -> list[str]:
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [hipcc(), *HIPFLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_calib(verbose)
    build_facade_demo(verbose)
    build_facade_cpu(verbose)
    return LIB


CALIB_SRC = os.path.join(CSRC, "calib", "probes.hip")
CALIB_LIB = os.path.join(LIBDIR, "libmlbvh_calib.so")


def build_calib(verbose: bool = False) -> str:
    """The roofline calibration probes (not part of libmlbvh's ABI)."""
    tmp = CALIB_LIB + ".tmp"
    cmd = [hipcc(), *HIPFLAGS, "-shared", CALIB_SRC, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"calibration probes build failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, CALIB_LIB)
    if verbose:
        print(f"built {CALIB_LIB}")
    return CALIB_LIB


FACADE_CPU_SRC = os.path.join(ROOT, "tests", "facade", "facade_cpu.cpp")
FACADE_CPU_BIN = os.path.join(ROOT, "tests", "facade", "_build", "facade_cpu")


def build_facade_cpu(verbose: bool = False) -> str:
    """The host-side facade test program (BVHCPUBuilder and the inline
    traversals on the CPU); needs no GPU to run."""
    os.makedirs(os.path.dirname(FACADE_CPU_BIN), exist_ok=True)
    cmd = [hipcc(), *[f for f in HIPFLAGS if f != "-fvisibility=hidden"], "-fopenmp", FACADE_CPU_SRC, "-o",
           FACADE_CPU_BIN, f"-L{LIBDIR}", "-lmlbvh", "-Wl,-rpath,$ORIGIN/../../../minimumlbvh_amd/_lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"facade CPU program build failed:\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {FACADE_CPU_BIN}")
    return FACADE_CPU_BIN


FACADE_SRC = os.path.join(ROOT, "tests", "facade", "facade_demo.hip")
FACADE_BIN = os.path.join(ROOT, "tests", "facade", "_build", "facade_demo")


def build_facade_demo(verbose: bool = False) -> str:
    """The C++-facade test application (reference-style host code over
    include/minimum_lbvh.h), linked against libmlbvh.so."""
    os.makedirs(os.path.dirname(FACADE_BIN), exist_ok=True)
    cmd = [hipcc(), *[f for f in HIPFLAGS if f != "-fvisibility=hidden"], FACADE_SRC, "-o", FACADE_BIN,
           f"-L{LIBDIR}", "-lmlbvh", "-Wl,-rpath,$ORIGIN/../../../minimumlbvh_amd/_lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"facade demo build failed:\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {FACADE_BIN}")
    return FACADE_BIN


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
