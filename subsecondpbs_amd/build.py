"""Builds the HIP engine library (libpbs_hip.so) in-tree for gfx950.

Host planning code is compiled with g++ and -ffp-contract=off (its float arithmetic decides
grid sizes and step counts); the kernels and the C ABI are compiled with hipcc.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "_lib")
LIB = os.path.join(LIBDIR, "libpbs_hip.so")
ARCH = os.environ.get("PBS_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions stay off: the compiler would otherwise form v_pk_* from float code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build failed: " + " ".join(cmd[:3]) + " ...")
    return r


def build(verbose: bool = False) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    obj_dir = os.path.join(LIBDIR, "obj")
    os.makedirs(obj_dir, exist_ok=True)
    hipcc = _hipcc()
    plan_o = os.path.join(obj_dir, "plan.o")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-c",
          os.path.join(CSRC, "plan.cpp"), "-o", plan_o])
    objs = [plan_o]
    for src in ("kernels.hip", "adjoint.hip", "engine.hip", "compat.hip"):
        o = os.path.join(obj_dir, src.replace(".hip", ".o"))
        _run([hipcc, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", o])
        objs.append(o)
    tmp = LIB + ".tmp"
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    caller = build_cpp_caller()
    if verbose:
        print("built", caller)
    return LIB


def build_cpp_caller() -> str:
    """The C++ caller of subsecondWrapper that the GPU tests run (tests/cpp/rtd_wrapper_main.cpp):
    compiled with g++ against include/raytracedicom_integration.h and linked to the library built
    above, found at run time through an rpath relative to the executable."""
    root = os.path.dirname(PKG)
    src = os.path.join(root, "tests", "cpp", "rtd_wrapper_main.cpp")
    out_dir = os.path.join(root, "tests", "cpp", "_bin")
    os.makedirs(out_dir, exist_ok=True)
    exe = os.path.join(out_dir, "rtd_wrapper_main")
    rocm_inc = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "include")
    _run(["g++", "-std=c++17", "-O2", "-Wall", "-D__HIP_PLATFORM_AMD__", "-I", rocm_inc,
          "-I", os.path.join(root, "include"), src, "-L", LIBDIR, "-lpbs_hip",
          "-Wl,-rpath,$ORIGIN/../../../subsecondpbs_amd/_lib", "-o", exe + ".tmp"])
    os.replace(exe + ".tmp", exe)
    return exe


if __name__ == "__main__":
    build(verbose=True)
