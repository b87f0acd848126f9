"""Build libecsg_hip.so (gfx950) in-tree with hipcc, and the command-line
driver ecsg_run (host C++, linked against it) beside it.

    python -m ecsg_amd.build

Flags: -ffp-contract=off keeps the double-precision arithmetic operation for
operation equal to the reference's; packed FP32 instructions are disabled
(-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "lib", "libecsg_hip.so")
SOURCES = [os.path.join(HERE, "csrc", "ecsg_capi.hip")]
DRIVER = os.path.join(HERE, "lib", "ecsg_run")
DRIVER_SOURCES = [os.path.join(HERE, "driver", "ecsg_run.cpp")]
ARCH = os.environ.get("ECSG_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build libecsg_hip.so)")


def flags():
    return ["--offload-arch=%s" % ARCH, "-O3", "-std=c++17", "-fPIC", "-shared",
            "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-result", "-I", os.path.join(ROOT, "include"), "-I", os.path.join(HERE, "csrc")]


def build(verbose: bool = False) -> str:
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    tmp = LIB + ".tmp%d" % os.getpid()
    cmd = [hipcc()] + flags() + SOURCES + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    out = subprocess.run(cmd, capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + out.stdout + out.stderr)
    os.replace(tmp, LIB)
    build_driver(verbose)
    return LIB


def build_driver(verbose: bool = False) -> str:
    """g++ build of ecsg_run, found at run time next to libecsg_hip.so."""
    cxx = os.environ.get("CXX") or shutil.which("g++") or "g++"
    tmp = DRIVER + ".tmp%d" % os.getpid()
    libdir = os.path.dirname(LIB)
    cmd = [cxx, "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), "-I", os.path.join(HERE, "driver")]
    cmd += DRIVER_SOURCES + ["-L", libdir, "-lecsg_hip", "-Wl,-rpath,$ORIGIN",
                             "-Wl,-rpath-link,/opt/rocm/lib", "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    out = subprocess.run(cmd, capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("driver build failed:\n" + out.stdout + out.stderr)
    os.replace(tmp, DRIVER)
    return DRIVER


if __name__ == "__main__":
    print(build(verbose=True))
    sys.exit(0)
