"""Builds libx3dpt.so in-tree: host C++ with g++, HIP kernels with hipcc for gfx950.

    python -m x3dcore2_amd.build        # or __graft_entry__.build()

Flags that matter for results: -ffp-contract=off on both compilers (no FMA contraction, so the
FP32 operation order of the reference is kept), HIP's default IEEE division and square root,
and packed-FP32 instructions disabled on the device.
"""

from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libx3dpt.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = os.environ.get("X3D_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]
GXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall",
             f"-I{ROCM}/include", "-D__HIP_PLATFORM_AMD__"]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def build_library(verbose: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    hipcc = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")
    gxx = shutil.which("g++") or "g++"
    objs, cmds = [], []
    for unit in ("scene", "objfile"):
        o = os.path.join(BUILD, unit + ".o")
        cmds.append([gxx, *GXX_FLAGS, "-c", os.path.join(CSRC, unit + ".cpp"), "-o", o])
        objs.append(o)
    for unit in ("render", "wavefront", "batch", "adjoint", "compose"):
        o = os.path.join(BUILD, unit + ".o")
        cmds.append([hipcc, *HIPCC_FLAGS, "-c", os.path.join(CSRC, unit + ".hip"), "-o", o])
        objs.append(o)
    # independent translation units compile concurrently
    with concurrent.futures.ThreadPoolExecutor(max_workers=len(cmds)) as pool:
        for f in [pool.submit(_run, c) for c in cmds]:
            f.result()
    tmp = LIB + ".tmp"
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp,
          f"-L{ROCM}/lib", "-lamdhip64", "-Wl,-rpath," + os.path.join(ROCM, "lib")])
    os.replace(tmp, LIB)
    build_host()
    return LIB


HOST = os.path.join(PKG, "host", "render_cornell")


def build_host() -> str:
    """Builds the C++ host program (host/render_cornell.cpp) against the C ABI and libx3dpt.so."""
    gxx = shutil.which("g++") or "g++"
    tmp = HOST + ".tmp"
    _run([gxx, "-std=c++17", "-O2", "-Wall", "-I" + os.path.join(ROOT, "include"), HOST + ".cpp", "-o", tmp,
          "-L" + PKG, "-lx3dpt", "-Wl,-rpath,$ORIGIN/.."])
    os.replace(tmp, HOST)
    return HOST


def build_oracle(with_reference: bool = True) -> None:
    """Builds the CPU checkers under oracle/ (test infrastructure)."""
    odir = os.path.join(ROOT, "oracle")
    _run(["make", "-s", "-C", odir, "-j8"])
    _run(["make", "-s", "-C", odir, "-j8", "sanitize"])
    if with_reference and os.path.isdir("/root/reference/corelib/src"):
        _run(["make", "-s", "-C", odir, "-j8", "ref"])


if __name__ == "__main__":
    build_library()
    if "--no-oracle" not in sys.argv:
        build_oracle()
