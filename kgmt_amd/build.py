"""Build recipe of libkgmt.so (HIP kernels for gfx950 + the C-ABI) and of the
CPU oracle library used by the tests.

    python -m kgmt_amd.build            # both
    python -m kgmt_amd.build --lib      # product library only

Flags that matter for parity: -ffp-contract=off (one IEEE operation per float
operator, as in the oracle's gcc build) and HIP's default correctly rounded
float division and square root. Packed FP32 instructions are disabled
(-packed-fp32-ops) for every HIP translation unit.
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libkgmt.so")
ORACLE_DIR = os.path.join(ROOT, "oracle")
ARCH = os.environ.get("KGMT_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["kgmt_kernels.hip", "kgmt_adjoint.hip", "kgmt_noalloc.hip", "kgmt_engine.cpp"]
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build libkgmt")


def build_lib(verbose: bool = False) -> str:
    hipcc = _hipcc()
    objs = []
    build_dir = os.path.join(HERE, "_build")
    os.makedirs(build_dir, exist_ok=True)
    for src in HIP_SOURCES:
        obj = os.path.join(build_dir, os.path.splitext(src)[0] + ".o")
        cmd = [hipcc, "-x", "hip", *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


def build_dropin(verbose: bool = False) -> str:
    """tests/cpp/dropin_main: a C++ caller of include/kgmt_planner.hpp linked
    against libkgmt.so (host code only)."""
    src = os.path.join(ROOT, "tests", "cpp", "dropin_main.cpp")
    out = os.path.join(ROOT, "tests", "cpp", "dropin_main")
    cmd = [_hipcc(), "-O2", "-std=c++17", src, "-o", out, f"-L{HERE}", "-lkgmt", f"-Wl,-rpath,{HERE}",
           "-Wl,-rpath,$ORIGIN/../../kgmt_amd"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return out


def build_oracle(verbose: bool = False) -> str:
    cmd = ["make", "-s", "-C", ORACLE_DIR]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return os.path.join(ORACLE_DIR, "libkgmt_oracle.so")


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib", action="store_true", help="build only libkgmt.so")
    ap.add_argument("--oracle", action="store_true", help="build only the oracle")
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args(argv)
    both = not (a.lib or a.oracle)
    if a.lib or both:
        print(build_lib(a.verbose))
        print(build_dropin(a.verbose))
    if a.oracle or both:
        print(build_oracle(a.verbose))
    return 0


if __name__ == "__main__":
    sys.exit(main())
