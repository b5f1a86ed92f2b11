"""Builds the engine in-tree: swifttn93_amd/lib/libtn93g.so (HIP, gfx950) and swifttn93_amd/bin/tn93g (CLI).

    python3 -m swifttn93_amd.build          (or swifttn93_amd.build.build() from Python)

Every HIP translation unit is compiled with -ffp-contract=off (tn93's floating-point expressions are evaluated as
written, no FMA contraction) and with packed-FP32 instructions disabled.
"""
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "lib" / "libtn93g.so"
CLI = PKG / "bin" / "tn93g"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("TN93G_ARCH", "gfx950")

HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument"]


def _run(cmd):
    print("+", " ".join(str(c) for c in cmd), flush=True)
    subprocess.run([str(c) for c in cmd], check=True)


def build(verbose=True):
    LIB.parent.mkdir(exist_ok=True)
    CLI.parent.mkdir(exist_ok=True)
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, CSRC / "tn93g.hip", "-lrocblas"])
    os.replace(tmp, LIB)
    ctmp = CLI.with_name("tn93g.tmp")
    _run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-pthread", "-o", ctmp, CSRC / "tn93g_cli.cpp",
          f"-L{LIB.parent}", "-ltn93g", "-Wl,-rpath,$ORIGIN/../lib", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(ctmp, CLI)
    return LIB, CLI


if __name__ == "__main__":
    build()
    sys.exit(0)
