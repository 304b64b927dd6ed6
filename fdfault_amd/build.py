"""Builds the native parts of fdfault_amd in-tree:

* ``fdfault_amd/_lib/libfdfault_hip.so`` -- the C-ABI library (HIP kernels for
  gfx950 + the C++ host problem compiler);
* ``fdfault_amd/_lib/fdfault`` -- the drop-in executable.

HIP code is compiled with -ffp-contract=off (the FP64 path keeps the
reference's rounding) and with packed-FP32 instructions disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_lib")
LIB = os.path.join(OUT, "libfdfault_hip.so")
EXE = os.path.join(OUT, "fdfault")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("FDF_OFFLOAD_ARCH", "gfx950")
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result", "-Wno-unused-command-line-argument"]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-fopenmp"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=ROOT)


def build(verbose: bool = True) -> str:
    os.makedirs(OUT, exist_ok=True)
    setup_o = os.path.join(OUT, "setup.o")
    _run(["g++"] + HOST_FLAGS + ["-c", os.path.join(SRC, "setup.cpp"), "-o", setup_o])
    runtime_o = os.path.join(OUT, "runtime.o")
    _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(SRC, "runtime.hip"), "-o", runtime_o])
    tmp = LIB + ".tmp"
    _run([HIPCC] + ["--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp, runtime_o, setup_o, "-lgomp"])
    os.replace(tmp, LIB)
    _run(["g++", "-O2", "-std=c++17", os.path.join(SRC, "main.cpp"), "-o", EXE, "-L" + OUT, "-lfdfault_hip",
          "-Wl,-rpath,$ORIGIN"])
    _run([HIPCC] + HIP_FLAGS + [os.path.join(SRC, "copy_bw.hip"), "-o", os.path.join(OUT, "copy_bw")])
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
