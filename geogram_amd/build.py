"""Build the native libraries in-tree.

* libvgc.so      — gfx950 kernels + C ABI (hipcc, the product)
* libvgc_emu.so  — CPU emulation of the same kernels (g++, tests only)

Usage: python -m geogram_amd.build [--only gpu|emu]
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")

HIP_FLAGS = [
    "-std=c++17", "-O3", "--offload-arch=gfx950", "-fPIC", "-shared",
    # no implicit FMA contraction: the error bounds of the sign tests and the
    # bit-for-bit agreement with the emulation assume the written operations
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
EMU_FLAGS = ["-std=c++17", "-O2", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=CSRC)


def build_gpu():
    out = os.path.join(HERE, "libvgc.so")
    _run(["hipcc", *HIP_FLAGS, "-o", out + ".tmp", "backend_hip.hip"])
    os.replace(out + ".tmp", out)
    return out


def build_emu():
    out = os.path.join(HERE, "libvgc_emu.so")
    _run(["g++", *EMU_FLAGS, "-o", out + ".tmp", "backend_emu.cpp"])
    os.replace(out + ".tmp", out)
    return out


def build(only=None):
    outs = []
    if only in (None, "gpu"):
        outs.append(build_gpu())
    if only in (None, "emu"):
        outs.append(build_emu())
    return outs


if __name__ == "__main__":
    only = None
    if "--only" in sys.argv:
        only = sys.argv[sys.argv.index("--only") + 1]
    build(only)
