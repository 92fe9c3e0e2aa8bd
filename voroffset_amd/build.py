"""Builds the native library libvoroffset_hip.so in-tree for gfx950.

    python -m voroffset_amd.build            # library only
    python -m voroffset_amd.build --cli      # library + offset3d_hip CLI

Every HIP translation unit is compiled with packed-FP32 instructions disabled
(``-Xclang -target-feature -Xclang -packed-fp32-ops``).
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libvoroffset_hip.so")
ARCH = os.environ.get("VO_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


def hipcc():
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found")
    return path


def run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_library(verbose=True):
    tmp = LIB + ".tmp"
    run([hipcc(), *HIP_FLAGS, "-shared", "-o", tmp, os.path.join(CSRC, "vo_api.hip")])
    os.replace(tmp, LIB)
    return LIB


def build_cli():
    out = os.path.join(HERE, "offset3d_hip")
    run([hipcc(), "-O2", "-std=c++17", "-o", out, os.path.join(CSRC, "offset3d.cpp"),
         f"-I{os.path.join(ROOT, 'include')}", f"-L{HERE}", "-lvoroffset_hip",
         f"-Wl,-rpath,{HERE}"])
    return out


if __name__ == "__main__":
    build_library()
    if "--cli" in sys.argv:
        build_cli()
