"""Builds the native library libvoroffset_hip.so in-tree for gfx950.

    python -m voroffset_amd.build            # library, vo_copy_probe, offset3d_hip CLI

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


def build_copy_probe():
    """The calibration copy kernel used for bandwidth and counter scaling."""
    out = os.path.join(HERE, "vo_copy_probe")
    run([hipcc(), *HIP_FLAGS, "-o", out + ".tmp", os.path.join(CSRC, "copy_probe.hip")])
    os.replace(out + ".tmp", out)
    return out


def build_cli():
    """offset3d_hip: host-only C++ (g++), linked against the library with an
    $ORIGIN rpath so the pair can move together."""
    out = os.path.join(HERE, "offset3d_hip")
    gxx = shutil.which("g++") or "g++"
    run([gxx, "-O2", "-std=c++17", "-Wall", "-o", out + ".tmp", os.path.join(CSRC, "offset3d.cpp"),
         f"-I{os.path.join(ROOT, 'include')}", f"-L{HERE}", "-lvoroffset_hip", "-Wl,-rpath,$ORIGIN"])
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    build_library()
    build_copy_probe()
    build_cli()
