"""Builds gpair_amd/libgpair_hip.so (HIP kernels + C ABI + C host loop) for gfx950.

    python -m gpair_amd.build

hipcc compiles the device code ahead of time (no runtime kernel compilation,
unlike the reference's clBuildProgram path, gpair1/trunk/gpu.c:270-304).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libgpair_hip.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GPAIR_ARCH", "gfx950")

# -packed-fp32-ops: no v_pk_*_f32 generation (wrong results seen with some forms)
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-Wno-unused-result"]
C_FLAGS = ["-O2", "-std=gnu11", "-fPIC", "-Wall"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src in ("gpair_api.hip", "gpair_nn.hip"):
        obj = os.path.join(BUILD, src + ".o")
        _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in ("gpair_cg.c", "gpair_host.c", "gpair_io.c"):
        obj = os.path.join(BUILD, src + ".o")
        _run(["gcc", *C_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    shutil.rmtree(BUILD, ignore_errors=True)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
