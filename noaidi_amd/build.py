"""Compile the native libraries in-tree.

  noaidi_amd/_build/libprima_hip.so        HIP kernels (gfx950) + C-ABI host side
  noaidi_amd/_build/libprima_instances.so  seeded instance generators (C++)

Run ``python -m noaidi_amd.build``. Uses hipcc/g++ directly; no cmake.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
ROOT = os.path.dirname(HERE)

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PRIMA_OFFLOAD_ARCH", "gfx950")

# Packed FP32 instructions are kept out of every device build.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_hip() -> str:
    os.makedirs(OUT, exist_ok=True)
    objs = []
    inc = ["-I", os.path.join(ROOT, "include"), "-I", CSRC]
    for src in ("pdhg_kernel.hip", "kernels_ops.hip", "refine_kernels.hip"):
        obj = os.path.join(OUT, src.replace(".hip", ".o"))
        _run([HIPCC, f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-c",
              *DEVICE_FLAGS, *inc, os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    obj = os.path.join(OUT, "host_api.o")
    _run([HIPCC, "-x", "c++", "-O2", "-std=c++17", "-fPIC", "-c", "-D__HIP_PLATFORM_AMD__",
          "-I", "/opt/rocm/include",
          *inc, os.path.join(CSRC, "host_api.cpp"), "-o", obj])
    objs.append(obj)
    lib = os.path.join(OUT, "libprima_hip.so")
    tmp = lib + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, lib)
    return lib


def build_instances() -> str:
    os.makedirs(OUT, exist_ok=True)
    lib = os.path.join(OUT, "libprima_instances.so")
    tmp = lib + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
          os.path.join(CSRC, "instances.cpp"), "-o", tmp])
    os.replace(tmp, lib)
    return lib


def build_all() -> None:
    build_instances()
    build_hip()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    if "all" in what:
        build_all()
    else:
        if "instances" in what:
            build_instances()
        if "hip" in what:
            build_hip()
