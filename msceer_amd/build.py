"""Builds libmsceer_hip.so (gfx950) and the command-line drivers, in-tree.

    python -m msceer_amd.build

hipcc cross-compiles for gfx950 without a GPU.  Packed FP32 instructions are
disabled for every HIP translation unit.
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
CLI = os.path.join(PKG, "cli")
LIB = os.path.join(PKG, "libmsceer_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]


def hip_sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_lib():
    objs = []
    obj_dir = os.path.join(PKG, "_build")
    os.makedirs(obj_dir, exist_ok=True)
    for src in hip_sources():
        obj = os.path.join(obj_dir, os.path.basename(src) + ".o")
        _run([HIPCC, *HIP_FLAGS, "-I", os.path.join(ROOT, "include"), "-c", src, "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    return LIB


def build_cli():
    out = []
    if not os.path.isdir(CLI):
        return out
    for f in sorted(os.listdir(CLI)):
        if not f.endswith(".cpp"):
            continue
        exe = os.path.join(CLI, os.path.splitext(f)[0])
        _run(["g++", "-std=c++17", "-O2", "-fopenmp", "-I", os.path.join(ROOT, "include"),
              os.path.join(CLI, f), "-o", exe, "-L", PKG, "-lmsceer_hip",
              "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath," + PKG])
        out.append(exe)
    return out


def build_all():
    build_lib()
    build_cli()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
