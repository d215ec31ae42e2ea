"""Builds the native pieces of caustic_amd in-tree for gfx950.

  caustic_amd/libcaustic_ot.so   HIP kernels + C ABI of the OT solver (include/caustic_ot.h)
  caustic_amd/libcaustic_wsn.so  warm-start network kernels + C ABI (include/caustic_wsn.h)
  caustic_amd/bin/caustic_gpu    drop-in CLI for the reference's caustic_GPU executable

Run ``python -m caustic_amd.build`` (hipcc cross-compiles without a GPU).
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libcaustic_ot.so")
WSN_LIB = os.path.join(PKG, "libcaustic_wsn.so")
CLI = os.path.join(PKG, "bin", "caustic_gpu")
ARCH = os.environ.get("CAUSTIC_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # packed FP32 instructions stay off: hipcc only emits scalar FP32 forms
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
    "-Wall",
    "-Wno-unused-result",
]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def _shared(sources, out):
    objs = []
    for src in sources:
        obj = os.path.join(CSRC, os.path.splitext(src)[0] + ".o")
        extra = ["-x", "hip"] if src.endswith(".cpp") else []
        _run(["hipcc", *HIP_FLAGS, *extra, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    _run(["hipcc", *HIP_FLAGS, "-shared", *objs, "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    for o in objs:
        os.remove(o)


def build_library():
    _shared(("cot_kernels.hip", "cot_solver.hip", "cot_reflect.hip", "cot_newton.hip", "cot_render.hip",
             "cot_splat.hip", "cot_api.cpp", "cot_ladder.cpp", "cot_pool.cpp"), LIB)
    _shared(("wsn_kernels.hip", "wsn_api.cpp"), WSN_LIB)


def build_cli():
    src = os.path.join(CSRC, "caustic_gpu_main.cpp")
    if not os.path.exists(src):
        return
    os.makedirs(os.path.dirname(CLI), exist_ok=True)
    _run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), src, "-o", CLI + ".tmp",
          "-L", PKG, "-lcaustic_ot", f"-Wl,-rpath,$ORIGIN/..",
          "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(CLI + ".tmp", CLI)


def main():
    build_library()
    build_cli()


if __name__ == "__main__":
    sys.exit(main())
