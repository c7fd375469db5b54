"""Builds the in-tree native artefacts of the package.

  liblulesh_mi355x.so   HIP kernels + C ABI (include/lulesh_mi355x.h), gfx950
  lulesh_mi355x         C++ command-line harness linked against the library

Both land next to this file so they travel with the source tree.  Device
code is compiled without FMA contraction (the kernels reproduce the
reference's rounding) and without packed-FP32 instructions; the element
phases are compiled a second time with contraction (LULESH_FUSED, the
fused-multiply-add mode of the float64 step).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(PKG, "liblulesh_mi355x.so")
HARNESS = os.path.join(PKG, "lulesh_mi355x")

HIP_SOURCES = ["lulesh_kernels.hip", "lulesh_force.hip", "lulesh_mixed.hip",
               "lulesh_abi.hip", "lulesh_batch.hip", "lulesh_mlp.hip",
               "lulesh_setup.cpp", "lulesh_output.cpp",
               "lulesh_mesh.cpp"]
HEADERS = ["lulesh_kernels.hpp", "lulesh_elem.hpp", "lulesh_cbrt.h",
           "lulesh_mesh.hpp", "lulesh_setup.hpp", "lulesh_mlp.hpp"]
# built a second time with multiply-adds fused (namespace lulesh::fused)
FUSED_SOURCES = ["lulesh_force.hip", "lulesh_kernels.hip"]
OBJ_DIR = os.path.join(PKG, "_build")

COMMON_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + INCLUDE, "-I" + CSRC,
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc",
                 shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd: list[str]) -> None:
    out = subprocess.run(cmd, capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" +
                           out.stdout + out.stderr)


def build_library() -> str:
    """Compiles every source to an object (in parallel), then links."""
    from concurrent.futures import ThreadPoolExecutor

    os.makedirs(OBJ_DIR, exist_ok=True)
    cc = hipcc()
    jobs = []
    for src in HIP_SOURCES:
        obj = os.path.join(OBJ_DIR, os.path.splitext(src)[0] + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        jobs.append((obj, [cc] + COMMON_FLAGS + ["-fPIC"] + lang +
                     ["-c", os.path.join(CSRC, src), "-o", obj]))
    fused_flags = [("-ffp-contract=fast" if f == "-ffp-contract=off" else f)
                   for f in COMMON_FLAGS] + ["-DLULESH_FUSED"]
    for src in FUSED_SOURCES:
        obj = os.path.join(OBJ_DIR, os.path.splitext(src)[0] + "_fused.o")
        jobs.append((obj, [cc] + fused_flags + ["-fPIC", "-x", "hip", "-c",
                                                os.path.join(CSRC, src), "-o", obj]))
    workers = max(1, min(len(jobs), os.cpu_count() or 1, 8))
    with ThreadPoolExecutor(max_workers=workers) as pool:
        list(pool.map(lambda j: _run(j[1]), jobs))
    tmp = LIB + ".tmp"
    _run([cc, "--offload-arch=gfx950", "-fPIC", "-shared", "-o", tmp] +
         [obj for obj, _ in jobs])
    os.replace(tmp, LIB)
    return LIB


def build_harness() -> str:
    src = os.path.join(PKG, "harness", "lulesh_main.cc")
    tmp = HARNESS + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I" + INCLUDE, "-o", tmp, src,
          "-L" + PKG, "-llulesh_mi355x", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, HARNESS)
    return HARNESS


def build_all() -> list[str]:
    return [build_library(), build_harness()]


if __name__ == "__main__":
    for p in build_all():
        print("built", p)
    sys.exit(0)
