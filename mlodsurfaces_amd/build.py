"""Builds mlodsurfaces_amd/libmlod.so in-tree with hipcc for gfx950.

    python -m mlodsurfaces_amd.build          # or: python mlodsurfaces_amd/build.py

Every HIP source is compiled to an object under mlodsurfaces_amd/_build/ (in parallel) and
linked into one shared library exporting the C ABI of include/mlod.h.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmlod.so")
ARCH = os.environ.get("MLOD_ARCH", "gfx950")

# Packed-FP32 instructions are kept out of all device code (-packed-fp32-ops).
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-Wno-unused-result"]
SOURCES = ["build.hip", "project.hip", "queries.hip", "neighbors.hip", "serve.hip", "adaptive.hip", "dc.hip", "noalloc.hip",
           "api.cpp"]
# The octree build reproduces the reference's float arithmetic (cell boxes, radii, areas,
# plain moments) bit for bit, and the neighbour search the kd-tree's float distances, so no
# multiply-add may be fused there.
NO_CONTRACT = {"build.hip", "neighbors.hip", "dc.hip"}


def hipcc():
    h = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(h):
        raise RuntimeError("hipcc not found")
    return h


def _compile(src: str) -> str:
    path = os.path.join(CSRC, src)
    obj = os.path.join(OUT_DIR, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc()] + HIP_FLAGS
    if src.endswith(".cpp"):
        cmd += ["-ffp-contract=off"]
    else:
        cmd += ["-x", "hip"]
    if src in NO_CONTRACT:
        cmd += ["-ffp-contract=off"]
    cmd += ["-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    return obj


def build(verbose: bool = True) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-pthread", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-6000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    try:
        build()
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
