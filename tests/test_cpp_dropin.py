"""The header-only C++ drop-in (include/mlod_apss.hpp) compiles against libmlod.so, and on the
GPU gives the same projections as the Python mirror of the interface."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

SRC = os.path.join(ROOT, "tests", "cpp", "apss_dropin.cpp")


def build_exe(tmp_path):
    from mlodsurfaces_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        from mlodsurfaces_amd import build
        build.build(verbose=False)
    exe = str(tmp_path / "apss_dropin")
    libdir = os.path.dirname(_lib.LIB_PATH)
    cmd = ["g++", "-std=c++17", "-O2", "-I", os.path.join(ROOT, "include"), SRC, "-o", exe, f"-L{libdir}", "-lmlod",
           f"-Wl,-rpath,{libdir}"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


def test_dropin_header_compiles_and_links(tmp_path):
    exe = build_exe(tmp_path)
    assert os.path.exists(exe)


@pytest.mark.gpu
def test_dropin_matches_python_api(tmp_path):
    from mlodsurfaces_amd import apss as A
    from mlodsurfaces_amd import datasets as ds
    exe = build_exe(tmp_path)
    p, n = ds.sphere(30000, seed=41)
    q = ds.jitter_queries(p[:3000], 0.01, seed=42)
    for name, a in (("p", p), ("n", n), ("q", q)):
        a.astype(np.float32).tofile(tmp_path / f"{name}.bin")
    r = subprocess.run([exe, *(str(tmp_path / f) for f in ("p.bin", "n.bin", "q.bin", "op.bin", "on.bin", "cells.bin"))],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    op = np.fromfile(tmp_path / "op.bin", np.float32).reshape(-1, 3)
    on = np.fromfile(tmp_path / "on.bin", np.float32).reshape(-1, 3)
    t = A.APSSOctree().build(p, n)
    K = A.Kernel(type=2, beginningSigma=0.01, factor=2.0, nbOfComputedSigmas=4).updateSigmas(True, t.radius())
    ap = A.Fastapss(t, K)
    ap.setScalingProtectionSphere(3.0)
    ap.copyPointsToGPU(len(q), q)
    for _ in range(3):
        pp, pn = ap.project(len(q), np.zeros_like(q), np.zeros_like(q), 1, K)
    assert np.array_equal(op, pp) and np.array_equal(on, pn)
    cells = np.fromfile(tmp_path / "cells.bin", np.int32).reshape(-1, 4)
    gen = A.OctreeGenerator(ap, K)
    total = []
    assert np.array_equal(cells, gen.generate(4, total))
    assert f"cells {len(cells)} total {len(total[0])}" in r.stdout
