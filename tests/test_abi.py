"""The C-ABI library loads and exports every entry point include/mlod.h declares; argument
checks that return before any HIP call behave (no compute runs here)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "mlod.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\*?(mlod_\w+)\s*\(", txt, flags=re.M)))


@pytest.fixture(scope="module")
def L():
    from mlodsurfaces_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        from mlodsurfaces_amd import build
        build.build(verbose=False)
    return _lib.lib()


def test_header_declares_entry_points():
    f = declared_functions()
    for name in ["mlod_tree_create", "mlod_project", "mlod_set_kernel", "mlod_queries_upload", "mlod_tree_destroy"]:
        assert name in f
    assert len(f) >= 15


def test_library_exports_every_declared_symbol(L):
    from mlodsurfaces_amd import _lib
    so = _lib.LIB_PATH
    out = subprocess.run(["nm", "-D", "--defined-only", so], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (mlod_\w+)", out))
    for name in declared_functions():
        assert name in exported, name
        assert hasattr(L, name)
        assert name in _lib.SIGNATURES, f"{name} missing from the Python binding"


def test_argument_errors_without_gpu(L):
    from mlodsurfaces_amd import _lib
    h = C.c_void_p()
    p = np.zeros((4, 3), np.float32)
    fp = p.ctypes.data_as(C.POINTER(C.c_float))
    assert L.mlod_tree_create(fp, fp, 0, None, None, C.byref(h)) == -1
    assert b"empty point set" in L.mlod_last_error()
    assert L.mlod_tree_create(fp, fp, 4, None, None, None) == -1
    assert L.mlod_set_kernel(None, None) == -1
    assert L.mlod_set_params(None, 2.0, 0, 0, 0) == -1
    assert L.mlod_project(None, 1, 1, -1.0, None, None, None) == -1
    assert L.mlod_queries_upload(None, fp, 4, None) == -1
    assert L.mlod_tree_destroy(None) == 0
    info = _lib.TreeInfoC()
    assert L.mlod_tree_get_info(None, C.byref(info)) == -1
    # dual contouring and mesh passes: NULL handles and arrays are rejected before any HIP call
    n = C.c_uint32()
    assert L.mlod_dc_build(None, None, 0, 0, 0, None) == -1
    assert L.mlod_dc_build_from_points(None, None, 0, None) == -1
    assert L.mlod_dc_extract(None, C.byref(n), C.byref(n), None) == -1
    assert L.mlod_dc_refine(None, None, 3, 2, C.byref(n), C.byref(n), None) == -1
    assert L.mlod_dc_refined_mesh(None, None, None, None, None) == -1
    assert L.mlod_mesh_smooth(None, 4, None, 0, None, None) == -1
    tri = np.array([[0, 1, 9]], np.uint32)
    assert L.mlod_mesh_smooth(fp, 4, tri.ctypes.data_as(C.c_void_p), 1, fp, None) == -1
    assert b"not a vertex" in L.mlod_last_error()
    assert L.mlod_mesh_subdivide(fp, 4, tri.ctypes.data_as(C.c_void_p), 1, fp, 4, tri.ctypes.data_as(C.c_void_p),
                                 C.byref(n), None) == -1
    assert b"fewer than" in L.mlod_last_error()


def test_kernel_sigma_generation_matches_reference_semantics():
    """Kernel::updateSigmas (Kernel.cpp:110-128): float geometric series, then x radius."""
    from mlodsurfaces_amd import Kernel
    k = Kernel(type=2, beginningSigma=0.01, factor=2.0, nbOfComputedSigmas=5)
    k.updateSigmas(True, 0.75)
    f32 = np.float32
    s = [f32(0.01)]
    for _ in range(4):
        s.append(f32(s[-1] * f32(2.0)))
    assert k.sigmas == [float(f32(v) * f32(0.75)) for v in s]
    k2 = Kernel(type=2, fixedSigmas=[0.1, 0.3])
    k2.updateSigmas(False, 2.0)
    assert k2.sigmas == [float(f32(0.1) * f32(2.0)), float(f32(0.3) * f32(2.0))]
    assert Kernel().type == Kernel.GAUSSIAN and Kernel().sigma == 2.0
