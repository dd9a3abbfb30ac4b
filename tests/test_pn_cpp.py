"""The header-only C++ point-set files and settings (include/mlod_pn.hpp) agree with the Python
ones (mlodsurfaces_amd.pointio) bit for bit: the normalised frame and points of a model, the
.params fields and generated sigmas, and the PLY they write."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, load_golden
from mlodsurfaces_amd import pointio

SRC = os.path.join(ROOT, "tests", "cpp", "pn_tool.cpp")


@pytest.fixture(scope="module")
def tool(tmp_path_factory):
    from mlodsurfaces_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        from mlodsurfaces_amd import build
        build.build(verbose=False)
    exe = str(tmp_path_factory.mktemp("pn") / "pn_tool")
    libdir = os.path.dirname(_lib.LIB_PATH)
    cmd = ["g++", "-std=c++17", "-O2", "-I", os.path.join(ROOT, "include"), SRC, "-o", exe, f"-L{libdir}", "-lmlod",
           f"-Wl,-rpath,{libdir}"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


@pytest.mark.parametrize("name,ext", [("real_Face.npz", "ply"), ("real_half-igea.npz", "pn")])
def test_cpp_matches_python(tool, tmp_path, name, ext):
    g = load_golden(name)
    model = str(tmp_path / f"m.{ext}")
    pointio.write_points(model, g["points"], g["normals"])
    params = tmp_path / "m.params"
    params.write_bytes(g["params"].tobytes())
    out = subprocess.run([tool, model, str(tmp_path / "c"), str(params)], check=True, capture_output=True,
                         text=True).stdout.splitlines()
    Pn, fr = pointio.normalize(g["points"])
    raw = np.fromfile(str(tmp_path / "c.norm"), np.float32)
    assert np.array_equal(raw[:3], fr.offset) and raw[3] == np.float32(fr.norme)
    assert np.array_equal(raw[4:].reshape(-1, 3), Pn)
    assert out[0] == f"points {len(Pn)} normals {len(Pn)}"
    # Pn::order of the normalised positions = pointio.morton_order
    assert np.array_equal(np.fromfile(str(tmp_path / "c.order"), np.float32).reshape(-1, 3), Pn[pointio.morton_order(Pn)])
    # the C++ save of the normalised set = the Python write of it in the file frame
    P2, N2 = pointio.read_points(str(tmp_path / "c.ply"))
    pointio.write_points(str(tmp_path / "py.ply"), Pn, g["normals"], fr)
    assert open(str(tmp_path / "c.ply"), "rb").read() == open(str(tmp_path / "py.ply"), "rb").read()
    assert np.array_equal(N2, g["normals"])
    prm = pointio.read_params(g["params"].tobytes())
    f = dict(kv.split(" ", 1) for kv in out[1].replace(" cutoff", "|cutoff").replace(" min_depth", "|min_depth")
             .replace(" iterations", "|iterations").replace(" computed", "|computed").replace(" n_sigmas", "|n_sigmas")
             .replace(" factor", "|factor").replace(" begin", "|begin").split("|"))
    assert int(f["type"]) == prm.kernel_type and np.float32(f["cutoff"]) == np.float32(prm.cutoff)
    assert int(f["min_depth"]) == prm.min_depth and int(f["iterations"]) == prm.max_iterations
    assert int(f["n_sigmas"]) == prm.n_computed_sigmas and bool(int(f["computed"])) == prm.computed_sigma
    k = prm.kernel(0.8125)
    vals = out[2].split()
    assert np.float32(vals[1]) == np.float32(k.sigma)
    assert np.array_equal(np.array(vals[3:], np.float32), np.array(k.sigmas, np.float32))
    # Pn::saveMesh with meshComments = pointio.write_mesh with mesh_comments, byte for byte
    pointio.write_mesh(str(tmp_path / "pyMesh.ply"), Pn[:4], g["normals"][:4], np.array([[0, 1, 2], [0, 2, 3]]), fr,
                       comments=pointio.mesh_comments(k, 0.8125))
    assert open(str(tmp_path / "cMesh.ply"), "rb").read() == open(str(tmp_path / "pyMesh.ply"), "rb").read()


def test_cpp_ascii_ply(tool, tmp_path):
    txt = ("ply\nformat ascii 1.0\nelement vertex 3\nproperty double x\nproperty double y\nproperty double z\n"
           "property float nx\nproperty float ny\nproperty float nz\nelement face 1\n"
           "property list uchar int vertex_indices\nend_header\n"
           "1.5 2.5 3.5 0 0 1\n-1 -2 -3 0 1 0\n0.25 0.5 0.75 1 0 0\n3 0 1 2\n")
    (tmp_path / "a.ply").write_text(txt)
    subprocess.run([tool, str(tmp_path / "a.ply"), str(tmp_path / "c")], check=True, capture_output=True)
    P, N = pointio.read_points(str(tmp_path / "a.ply"))
    Pn, fr = pointio.normalize(P)
    raw = np.fromfile(str(tmp_path / "c.norm"), np.float32)
    assert np.array_equal(raw[4:].reshape(-1, 3), Pn) and raw[3] == np.float32(fr.norme)


def test_mesh_comments_kernel_types():
    """The reference's kernel lines (scene.cpp:1498-1530) for the three kernel types."""
    from mlodsurfaces_amd.apss import Kernel
    c = pointio.mesh_comments(Kernel(type=2, sigmas=[0.1, 0.2], beginningSigma=0.01, factor=2.0), 1.0)
    assert c[2:7] == ["Multiple Gaussian", "First sigma  0.01", "Factor 2", "Number of sigmas 2",
                      "Constant_shift (unused) 0"]
    assert pointio.mesh_comments(Kernel(type=1, exponent=4.0, constant_shift=1e-3), 1.0)[2:5] == \
        ["Singular", "Exponent 4", "Constant_shift 0.001"]
    assert pointio.mesh_comments(Kernel(type=0, sigma=0.05), 0.5)[2:4] == ["Gaussian", "Sigma 0.1"]
    assert pointio.mesh_comments(Kernel(type=0), 1.0)[-2:] == ["Octree:", "Built using the adaptive octree method"]
    assert pointio.mesh_comments(Kernel(type=0), 1.0, "points")[-1] == "Built on input points"


@pytest.mark.gpu
def test_cpp_surface_mesh_matches_python(tmp_path):
    """tests/cpp/dc_mesh.cpp (mlod_dc.hpp's computeDualContouring and Pn::saveMesh) writes the
    same mesh file, byte for byte, as pointio.project_model(..., mesh_depth=0) for the Face model."""
    from mlodsurfaces_amd import _lib
    exe = str(tmp_path / "dc_mesh")
    libdir = os.path.dirname(_lib.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-O2", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "dc_mesh.cpp"), "-o", exe, f"-L{libdir}", "-lmlod",
                    f"-Wl,-rpath,{libdir}"], check=True, capture_output=True, text=True)
    g = load_golden("real_Face.npz")
    model = str(tmp_path / "Face.ply")
    pointio.write_points(model, g["points"], g["normals"])
    (tmp_path / "Face.params").write_bytes(g["params"].tobytes())
    r = subprocess.run([exe, model, str(tmp_path / "Face.params"), str(tmp_path / "cpp.ply")], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    pointio.project_model(model, log=lambda m: None, mesh_depth=0)
    a = open(str(tmp_path / "cpp.ply"), "rb").read()
    b = open(str(tmp_path / "FaceMesh.ply"), "rb").read()
    assert len(a) > 100000 and a == b
    # the octree built on the points (computeDualContouringOnPoints, uniformizeFromPoints)
    r = subprocess.run([exe, model, str(tmp_path / "Face.params"), str(tmp_path / "cpp_pts.ply"), "points"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    pointio.project_model(model, log=lambda m: None, mesh_depth=0, mesh_octree="points")
    a = open(str(tmp_path / "cpp_pts.ply"), "rb").read()
    b = open(str(tmp_path / "FaceMesh.ply"), "rb").read()
    assert len(a) > 100000 and a == b


@pytest.mark.gpu
def test_cpp_model_projection_matches_python(tmp_path):
    """tests/cpp/pn_project.cpp (the C++ headers end to end) writes the same PLY bytes as
    pointio.project_model for the Face model with its settings."""
    from mlodsurfaces_amd import _lib
    exe = str(tmp_path / "pn_project")
    libdir = os.path.dirname(_lib.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-O2", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "pn_project.cpp"), "-o", exe, f"-L{libdir}", "-lmlod",
                    f"-Wl,-rpath,{libdir}"], check=True, capture_output=True, text=True)
    g = load_golden("real_Face.npz")
    model = str(tmp_path / "Face.ply")
    pointio.write_points(model, g["points"], g["normals"])
    (tmp_path / "Face.params").write_bytes(g["params"].tobytes())
    r = subprocess.run([exe, model, str(tmp_path / "Face.params"), str(tmp_path / "cpp.ply")], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    pointio.project_model(model, log=lambda m: None)
    assert open(str(tmp_path / "cpp.ply"), "rb").read() == open(str(tmp_path / "FacePoints.ply"), "rb").read()
