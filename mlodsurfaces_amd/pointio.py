"""Point-set files and per-model settings of the reference application, around the projection.

* ``read_points`` / ``write_points``: oriented point clouds as PLY (binary little-endian or
  ASCII, vertex properties x y z [nx ny nz]) and PN (raw float32 records x y z nx ny nz),
  as ``Pn::load`` / ``Pn::save`` read and write them (sources/apss/Pn.cpp:102-164, 201-245).
  A written PLY has the reference's header and record layout; positions are mapped back from the
  normalised frame with ``p * norme + offset`` (``Pn::save``, ``scene::saveProjectedPoints``,
  sources/scene/scene.cpp:1622-1661).
* ``normalize``: the frame the application works in (``Pn::normalize``, Pn.cpp:262-290), in the
  same float32 operations and order: offset = running float sum / n, box about the offset grown by
  0.5 % of its diagonal, norme = its largest |coordinate|, p' = (p - offset) / norme.
* ``read_params``: the 134-byte ``.params`` file of a model (``scene::loadParameters``,
  scene.cpp:1183-1250) and ``Params.kernel(radius)``, the kernel it describes with sigmas
  generated and scaled as ``Kernel::updateSigmas`` does (sources/apss/Kernel.cpp:110-128).
* ``write_timings``: the ``.timings`` summary (``MainWindow::save_timings``,
  sources/interface/mainwindow.cpp:634-705) for a projection run.

``python -m mlodsurfaces_amd.pointio model.ply`` runs the reference's automatic projection of a
model's own points with its ``.params`` (normalise, octree, ``maxNbIterations`` steps) on the GPU
and writes ``<model>Points.ply`` and ``<model>.timings``.
"""
from __future__ import annotations

import argparse
import math
import os
import struct
import sys
import time
from dataclasses import dataclass, field

import numpy as np

__all__ = ["read_points", "write_points", "write_mesh", "mesh_comments", "read_mesh", "normalize", "Frame",
           "morton_codes", "morton_order", "read_params", "Params", "write_timings", "compare_point_set"]

_PLY_TYPES = {
    "char": "i1", "int8": "i1", "uchar": "u1", "uint8": "u1", "short": "i2", "int16": "i2", "ushort": "u2",
    "uint16": "u2", "int": "i4", "int32": "i4", "uint": "u4", "uint32": "u4", "float": "f4", "float32": "f4",
    "double": "f8", "float64": "f8",
}


def _parse_ply_header(f):
    if f.readline().strip() != b"ply":
        raise ValueError("not a PLY file")
    fmt, elements = None, []
    while True:
        line = f.readline()
        if not line:
            raise ValueError("PLY header without end_header")
        tok = line.decode("ascii", "replace").split()
        if not tok or tok[0] in ("comment", "obj_info"):
            continue
        if tok[0] == "format":
            fmt = tok[1]
        elif tok[0] == "element":
            elements.append({"name": tok[1], "count": int(tok[2]), "props": []})
        elif tok[0] == "property":
            if not elements:
                raise ValueError("PLY property before any element")
            if tok[1] == "list":
                elements[-1]["props"].append((tok[4], ("list", _PLY_TYPES[tok[2]], _PLY_TYPES[tok[3]])))
            else:
                elements[-1]["props"].append((tok[2], _PLY_TYPES[tok[1]]))
        elif tok[0] == "end_header":
            break
    if fmt not in ("binary_little_endian", "ascii", "binary_big_endian"):
        raise ValueError(f"unsupported PLY format {fmt}")
    return fmt, elements


def _read_ply(path):
    with open(path, "rb") as f:
        fmt, elements = _parse_ply_header(f)
        body = f.read()
    vertex = None
    pos = 0
    text = body.decode("ascii", "replace").split() if fmt == "ascii" else None
    tpos = 0
    for el in elements:
        scalar = all(not isinstance(t, tuple) for _, t in el["props"])
        if fmt == "ascii":
            if el["name"] == "vertex":
                if not scalar:
                    raise ValueError("list properties in the vertex element are not supported")
                k = len(el["props"])
                vals = np.array(text[tpos:tpos + k * el["count"]], dtype=np.float64).reshape(el["count"], k)
                vertex = {name: vals[:, j] for j, (name, _) in enumerate(el["props"])}
                tpos += k * el["count"]
            else:
                for _ in range(el["count"]):
                    for _, t in el["props"]:
                        if isinstance(t, tuple):
                            n = int(text[tpos])
                            tpos += 1 + n
                        else:
                            tpos += 1
            continue
        end = "<" if fmt == "binary_little_endian" else ">"
        if scalar:
            dt = np.dtype([(name, end + t) for name, t in el["props"]])
            arr = np.frombuffer(body, dtype=dt, count=el["count"], offset=pos)
            pos += dt.itemsize * el["count"]
            if el["name"] == "vertex":
                vertex = {name: arr[name].astype(np.float64) for name, _ in el["props"]}
        else:
            if el["name"] == "vertex":
                raise ValueError("list properties in the vertex element are not supported")
            for _ in range(el["count"]):  # faces and the like after the vertices: skipped
                for _, t in el["props"]:
                    if isinstance(t, tuple):
                        n = np.frombuffer(body, end + t[1], 1, pos)[0]
                        pos += np.dtype(t[1]).itemsize + int(n) * np.dtype(t[2]).itemsize
                    else:
                        pos += np.dtype(t).itemsize
        if el["name"] == "vertex":
            break
    if vertex is None or not all(c in vertex for c in "xyz"):
        raise ValueError("PLY file without vertex x y z")
    P = np.stack([vertex["x"], vertex["y"], vertex["z"]], 1).astype(np.float32)
    N = None
    if all(c in vertex for c in ("nx", "ny", "nz")):
        N = np.stack([vertex["nx"], vertex["ny"], vertex["nz"]], 1).astype(np.float32)
    return P, N


def read_points(path: str):
    """Positions and normals (float32, n x 3; normals None when the file has none) of a PLY or
    PN file (Pn::load, Pn.cpp:102-164: a PN file is records of six float32; a PLY file gives
    the vertex properties x y z and nx ny nz)."""
    if path.endswith("pn"):
        raw = np.fromfile(path, dtype="<f4")
        n = len(raw) // 6
        rec = raw[:6 * n].reshape(n, 6)
        return np.ascontiguousarray(rec[:, :3]), np.ascontiguousarray(rec[:, 3:])
    if path.endswith("ply"):
        return _read_ply(path)
    raise ValueError(f"{path}: neither .ply nor .pn")


@dataclass
class Frame:
    """The normalised frame: file position = p * norme + offset."""
    offset: np.ndarray = field(default_factory=lambda: np.zeros(3, np.float32))
    norme: float = 1.0


def morton_codes(P: np.ndarray, size_of_grid: int = 64) -> np.ndarray:
    """The Morton codes of Pn::order (sources/apss/Pn.cpp:388-417) for points P: per axis the cell
    floor((p - mini) / norme * size) clamped to [0, size - 1], with mini / maxi the smallest /
    largest coordinate of the point set's box over all axes and norme = maxi - mini (float32), bits
    interleaved as _pdep_u32 with the masks 0x92492492 (x), 0x49249249 (y), 0x24924924 (z): bit k
    of x at 3k + 1, of y at 3k, of z at 3k + 2."""
    P = np.ascontiguousarray(P, np.float32)
    lo, hi = P.min(0), P.max(0)
    mini, maxi = np.float32(min(lo)), np.float32(max(hi))
    norme = np.float32(maxi - mini)
    v = np.floor(((P - mini) / norme) * np.float32(size_of_grid)).astype(np.int64)
    v = np.clip(v, 0, size_of_grid - 1).astype(np.uint32)
    code = np.zeros(len(P), np.uint32)
    for k in range(11):
        for axis, off in ((0, 1), (1, 0), (2, 2)):
            pos = 3 * k + off
            if pos < 32:
                code |= ((v[:, axis] >> np.uint32(k)) & np.uint32(1)) << np.uint32(pos)
    return code


def morton_order(P: np.ndarray, size_of_grid: int = 64) -> np.ndarray:
    """The permutation of Pn::order (Pn.cpp:388-445): points sorted by their Morton code
    (morton_codes).  The reference's std::sort leaves points of equal code in an unspecified
    order; here they keep their input order.  The reference reorders the positions only."""
    return np.argsort(morton_codes(P, size_of_grid), kind="stable")


def normalize(P: np.ndarray):
    """Pn::normalize (Pn.cpp:262-290) in its float32 operations and order: returns the points in
    the normalised frame and the Frame (offset, norme) that maps them back."""
    P = np.ascontiguousarray(P, np.float32)
    n = len(P)
    if n == 0:
        raise ValueError("empty point set")
    off = np.add.accumulate(P, axis=0, dtype=np.float32)[-1]  # p0 + p1 + ... in float32, in order
    off = (off / np.float32(n)).astype(np.float32)
    maxi = (P.max(0) - off).astype(np.float32)
    mini = (P.min(0) - off).astype(np.float32)
    d = (maxi - mini).astype(np.float32)
    length = np.float32(math.sqrt(np.float32(np.float32(np.float32(d[0] * d[0]) + np.float32(d[1] * d[1])) + np.float32(d[2] * d[2]))))
    small = np.float32(length * np.float32(0.005))
    maxi = (maxi + small).astype(np.float32)
    mini = (mini - small).astype(np.float32)
    norme = np.float32(max(np.abs(maxi).max(), np.abs(mini).max()))
    Pn = ((P - off).astype(np.float32) / norme).astype(np.float32)
    return Pn, Frame(off, float(norme))


def write_points(path: str, P: np.ndarray, N: np.ndarray | None = None, frame: Frame | None = None):
    """Writes positions (mapped back with frame: p * norme + offset, in float32) and normals as
    Pn::save / scene::saveProjectedPoints do: a binary little-endian PLY with the properties
    x y z [nx ny nz], or a PN file of six-float records."""
    P = np.ascontiguousarray(P, np.float32)
    if frame is not None:
        P = (P * np.float32(frame.norme) + np.asarray(frame.offset, np.float32)).astype(np.float32)
    if N is not None:
        N = np.ascontiguousarray(N, np.float32)
        if N.shape != P.shape:
            raise ValueError("positions and normals differ in shape")
    if path.endswith("pn"):
        if N is None:
            N = np.zeros_like(P)
        np.concatenate([P, N], 1).astype("<f4").tofile(path)
        return
    if not path.endswith("ply"):
        raise ValueError(f"{path}: neither .ply nor .pn")
    head = ["ply", "format binary_little_endian 1.0", f"element vertex {len(P)}",
            "property float x", "property float y", "property float z"]
    if N is not None:
        head += ["property float nx", "property float ny", "property float nz"]
    head.append("end_header")
    rec = P if N is None else np.concatenate([P, N], 1)
    with open(path, "wb") as f:
        f.write(("\n".join(head) + "\n").encode("ascii"))
        f.write(rec.astype("<f4").tobytes())


def write_mesh(path: str, V: np.ndarray, N: np.ndarray, F: np.ndarray, frame: Frame | None = None,
               comments: list | None = None):
    """The surface mesh file of scene::computeDualContouring (scene.cpp:1489-1560): a binary
    little-endian PLY with comment lines, vertices x y z nx ny nz (positions mapped back with the
    frame) and faces as a uchar count 3 and three int indices."""
    V = np.ascontiguousarray(V, np.float32)
    if frame is not None:
        V = (V * np.float32(frame.norme) + np.asarray(frame.offset, np.float32)).astype(np.float32)
    N = np.ascontiguousarray(N, np.float32)
    F = np.ascontiguousarray(F, np.uint32)
    head = ["ply", "format binary_little_endian 1.0"] + [f"comment {c}" for c in (comments or [])] + [
        f"element vertex {len(V)}", "property float x", "property float y", "property float z",
        "property float nx", "property float ny", "property float nz", f"element face {len(F)}",
        "property list uchar int vertex_indices", "end_header"]
    face = np.zeros(len(F), dtype=np.dtype([("n", "u1"), ("v", "<i4", (3,))]))
    face["n"] = 3
    face["v"] = F.astype(np.int32)
    with open(path, "wb") as f:
        f.write(("\n".join(head) + "\n").encode("ascii"))
        f.write(np.concatenate([V, N], 1).astype("<f4").tobytes())
        f.write(face.tobytes())


_OCTREE_LINES = {"adaptive": "Built using the adaptive octree method", "points": "Built on input points",
                 "upsampled": "Built on upsampled points"}


def mesh_comments(kernel, radius: float, octree: str = "adaptive") -> list:
    """The comment lines the reference writes into its mesh file (scene.cpp:1498-1530): the
    kernel's parameters (as C++ streams print floats: six significant digits) and the octree
    method ("adaptive", "points" for the octree on the input points, "upsampled"), for a kernel
    whose sigmas were scaled by a cloud of bounding-box radius `radius`."""
    f32 = np.float32
    g = lambda v: format(float(f32(v)), "g")  # noqa: E731
    c = ["Created by mlodsurfaces_amd", "Kernel parameters:"]
    if kernel.type == 2:
        c += ["Multiple Gaussian", f"First sigma  {g(kernel.beginningSigma)}", f"Factor {g(kernel.factor)}",
              f"Number of sigmas {len(kernel.sigmas)}", f"Constant_shift (unused) {g(kernel.constant_shift)}"]
    elif kernel.type == 1:
        c += ["Singular", f"Exponent {g(kernel.exponent)}", f"Constant_shift {g(kernel.constant_shift)}"]
    elif kernel.type == 0:
        c += ["Gaussian", f"Sigma {g(f32(kernel.sigma) / f32(radius))}", f"Constant_shift {g(kernel.constant_shift)}"]
    else:
        c += ["Kernel unknown"]
    return c + ["Octree:", _OCTREE_LINES[octree]]


def read_mesh(path: str):
    """Vertices, normals (or None) and triangles of a PLY mesh with a vertex and a face element."""
    with open(path, "rb") as f:
        fmt, elements = _parse_ply_header(f)
        body = f.read()
    if fmt != "binary_little_endian":
        raise ValueError("read_mesh reads binary little-endian PLY")
    P, N = _read_ply(path)
    pos = 0
    F = None
    for el in elements:
        if el["name"] == "vertex":
            pos += sum(np.dtype(t).itemsize for _, t in el["props"]) * el["count"]
        elif el["name"] == "face":
            (_, (_, ct, it)), = el["props"]
            dt = np.dtype([("n", ct), ("v", "<" + it, (3,))])
            arr = np.frombuffer(body, dtype=dt, count=el["count"], offset=pos)
            if np.any(arr["n"] != 3):
                raise ValueError("read_mesh reads triangles")
            F = arr["v"].astype(np.uint32)
    return P, N, F


# scene::loadParameters field order (scene.cpp:1183-1250): camera quaternion (4 floats) and
# translation (3 floats) and distance; sigma; five fixed sigmas; beginningSigma; factor;
# constant_shift; exponent; cutoff (the scaling-protection sphere); nbOfComputedSigmas;
# maxNbIterations; maxDepth; minDepth; method (1 = MLoD / fast, 0 = global APSS); kernel type
# (Kernel.h numbering); computedSigma; cut; nbOfPoints; maxUpsamples; six floats of offset.
_PARAMS = struct.Struct("<4f3ff" + "f5f" + "4f" + "f" + "5I" + "i" + "??" + "2I" + "6f")


@dataclass
class Params:
    camera_quat: tuple
    camera_translation: tuple
    camera_dist: float
    sigma: float
    fixed_sigmas: tuple
    beginning_sigma: float
    factor: float
    constant_shift: float
    exponent: float
    cutoff: float
    n_computed_sigmas: int
    max_iterations: int
    max_depth: int
    min_depth: int
    method: int
    kernel_type: int
    computed_sigma: bool
    cut: bool
    n_points: int
    max_upsamples: int
    offset: tuple

    def kernel(self, radius: float):
        """The apss.Kernel these settings describe for a cloud of bounding-box radius `radius`:
        sigmas generated and scaled as Kernel::updateSigmas (Kernel.cpp:110-128) does after
        scene::loadParameters."""
        from .apss import Kernel
        k = Kernel(type=int(self.kernel_type), sigma=float(self.sigma), exponent=float(self.exponent),
                   constant_shift=float(self.constant_shift), beginningSigma=float(self.beginning_sigma),
                   factor=float(self.factor), nbOfComputedSigmas=int(self.n_computed_sigmas),
                   fixedSigmas=list(self.fixed_sigmas))
        return k.updateSigmas(self.computed_sigma, radius)


def read_params(path_or_bytes) -> Params:
    """Decodes a model's .params file (scene::loadParameters, scene.cpp:1183-1250)."""
    b = path_or_bytes if isinstance(path_or_bytes, (bytes, bytearray)) else open(path_or_bytes, "rb").read()
    if len(b) < _PARAMS.size:
        raise ValueError(f".params file of {len(b)} bytes, {_PARAMS.size} expected")
    v = _PARAMS.unpack_from(b)
    if v[24] not in (0, 1, 2):
        raise ValueError(f"unknown kernel type {v[24]}")
    return Params(camera_quat=tuple(v[0:4]), camera_translation=tuple(v[4:7]), camera_dist=v[7], sigma=v[8],
                  fixed_sigmas=tuple(v[9:14]), beginning_sigma=v[14], factor=v[15], constant_shift=v[16],
                  exponent=v[17], cutoff=v[18], n_computed_sigmas=v[19], max_iterations=v[20], max_depth=v[21],
                  min_depth=v[22], method=v[23], kernel_type=v[24], computed_sigma=bool(v[25]), cut=bool(v[26]),
                  n_points=v[27], max_upsamples=v[28], offset=tuple(v[29:35]))


def write_timings(path: str, model: str, n_input: int, n_projected: int, step_ms: list, gpu: bool = True,
                  method: str = "FastAPSS", max_depth: int = 0):
    """The .timings summary of MainWindow::save_timings (mainwindow.cpp:634-705) for the parts
    this library runs: per-point projection time (mean +- deviation over the steps, ms per
    projected point) and the octree depth; the adaptive-octree and upsampling lines are 0."""
    t = np.asarray(step_ms, np.float64)
    avg = t.sum() / (len(t) * n_projected)
    sd = math.sqrt(max(float((t * t).sum()) / (len(t) * n_projected * n_projected) - avg * avg, 0.0))
    lines = [f"File: {model}", f"Nb of input points: {n_input}", f"Nb of projected points: {n_projected}",
             f"{'GPU' if gpu else 'CPU'} - Mode: {method}",
             f"Timings for projection per point: {avg:g}+-{sd:g}ms", "Timings for Morton Code: 0+-0ms",
             f"Octrees max depth: {max_depth}", "Adaptive octree: 0", "Usual octree: 0", "Uniformize octree: 0",
             "Crossed nodes: 0", "Nodes created on projected points: 0"]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def compare_point_set(apss, points, frame: Frame | None = None, n_iterations: int = 1):
    """scene::comparePointSet (scene.cpp:1663-1720): a second point set (a path, or (n, 3) points
    in the file frame) mapped into the model's normalised frame, projected `n_iterations` steps
    with `apss`'s tree and settings, and its signed distances (p - p')·n' to the surface, also
    divided by their largest magnitude (the values the reference colours with, in [-1, 1]).
    The reference's GPU plugin projects its uploaded queries there rather than the loaded set;
    this projects the loaded set.  Returns (values, distances), distances in the normalised frame."""
    P = read_points(points)[0] if isinstance(points, str) else np.asarray(points, np.float32)
    if frame is not None:
        P = ((P - np.asarray(frame.offset, np.float32)) / np.float32(frame.norme)).astype(np.float32)
    op, on = apss.project_points(P, n_iterations)
    d = np.einsum("ij,ij->i", (P - op).astype(np.float64), on.astype(np.float64)).astype(np.float32)
    norme = max(abs(float(d.max())), abs(float(d.min()))) if len(d) else 0.0
    return (d / np.float32(norme) if norme > 0 else d), d


def project_model(model: str, params_path: str | None = None, out: str | None = None, precision="fp32c",
                  semantics="ref_gpu", log=print, mesh_depth: int | None = None, mesh_octree: str = "adaptive"):
    """The reference's automatic projection of a model's own points (mainwindow.cpp:180-220,
    scene.cpp:916-966) with its .params settings: read, normalise, build the octree, project the
    points maxNbIterations steps, write <model>Points.ply (file frame) and <model>.timings; with
    mesh_depth (0: the .params maxDepth), also the dual-contouring surface mesh <model>Mesh.ply
    (scene::computeDualContouring) on the adaptive octree, or with mesh_octree="points" on the
    octree built on the normalised input points (the reference's usual octree)."""
    from . import apss as A
    from . import hiprt
    P, N = read_points(model)
    if N is None:
        raise ValueError(f"{model}: the projection needs normals")
    prm = read_params(params_path or os.path.splitext(model)[0] + ".params")
    Pn, frame = normalize(P)
    tree = A.APSSOctree().build(Pn, N)
    K = prm.kernel(tree.radius())
    cls = A.Fastapss if prm.method else A.GlobalAPSS
    ap = cls(tree, K)
    ap.setScalingProtectionSphere(prm.cutoff)
    ap.setMinDepth(prm.min_depth)
    ap.set_semantics(semantics)
    ap.set_precision(precision)
    ap.copyPointsToGPU(len(Pn), Pn)
    op = np.zeros_like(Pn)
    on = np.zeros_like(Pn)
    steps = []
    for it in range(max(prm.max_iterations, 1)):
        hiprt.device_synchronize()
        t0 = time.perf_counter()
        last = it == max(prm.max_iterations, 1) - 1
        ap.project(len(Pn), op if last else None, on if last else None, 1, K)
        hiprt.device_synchronize()
        steps.append((time.perf_counter() - t0) * 1e3)
    base = os.path.splitext(out or model)[0]
    write_points(base + "Points.ply", op, on, frame)
    if mesh_depth is not None:
        from .dc import surface_mesh
        V, N, F, info = surface_mesh(ap, mesh_depth if mesh_depth > 0 else prm.max_depth,
                                     points=Pn if mesh_octree == "points" else None)
        write_mesh(base + "Mesh.ply", V, N, F, frame, comments=mesh_comments(K, tree.radius(), mesh_octree))
        log(f"{model}: surface mesh of {info['triangles']} triangles at depth {mesh_depth or prm.max_depth} -> "
            f"{base}Mesh.ply")
    write_timings(base + ".timings", model, len(P), len(Pn), steps, gpu=True,
                  method="FastAPSS" if prm.method else "Global APSS", max_depth=prm.max_depth)
    log(f"{model}: {len(P)} points, {len(K.sigmas) or 1} sigmas, cutoff {prm.cutoff:g}, min depth {prm.min_depth}, "
        f"{len(steps)} steps in {sum(steps):.1f} ms -> {base}Points.ply, {base}.timings")
    return op, on, frame, prm


def main(argv=None):
    ap = argparse.ArgumentParser(description="project a model's own points with its .params settings")
    ap.add_argument("model", help=".ply or .pn file")
    ap.add_argument("--params", default=None, help=".params file (default: next to the model)")
    ap.add_argument("--out", default=None, help="output base name (default: the model's)")
    ap.add_argument("--precision", default="fp32c", choices=["fp32", "fp32c", "fp64"])
    ap.add_argument("--semantics", default="ref_gpu", choices=["ref_gpu", "ref_cpu"])
    ap.add_argument("--mesh", type=int, default=None, metavar="DEPTH",
                    help="also write the surface mesh at this depth (0: the .params maxDepth)")
    ap.add_argument("--mesh-octree", default="adaptive", choices=["adaptive", "points"],
                    help="octree of the mesh: the adaptive octree, or the octree built on the input points")
    a = ap.parse_args(argv)
    project_model(a.model, a.params, a.out, a.precision, a.semantics, log=lambda m: print(m, file=sys.stderr),
                  mesh_depth=a.mesh, mesh_octree=a.mesh_octree)


if __name__ == "__main__":
    main()
