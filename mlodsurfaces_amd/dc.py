"""Dual contouring of the adaptive octree on the GPU (libmlod.so's mlod_dc_*): the reference's
mesh output from the projection (scene::computeDualContouring, sources/scene/scene.cpp:1255-1460).

``DualContouring(cells, max_depth, cut)`` uniformizes the cells (upsample.cpp:262-272);
``project_corners(tree)`` takes the corner field from the tree's projection (or
``set_corner_values`` from the caller); ``extract()`` returns the mesh.  ``smooth`` and
``subdivide`` are the reference's mesh passes.  ``surface_mesh`` runs the reference's sequence on a
model: adaptive octree, uniformize, corner field, extraction, vertices projected three steps onto
the surface, then twice smoothing, subdivision and projection (scene.cpp:1417-1452)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib

__all__ = ["DualContouring", "grid_centres", "smooth", "subdivide", "surface_mesh"]


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class DualContouring:
    """uniformize(max_depth, cells, ..., cut) of octree cells (upsample.cpp:262-272), or with
    from_points uniformizeFromPoints(cells) (253-259: cells of an octree built on points, no
    re-rooting, no cut); then the corner field and the extraction."""

    def __init__(self, cells, max_depth=0, cut=False, stream=None, from_points=False):
        ids = np.ascontiguousarray(cells, np.int32).reshape(-1, 4)
        h = C.c_void_p()
        check(lib().mlod_dc_create(C.byref(h)))
        self.h = h
        self._stream = stream
        if from_points:
            check(lib().mlod_dc_build_from_points(self.h, _p(ids), len(ids), stream))
        else:
            check(lib().mlod_dc_build(self.h, _p(ids), len(ids), int(max_depth), int(bool(cut)), stream))

    def close(self):
        if getattr(self, "h", None):
            lib().mlod_dc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        o = np.zeros(9, np.float64)
        check(lib().mlod_dc_info(self.h, _p(o)))
        return dict(nodes=int(o[0]), leaves=int(o[1]), corners=int(o[2]), max_depth=int(o[3]), diag=float(o[4]),
                    origin=o[5:8].astype(np.float32), extracted=bool(o[8]))

    def nodes(self):
        n = self.info()["nodes"]
        ids = np.zeros((n, 4), np.int32)
        leaf = np.zeros(n, np.uint8)
        check(lib().mlod_dc_nodes(self.h, _p(ids), _p(leaf), self._stream))
        return ids, leaf.astype(bool)

    def corners(self):
        """(integer corner positions at the finest depth, positions in the projection frame)"""
        n = self.info()["corners"]
        xyz = np.zeros((n, 3), np.int32)
        pos = np.zeros((n, 3), np.float32)
        check(lib().mlod_dc_corners(self.h, _p(xyz), _p(pos), self._stream))
        return xyz, pos

    def project_corners(self, tree):
        """Corner values from one projection step of every corner on `tree` (its current kernel,
        cutoff, min depth, mode, semantics and form)."""
        check(lib().mlod_dc_project_corners(self.h, tree.handle, self._stream))

    def set_corner_values(self, f):
        f = np.ascontiguousarray(f, np.float32)
        if len(f) != self.info()["corners"]:
            raise ValueError("one value per corner")
        check(lib().mlod_dc_set_corner_values(self.h, _p(f), self._stream))

    def corner_values(self):
        f = np.zeros(self.info()["corners"], np.float32)
        check(lib().mlod_dc_corner_values(self.h, _p(f), self._stream))
        return f

    def leaf_centres(self, depth=None):
        """Centres of the leaves at `depth` of the uniformized octree (default: its finest depth,
        which holds the generator's deepest cells), in the projection frame, in node order:
        new2bbox applied to (2 p + 1) / 2^(d + 1) in float32 (scene.cpp:1300-1315)."""
        inf = self.info()
        ids, leaf = self.nodes()
        d = inf["max_depth"] if depth is None else int(depth)
        L = ids[leaf & (ids[:, 0] == d)][:, 1:]
        m00 = np.float32(2.0 * inf["diag"])
        m03 = (np.float32(2.0) * inf["origin"] - np.float32(1.0)).astype(np.float32)
        u = (2 * L + 1).astype(np.float32) * np.float32(2.0 ** -(d + 1))
        return (m00 * u + m03).astype(np.float32)

    def refine(self, apss, steps=3, rounds=2):
        """The extracted mesh kept on the device: vertices projected `steps` steps with `apss`'s
        tree and settings, then `rounds` times smoothing, subdivision and projection
        (scene.cpp:1417-1452); returns (vertices, normals, triangles) in the projection frame."""
        return _refine(self, apss, steps, rounds)

    def extract(self):
        """(vertices in the projection frame, triangles, node of each vertex, unit-frame vertices)"""
        nv, nt = C.c_uint32(), C.c_uint32()
        check(lib().mlod_dc_extract(self.h, C.byref(nv), C.byref(nt), self._stream))
        v = np.zeros((nv.value, 3), np.float32)
        u = np.zeros((nv.value, 3), np.float32)
        vn = np.zeros(nv.value, np.uint32)
        t = np.zeros((nt.value, 3), np.uint32)
        check(lib().mlod_dc_mesh(self.h, _p(v), _p(u), _p(vn), _p(t), self._stream))
        return v, t, vn, u


def _refine(dc, apss, steps, rounds):
    apss._sync_params()
    nv, nt = C.c_uint32(), C.c_uint32()
    check(lib().mlod_dc_refine(dc.h, apss.m_apssOctree.handle, int(steps), int(rounds), C.byref(nv), C.byref(nt),
                               dc._stream))
    V = np.zeros((nv.value, 3), np.float32)
    N = np.zeros((nv.value, 3), np.float32)
    F = np.zeros((nt.value, 3), np.uint32)
    check(lib().mlod_dc_refined_mesh(dc.h, _p(V), _p(N), _p(F), dc._stream))
    return V, N, F


def grid_centres(apss, depth, cap=None, seed=0):
    """The query points of the dual-contouring grid: the adaptive octree of `apss` generated with
    max depth `depth` (its crossed cells), uniformized, and the centres of the leaves at the
    uniformized octree's finest depth (the generator's deepest cells, whatever the re-rooting),
    in node order; with `cap`, a seeded random subset of at most `cap` of them, kept in node
    order."""
    from .apss import OctreeGenerator
    gen = OctreeGenerator(apss, getattr(apss, "kernel", None))
    total = []
    gen.generate(depth, total)
    gen.close()
    dc = DualContouring(total[0], depth)
    q = dc.leaf_centres()
    dc.close()
    if cap is not None and len(q) > cap:
        keep = np.sort(np.random.default_rng(seed).choice(len(q), int(cap), replace=False))
        q = q[keep]
    return q


def smooth(V, F):
    """mesh::smoothMesh (upsample.h:55-76) on the GPU, the reference's result bit for bit."""
    V = np.ascontiguousarray(V, np.float32)
    F = np.ascontiguousarray(F, np.uint32)
    out = np.zeros_like(V)
    check(lib().mlod_mesh_smooth(_p(V), len(V), _p(F), len(F), _p(out), None))
    return out


def subdivide(V, F):
    """mesh::subdivideMesh (upsample.h:78-133) on the GPU: (vertices, triangles)."""
    V = np.ascontiguousarray(V, np.float32)
    F = np.ascontiguousarray(F, np.uint32)
    cap = len(V) + 3 * len(F)
    out = np.zeros((cap, 3), np.float32)
    T = np.zeros((4 * len(F), 3), np.uint32)
    n = C.c_uint32()
    check(lib().mlod_mesh_subdivide(_p(V), len(V), _p(F), len(F), _p(out), cap, _p(T), C.byref(n), None))
    return out[:n.value], T


def surface_mesh(apss, depth, cut=False, reproject_steps=3, refine=2, log=None, points=None):
    """The reference's mesh of the surface for an APSS plugin `apss` (Fastapss or GlobalAPSS with
    its kernel and settings installed): OctreeGenerator.generate(depth) with its crossed cells,
    uniformize, corner field by projection, dual contouring, then the vertices projected
    `reproject_steps` steps, then `refine` rounds of smoothing, subdivision and projection
    (scene.cpp:1417-1452).  With `points` (the input points for the reference's usual octree, or
    its projected points for the octree on upsampled points), the cells are those of
    OctreeGenerator.generateFromPoints(depth, points), uniformized from points
    (scene.cpp:1279-1299).  Returns (vertices, normals, triangles, info)."""
    import time
    from .apss import OctreeGenerator
    t0 = time.perf_counter()
    gen = OctreeGenerator(apss, apss.kernel if hasattr(apss, "kernel") else None)
    if points is None:
        total = []
        gen.generate(depth, total)
        cells = total[0]
    else:
        cells = gen.generateFromPoints(depth, points)
    gen.close()
    t1 = time.perf_counter()
    dc = DualContouring(cells, depth, cut, from_points=points is not None)
    inf = dc.info()
    t2 = time.perf_counter()
    apss._sync_params()
    dc.project_corners(apss.m_apssOctree)
    t3 = time.perf_counter()
    nv, nt = C.c_uint32(), C.c_uint32()
    check(lib().mlod_dc_extract(dc.h, C.byref(nv), C.byref(nt), dc._stream))
    t4 = time.perf_counter()
    V, N, F = dc.refine(apss, reproject_steps, refine)
    t5 = time.perf_counter()
    info = dict(cells=len(cells), nodes=inf["nodes"], leaves=inf["leaves"], corners=inf["corners"],
                dc_vertices=nv.value, dc_triangles=nt.value, vertices=len(V), triangles=len(F),
                ms_generate=(t1 - t0) * 1e3, ms_uniformize=(t2 - t1) * 1e3, ms_corners=(t3 - t2) * 1e3,
                ms_extract=(t4 - t3) * 1e3, ms_refine=(t5 - t4) * 1e3)
    if log:
        log(f"surface mesh depth {depth}: {info['cells']} crossed cells -> {info['nodes']} nodes, {info['leaves']} "
            f"leaves, {info['corners']} corners -> {info['vertices']} vertices, {info['triangles']} triangles")
    dc.close()
    return V, N, F, info
