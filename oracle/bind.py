"""ctypes bindings for the oracle (liboracle.so) and the reference CPU path (_ref/libmlodref.so).

TEST INFRASTRUCTURE — only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
import this module, and only to check or to time the CPU baseline; the product never does.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libmlodref.so")
# the same reference sources with the reference's release flags (the CPU baseline's build)
REF_SO_RELEASE = os.path.join(HERE, "_ref", "libmlodref_release.so")

MODES = {"mlod": 0, "global": 1, "knn": 2, "ball": 3}
SEMANTICS = {"ref_gpu": 0, "ref_cpu": 1}


class KernelC(C.Structure):
    _fields_ = [("type", C.c_int32), ("sigma", C.c_float), ("exponent", C.c_float), ("constant_shift", C.c_float),
                ("n_sigmas", C.c_int32), ("sigmas", C.c_float * 32)]


class ParamsC(C.Structure):
    _fields_ = [("cutoff", C.c_float), ("min_depth", C.c_uint32), ("semantics", C.c_int32), ("mode", C.c_int32),
                ("knn", C.c_uint32), ("ball_radius", C.c_float)]


def make_kernel(type=2, sigma=0.0, exponent=1.0, constant_shift=0.0, sigmas=()):
    k = KernelC()
    k.type = int(type)
    k.sigma = float(sigma)
    k.exponent = float(exponent)
    k.constant_shift = float(constant_shift)
    k.n_sigmas = len(sigmas)
    for i, s in enumerate(sigmas):
        k.sigmas[i] = float(s)
    return k


def make_params(cutoff=2.0, min_depth=0, semantics="ref_cpu", mode="mlod", knn=20, ball_radius=0.1):
    p = ParamsC()
    p.cutoff = float(cutoff)
    p.min_depth = int(min_depth)
    p.semantics = SEMANTICS[semantics] if isinstance(semantics, str) else int(semantics)
    p.mode = MODES[mode] if isinstance(mode, str) else int(mode)
    p.knn = int(knn)
    p.ball_radius = float(ball_radius)
    return p


def _ptr(a, t):
    return a.ctypes.data_as(C.POINTER(t))


_vp = C.c_void_p
_f = C.POINTER(C.c_float)
_d = C.POINTER(C.c_double)
_i = C.POINTER(C.c_int32)
_u = C.POINTER(C.c_uint32)
_u64 = C.POINTER(C.c_uint64)

_orc = None
_ref = {}

# void (*)(const float* xyz, uint32_t n, float* out_pts, float* out_nmls, void* user)
PROJ_CB = C.CFUNCTYPE(None, _f, C.c_uint32, _f, _f, _vp)


def oracle_lib():
    global _orc
    if _orc is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"{ORACLE_SO} missing: run `make -C oracle`")
        L = C.CDLL(ORACLE_SO)
        L.orc_tree_build.restype = _vp
        L.orc_tree_build.argtypes = [_f, _f, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int32, C.c_uint32, _f]
        L.orc_tree_free.argtypes = [_vp]
        L.orc_tree_num_nodes.restype = C.c_uint32
        L.orc_tree_num_nodes.argtypes = [_vp]
        L.orc_tree_ties.restype = C.c_uint64
        L.orc_tree_ties.argtypes = [_vp]
        L.orc_tree_bbox.argtypes = [_vp, _f]
        L.orc_tree_nodes.argtypes = [_vp, _i, _f, _f, _f, _i, _i, _i, _i, _u, _u64]
        L.orc_tree_leaf_indices.argtypes = [_vp, _u]
        L.orc_tree_areas.argtypes = [_vp, _f]
        L.orc_tree_node_moments.argtypes = [_vp, _f, _d]
        L.orc_weight.restype = C.c_double
        L.orc_weight.argtypes = [C.POINTER(KernelC), _f, _f, C.c_int32]
        L.orc_project.restype = C.c_int32
        L.orc_project.argtypes = [_vp, C.POINTER(KernelC), C.POINTER(ParamsC), C.c_int32, _f, C.c_uint32, C.c_uint32,
                                  C.c_float, _d, _d, C.c_int32]
        L.orc_fit.argtypes = [_d, _d, C.c_double, C.c_int32, _d, _d]
        L.orc_stats.restype = C.c_int32
        L.orc_stats.argtypes = [_vp, C.POINTER(KernelC), C.POINTER(ParamsC), _f, C.c_uint32, _d]
        L.orc_gen_create.restype = _vp
        L.orc_gen_create.argtypes = [PROJ_CB, _vp]
        L.orc_gen_free.argtypes = [_vp]
        L.orc_gen_run.argtypes = [_vp, C.c_uint32]
        L.orc_gen_from_points.argtypes = [_vp, _f, C.c_uint32, C.c_uint32, _f]
        L.orc_gen_info.argtypes = [_vp, _u64]
        L.orc_gen_ids.argtypes = [_vp, C.c_int32, _i]
        L.orc_dc_build.restype = _vp
        L.orc_dc_build.argtypes = [_i, C.c_uint32, C.c_uint32, C.c_int32]
        L.orc_dc_build_points.restype = _vp
        L.orc_dc_build_points.argtypes = [_i, C.c_uint32]
        L.orc_dc_free.argtypes = [_vp]
        L.orc_dc_num_nodes.restype = C.c_uint32
        L.orc_dc_num_nodes.argtypes = [_vp]
        L.orc_dc_frame.argtypes = [_vp, _f, _i]
        L.orc_dc_nodes.argtypes = [_vp, _i, C.POINTER(C.c_uint8)]
        L.orc_dc_set_corners.argtypes = [_vp, _i, _f, C.c_uint32]
        L.orc_dc_extract.argtypes = [_vp, _u, _u]
        L.orc_dc_mesh.argtypes = [_vp, _f, _u, _u]
        L.orc_mesh_smooth.argtypes = [_f, C.c_uint32, _u, C.c_uint32]
        L.orc_mesh_subdivide.restype = C.c_uint32
        L.orc_mesh_subdivide.argtypes = [_f, C.c_uint32, _u, C.c_uint32]
        L.orc_cell_center.argtypes = [_i, _f, _f]
        L.orc_cell_ancestor.argtypes = [_i, C.c_int32, _i]
        _orc = L
    return _orc


def ref_lib(release=False):
    so = REF_SO_RELEASE if release else REF_SO
    if so not in _ref:
        if not os.path.exists(so):
            raise RuntimeError(f"{so} missing: run `make -C oracle ref` where the reference sources exist")
        L = C.CDLL(so)
        L.ref_tree_build.restype = _vp
        L.ref_tree_build.argtypes = [_f, _f, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, _f]
        L.ref_tree_free.argtypes = [_vp]
        L.ref_tree_num_nodes.restype = C.c_uint32
        L.ref_tree_num_nodes.argtypes = [_vp]
        L.ref_tree_nodes.argtypes = [_vp, _i, _f, _f, _f, _i, _i, _i, _i, _u, _u64, _u, _f]
        L.ref_tree_areas.argtypes = [_vp, _f]
        L.ref_tree_bbox.argtypes = [_vp, _f]
        L.ref_project.restype = C.c_int32
        L.ref_project.argtypes = [_vp, C.POINTER(KernelC), C.POINTER(ParamsC), _f, C.c_uint32, C.c_uint32, C.c_float, _f,
                                  _f, C.c_int32]
        L.ref_knn.argtypes = [_vp, _f, C.c_uint32, C.c_uint32, _f]
        _ref[so] = L
    return _ref[so]


def have_ref():
    return os.path.exists(REF_SO)


def _c32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class OracleTree:
    """The oracle's restated octree (CPUUtils.cpp:540-619)."""

    def __init__(self, points, normals, max_depth=12, max_pts_per_leaf=1, tie_policy=0, seed=12345, root_box=None):
        self.L = oracle_lib()
        self.p = _c32(points)
        self.n = _c32(normals)
        self.box = None if root_box is None else np.ascontiguousarray(root_box, np.float32).reshape(6)
        self.h = self.L.orc_tree_build(_ptr(self.p, C.c_float), _ptr(self.n, C.c_float), len(self.p), max_depth,
                                       max_pts_per_leaf, tie_policy, seed,
                                       None if self.box is None else _ptr(self.box, C.c_float))

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_tree_free(self.h)
            self.h = None

    @property
    def num_nodes(self):
        return int(self.L.orc_tree_num_nodes(self.h))

    @property
    def ties(self):
        return int(self.L.orc_tree_ties(self.h))

    def bbox(self):
        b = np.zeros(6, np.float32)
        self.L.orc_tree_bbox(self.h, _ptr(b, C.c_float))
        return b

    def nodes(self):
        nn = self.num_nodes
        d = dict(depth=np.zeros(nn, np.int32), bb=np.zeros((nn, 6), np.float32), center=np.zeros((nn, 3), np.float32),
                 radius=np.zeros(nn, np.float32), father=np.zeros(nn, np.int32), first_child=np.zeros(nn, np.int32),
                 next_brother=np.zeros(nn, np.int32), num_children=np.zeros(nn, np.int32),
                 leaf_count=np.zeros(nn, np.uint32), leaf_offset=np.zeros(nn, np.uint64))
        self.L.orc_tree_nodes(self.h, _ptr(d["depth"], C.c_int32), _ptr(d["bb"], C.c_float), _ptr(d["center"], C.c_float),
                              _ptr(d["radius"], C.c_float), _ptr(d["father"], C.c_int32),
                              _ptr(d["first_child"], C.c_int32), _ptr(d["next_brother"], C.c_int32),
                              _ptr(d["num_children"], C.c_int32), _ptr(d["leaf_count"], C.c_uint32),
                              _ptr(d["leaf_offset"], C.c_uint64))
        li = np.zeros(int(d["leaf_count"].sum()), np.uint32)
        self.L.orc_tree_leaf_indices(self.h, _ptr(li, C.c_uint32))
        d["leaf_idx"] = li
        return d

    def areas(self):
        a = np.zeros(len(self.p), np.float32)
        self.L.orc_tree_areas(self.h, _ptr(a, C.c_float))
        return a

    def moments(self):
        nn = self.num_nodes
        f = np.zeros((nn, 9), np.float32)
        d = np.zeros((nn, 9), np.float64)
        self.L.orc_tree_node_moments(self.h, _ptr(f, C.c_float), _ptr(d, C.c_double))
        return f, d

    def project(self, queries, kernel, params, n_iter=1, step_max=-1.0, precision=1, nthreads=None):
        q = _c32(queries)
        m = len(q)
        op = np.zeros((m, 3), np.float64)
        on = np.zeros((m, 3), np.float64)
        nt = nthreads or os.cpu_count() or 1
        r = self.L.orc_project(self.h, C.byref(kernel), C.byref(params), int(precision), _ptr(q, C.c_float), m, n_iter,
                               float(step_max), _ptr(op, C.c_double), _ptr(on, C.c_double), nt)
        assert r == 0
        return op, on

    def stats(self, queries, kernel, params):
        """FP64 weighted sums (w, wp, wn, wpn, wpp) of one MLoD or global pass per query."""
        q = _c32(queries)
        out = np.zeros((len(q), 9), np.float64)
        r = self.L.orc_stats(self.h, C.byref(kernel), C.byref(params), _ptr(q, C.c_float), len(q),
                             _ptr(out, C.c_double))
        assert r == 0
        return out


def oracle_weight(kernel, p, q, precision=1):
    L = oracle_lib()
    p = _c32(p)
    q = _c32(q)
    return L.orc_weight(C.byref(kernel), _ptr(p, C.c_float), _ptr(q, C.c_float), precision)


def oracle_fit(sums9, q, step_max=1e30, ref_gpu=False):
    L = oracle_lib()
    s = np.ascontiguousarray(sums9, np.float64)
    qq = np.ascontiguousarray(q, np.float64)
    op = np.zeros(3)
    on = np.zeros(3)
    L.orc_fit(_ptr(s, C.c_double), _ptr(qq, C.c_double), float(step_max), int(ref_gpu), _ptr(op, C.c_double),
              _ptr(on, C.c_double))
    return op, on


class RefTree:
    """The reference's own CPU APSSOctree (compiled into oracle/_ref)."""

    def __init__(self, points, normals, max_depth=12, max_pts_per_leaf=1, seed=12345, root_box=None, release=False):
        self.L = ref_lib(release)
        self.p = _c32(points)
        self.n = _c32(normals)
        self.box = None if root_box is None else np.ascontiguousarray(root_box, np.float32).reshape(6)
        self.h = self.L.ref_tree_build(_ptr(self.p, C.c_float), _ptr(self.n, C.c_float), len(self.p), max_depth,
                                       max_pts_per_leaf, seed, None if self.box is None else _ptr(self.box, C.c_float))

    def __del__(self):
        if getattr(self, "h", None):
            self.L.ref_tree_free(self.h)
            self.h = None

    @property
    def num_nodes(self):
        return int(self.L.ref_tree_num_nodes(self.h))

    def bbox(self):
        b = np.zeros(6, np.float32)
        self.L.ref_tree_bbox(self.h, _ptr(b, C.c_float))
        return b

    def nodes(self):
        nn = self.num_nodes
        d = dict(depth=np.zeros(nn, np.int32), bb=np.zeros((nn, 6), np.float32), center=np.zeros((nn, 3), np.float32),
                 radius=np.zeros(nn, np.float32), father=np.zeros(nn, np.int32), first_child=np.zeros(nn, np.int32),
                 next_brother=np.zeros(nn, np.int32), num_children=np.zeros(nn, np.int32),
                 leaf_count=np.zeros(nn, np.uint32), leaf_offset=np.zeros(nn, np.uint64),
                 leaf_idx=np.zeros(len(self.p), np.uint32), moments=np.zeros((nn, 9), np.float32))
        self.L.ref_tree_nodes(self.h, _ptr(d["depth"], C.c_int32), _ptr(d["bb"], C.c_float), _ptr(d["center"], C.c_float),
                              _ptr(d["radius"], C.c_float), _ptr(d["father"], C.c_int32),
                              _ptr(d["first_child"], C.c_int32), _ptr(d["next_brother"], C.c_int32),
                              _ptr(d["num_children"], C.c_int32), _ptr(d["leaf_count"], C.c_uint32),
                              _ptr(d["leaf_offset"], C.c_uint64), _ptr(d["leaf_idx"], C.c_uint32),
                              _ptr(d["moments"], C.c_float))
        return d

    def areas(self):
        a = np.zeros(len(self.p), np.float32)
        self.L.ref_tree_areas(self.h, _ptr(a, C.c_float))
        return a

    def project(self, queries, kernel, params, n_iter=1, step_max=-1.0, nthreads=None):
        q = _c32(queries)
        m = len(q)
        op = np.zeros((m, 3), np.float32)
        on = np.zeros((m, 3), np.float32)
        nt = nthreads or os.cpu_count() or 1
        r = self.L.ref_project(self.h, C.byref(kernel), C.byref(params), _ptr(q, C.c_float), m, n_iter, float(step_max),
                               _ptr(op, C.c_float), _ptr(on, C.c_float), nt)
        assert r == 0
        return op, on

    def knn(self, queries, k):
        q = _c32(queries)
        out = np.zeros((len(q), k, 3), np.float32)
        self.L.ref_knn(self.h, _ptr(q, C.c_float), len(q), k, _ptr(out, C.c_float))
        return out


class OracleGenerator:
    """The oracle's restated OctreeGenerator (upsample.cpp:128-144, 228-251).

    ``project(centres float32 (n, 3)) -> (points (n, 3), normals (n, 3))`` is the projection
    step of find_stable_nodesGPU (one step, unbounded step length)."""

    def __init__(self, project=None):
        self.L = oracle_lib()
        self._project = project
        self.calls = []

        def cb(xyz, n, out_p, out_n, user):
            c = np.ctypeslib.as_array(xyz, shape=(n, 3)).copy()
            p, nm = self._project(c)
            np.ctypeslib.as_array(out_p, shape=(n, 3))[:] = np.asarray(p, np.float32)
            np.ctypeslib.as_array(out_n, shape=(n, 3))[:] = np.asarray(nm, np.float32)
            self.calls.append(n)

        self._cb = PROJ_CB(cb)
        self.h = self.L.orc_gen_create(self._cb, None)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_gen_free(self.h)
            self.h = None

    def generate(self, max_depth):
        self.L.orc_gen_run(self.h, int(max_depth))
        return self.leaves()

    def generate_from_points(self, points, max_depth, box6=None):
        p = _c32(points)
        b = None if box6 is None else _ptr(np.ascontiguousarray(box6, np.float32), C.c_float)
        self.L.orc_gen_from_points(self.h, _ptr(p, C.c_float), len(p), int(max_depth), b)
        return self.leaves()

    def info(self):
        o = np.zeros(6, np.uint64)
        self.L.orc_gen_info(self.h, _ptr(o, C.c_uint64))
        return dict(leaves=int(o[0]), total=int(o[1]), max_rel_depth=int(o[2]), projected=int(o[3]), calls=int(o[4]),
                    bad=int(o[5]))

    def _ids(self, which, n):
        out = np.zeros((n, 4), np.int32)
        self.L.orc_gen_ids(self.h, which, _ptr(out, C.c_int32))
        return out

    def leaves(self):
        return self._ids(0, self.info()["leaves"])

    def total(self):
        return self._ids(1, self.info()["total"])


def oracle_cell_center(id4, box6=None):
    L = oracle_lib()
    i = np.ascontiguousarray(id4, np.int32)
    out = np.zeros(3, np.float32)
    b = None if box6 is None else _ptr(np.ascontiguousarray(box6, np.float32), C.c_float)
    L.orc_cell_center(_ptr(i, C.c_int32), b, _ptr(out, C.c_float))
    return out


def oracle_cell_ancestor(id4, depth):
    L = oracle_lib()
    i = np.ascontiguousarray(id4, np.int32)
    out = np.zeros(4, np.int32)
    L.orc_cell_ancestor(_ptr(i, C.c_int32), int(depth), _ptr(out, C.c_int32))
    return out


class OracleDC:
    """Dual contouring of the oracle (dc_oracle.cpp): uniformize the given octree cells (ids4:
    depth, x, y, z) — or uniformizeFromPoints them with from_points —, set corner values at
    max-depth integer corner positions, extract."""

    def __init__(self, ids4, max_depth=0, cut=False, from_points=False):
        self.L = oracle_lib()
        ids = np.ascontiguousarray(ids4, np.int32).reshape(-1, 4)
        if from_points:
            self.h = self.L.orc_dc_build_points(_ptr(ids, C.c_int32), len(ids))
        else:
            self.h = self.L.orc_dc_build(_ptr(ids, C.c_int32), len(ids), int(max_depth), int(bool(cut)))

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_dc_free(self.h)
            self.h = None

    def nodes(self):
        n = self.L.orc_dc_num_nodes(self.h)
        ids = np.zeros((n, 4), np.int32)
        leaf = np.zeros(n, np.uint8)
        self.L.orc_dc_nodes(self.h, _ptr(ids, C.c_int32), _ptr(leaf, C.c_uint8))
        return ids, leaf.astype(bool)

    def frame(self):
        o = np.zeros(4, np.float32)
        d = C.c_int32()
        self.L.orc_dc_frame(self.h, _ptr(o, C.c_float), C.byref(d))
        return float(o[0]), o[1:].copy(), int(d.value)

    def set_corners(self, xyz, f):
        xyz = np.ascontiguousarray(xyz, np.int32).reshape(-1, 3)
        f = np.ascontiguousarray(f, np.float32)
        self.L.orc_dc_set_corners(self.h, _ptr(xyz, C.c_int32), _ptr(f, C.c_float), len(f))

    def extract(self):
        nv, nt = C.c_uint32(), C.c_uint32()
        self.L.orc_dc_extract(self.h, C.byref(nv), C.byref(nt))
        v = np.zeros((nv.value, 3), np.float32)
        vn = np.zeros(nv.value, np.uint32)
        t = np.zeros((nt.value, 3), np.uint32)
        self.L.orc_dc_mesh(self.h, _ptr(v, C.c_float), _ptr(vn, C.c_uint32), _ptr(t, C.c_uint32))
        return v, vn, t


def oracle_mesh_smooth(V, F):
    L = oracle_lib()
    P = np.ascontiguousarray(V, np.float32).copy()
    T = np.ascontiguousarray(F, np.uint32)
    L.orc_mesh_smooth(_ptr(P, C.c_float), len(P), _ptr(T, C.c_uint32), len(T))
    return P


def oracle_mesh_subdivide(V, F):
    L = oracle_lib()
    nv, nt = len(V), len(F)
    P = np.zeros((nv + 3 * nt, 3), np.float32)
    P[:nv] = V
    T = np.zeros((4 * nt, 3), np.uint32)
    T[:nt] = F
    n = L.orc_mesh_subdivide(_ptr(P, C.c_float), nv, _ptr(T, C.c_uint32), nt)
    return P[:n], T
