"""Host-side mirror of the reference's projection interface over libmlod.so.

Same class and method names, argument meaning and defaults as the reference's
``Kernel`` (sources/apss/Kernel.h:24-60), ``APSSOctree`` (sources/apss/CPUUtils.h:166-229)
and the ``APSS`` plugin interface with its ``Fastapss`` / ``GlobalAPSS`` implementations
(sources/apss/GPUUtils.h:22-67, Fastapss.h:32-51, GlobalAPSS.h:19-43), with numpy
(n, 3) float32 arrays in place of ``glm::vec3*``.  Differences:

* the octree is built on the GPU by ``APSSOctree.build`` (the reference builds it on the
  CPU and flattens it in ``copyToGPU``);
* errors raise ``MlodError`` instead of ``exit()`` (gpuErrchk, GPUUtils.cu:32-41);
* ``projectCPU`` is not provided: there is no CPU path in the product;
* ``set_semantics`` / ``set_precision`` choose the reference function (REF_GPU = Fastapss,
  REF_CPU = APSSOctree::projectCPU) and the arithmetic form.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib

__all__ = ["Kernel", "APSSOctree", "APSS", "Fastapss", "GlobalAPSS", "MlodError"]

MlodError = _lib.MlodError


def _f32(a, name):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.ndim != 2 or a.shape[1] != 3:
        raise ValueError(f"{name} must have shape (n, 3)")
    return a


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Kernel:
    """Weight kernel (Kernel.h:24-60).  Types: GAUSSIAN = 0, SINGULAR = 1 (rational),
    GAUSSIAN_MULTIPLE = 2 (Kernel.h:37).  Defaults as Kernel::Kernel (Kernel.cpp:13-16)."""

    GAUSSIAN, SINGULAR, GAUSSIAN_MULTIPLE = 0, 1, 2

    def __init__(self, type=0, sigma=2.0, exponent=1.0, constant_shift=0.0, sigmas=None, beginningSigma=0.0,
                 factor=0.0, nbOfComputedSigmas=1, fixedSigmas=None):
        self.type = int(type)
        self.sigma = float(sigma)
        self.exponent = float(exponent)
        self.constant_shift = float(constant_shift)
        self.sigmas = [float(s) for s in (sigmas or [])]
        self.beginningSigma = float(beginningSigma)
        self.factor = float(factor)
        self.nbOfComputedSigmas = int(nbOfComputedSigmas)
        self.fixedSigmas = [float(s) for s in (fixedSigmas or [])]

    def copy(self):
        return Kernel(self.type, self.sigma, self.exponent, self.constant_shift, list(self.sigmas), self.beginningSigma,
                      self.factor, self.nbOfComputedSigmas, list(self.fixedSigmas))

    def updateSigmas(self, computedSigmas=True, radius=1.0):
        """Kernel::updateSigmas (Kernel.cpp:110-128): geometric sigmas sigma_i = beginningSigma
        * factor^i (or the fixed list), then sigma and every sigma_i scaled by ``radius``.
        Arithmetic in float32 as the reference."""
        f32 = np.float32
        if computedSigmas:
            s = [f32(self.beginningSigma)]
            for _ in range(1, self.nbOfComputedSigmas):
                s.append(f32(s[-1] * f32(self.factor)))
        else:
            s = [f32(v) for v in self.fixedSigmas]
        self.sigma = float(f32(self.sigma) * f32(radius))
        self.sigmas = [float(f32(v) * f32(radius)) for v in s]
        return self

    def scaled(self, radius):
        """scene::setNewKernel (scene.cpp:1025-1033): sigma and sigmas times the octree bbox
        radius (constant_shift unscaled)."""
        k = self.copy()
        f32 = np.float32
        k.sigma = float(f32(k.sigma) * f32(radius))
        k.sigmas = [float(f32(v) * f32(radius)) for v in k.sigmas]
        return k

    def to_c(self):
        k = _lib.KernelC()
        k.type = self.type
        k.sigma = self.sigma
        k.exponent = self.exponent
        k.constant_shift = self.constant_shift
        k.n_sigmas = len(self.sigmas)
        if len(self.sigmas) > _lib.MAX_SIGMAS:
            raise ValueError("at most 32 sigmas")
        for i, s in enumerate(self.sigmas):
            k.sigmas[i] = s
        return k


class APSSOctree:
    """The octree the projection walks (APSSOctree, CPUUtils.h:166-229), built on the GPU.

    ``build(points, normals, updateRootBoundingBox=True, maxDepth=12,
    maxNumberOfPointsPerLeaf=1)`` as CPUUtils.h:176-180; ``tie_policy`` fixes the child of
    exact-midpoint ties (0 low, 1 high) where the reference draws rand()."""

    def __init__(self):
        self._h = None
        self._n = 0
        self._root_box = None
        self.m_scalingProtectionSphere = 2.0
        self.m_minDepth = 0
        self.m_knnMode = False
        self.m_knn = 20
        self.m_ballMode = False
        self.m_ballRadius = 0.1

    def setRootBoundingBox(self, bb, BB):
        """The root box a build with updateRootBoundingBox=False keeps (the reference's root node
        box, octreeNode::setBoundingBox, CPUUtils.h:118)."""
        self._root_box = np.concatenate([np.asarray(bb, np.float32).reshape(3), np.asarray(BB, np.float32).reshape(3)])

    def _opts(self, updateRootBoundingBox, maxDepth, maxNumberOfPointsPerLeaf, tie_policy):
        opts = _lib.BuildOptsC(int(maxDepth), int(maxNumberOfPointsPerLeaf), int(tie_policy))
        if not updateRootBoundingBox:
            box = getattr(self, "_root_box", None)
            if box is None:
                raise _lib.MlodError(_lib.ERR_STATE, "updateRootBoundingBox=False needs a root box "
                                                     "(setRootBoundingBox or a previous build)")
            opts.use_root_box = 1
            for a in range(6):
                opts.root_box[a] = float(box[a])
        return opts

    def _built(self, h, n):
        self._h = h
        self._n = int(n)
        self._root_box = np.asarray(self.info()["bbox"], np.float32)

    def build(self, points, normals, updateRootBoundingBox=True, maxDepth=12, maxNumberOfPointsPerLeaf=1, tie_policy=0,
              stream=None):
        """APSSOctree::build (CPUUtils.h:176-180, CPUUtils.cpp:540-569).  With
        updateRootBoundingBox=False the root keeps its box: the one of this octree's previous
        build, or the one given to setRootBoundingBox."""
        p = _f32(points, "points")
        n = _f32(normals, "normals")
        if len(p) != len(n):
            raise ValueError("points and normals differ in length")
        opts = self._opts(updateRootBoundingBox, maxDepth, maxNumberOfPointsPerLeaf, tie_policy)
        self.free()
        h = C.c_void_p()
        check(lib().mlod_tree_create(_fp(p), _fp(n), len(p), C.byref(opts), stream, C.byref(h)))
        self._built(h, len(p))
        return self

    def build_device(self, d_points, d_normals, n, maxDepth=12, tie_policy=0, stream=None, updateRootBoundingBox=True,
                     maxNumberOfPointsPerLeaf=1):
        """Builds from device pointers (n x 3 float32 each)."""
        opts = self._opts(updateRootBoundingBox, maxDepth, maxNumberOfPointsPerLeaf, tie_policy)
        self.free()
        h = C.c_void_p()
        check(lib().mlod_tree_create_device(C.c_void_p(d_points), C.c_void_p(d_normals), int(n), C.byref(opts), stream,
                                            C.byref(h)))
        self._built(h, n)
        return self

    @property
    def handle(self):
        if self._h is None:
            raise _lib.MlodError(-5, "octree not built")
        return self._h

    def free(self):
        if self._h is not None:
            lib().mlod_tree_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def info(self):
        inf = _lib.TreeInfoC()
        check(lib().mlod_tree_get_info(self.handle, C.byref(inf)))
        return dict(n_points=inf.n_points, n_nodes=inf.n_nodes, depth=inf.depth, level_offset=list(inf.level_offset),
                    bbox=list(inf.bbox), radius=inf.radius, ties=inf.ties, device_bytes=inf.device_bytes)

    def getBoundingBox(self):
        b = self.info()["bbox"]
        return np.array(b[:3], np.float32), np.array(b[3:], np.float32)

    def radius(self):
        return self.info()["radius"]

    def getNbOfLeaves(self):
        return self._n

    def nbOfNodes(self):
        return self.info()["n_nodes"]

    def setScalingProtectionSphere(self, v):
        self.m_scalingProtectionSphere = float(v)

    def getScalingProtectionSphere(self):
        return self.m_scalingProtectionSphere

    def setMinDepth(self, v):
        self.m_minDepth = int(v)

    def getMinDepth(self):
        return self.m_minDepth

    def knnMode(self, setMode, knn=20):
        """APSSOctree::knnMode (CPUUtils.cpp:674-680)."""
        self.m_knnMode = bool(setMode)
        if self.m_knnMode:
            self.m_ballMode = False
        self.m_knn = int(knn)

    def ballMode(self, setMode, ballRadius=0.1):
        """APSSOctree::ballMode (CPUUtils.cpp:682-688); a negative radius means 3 sigmas[0]."""
        self.m_ballMode = bool(setMode)
        if self.m_ballMode:
            self.m_knnMode = False
        self.m_ballRadius = float(ballRadius)

    def isAPSS(self):
        return self.m_knnMode or self.m_ballMode

    def dump(self):
        """BFS-ordered node tables (the flattening of Fastapss copy(), Fastapss.cu:97-150)."""
        inf = self.info()
        nn, n = inf["n_nodes"], self._n
        d = dict(depth=np.zeros(nn, np.int32), bb=np.zeros((nn, 6), np.float32), center=np.zeros((nn, 3), np.float32),
                 radius=np.zeros(nn, np.float32), father=np.zeros(nn, np.int32), first_child=np.zeros(nn, np.int32),
                 next_brother=np.zeros(nn, np.int32), num_children=np.zeros(nn, np.int32),
                 leaf_count=np.zeros(nn, np.uint32), leaf_idx=np.zeros(n, np.uint32), areas=np.zeros(n, np.float32),
                 moments=np.zeros((nn, 9), np.float32))
        P = lambda a, t: a.ctypes.data_as(C.POINTER(t))  # noqa: E731
        check(lib().mlod_tree_dump(self.handle, P(d["depth"], C.c_int32), P(d["bb"], C.c_float),
                                   P(d["center"], C.c_float), P(d["radius"], C.c_float), P(d["father"], C.c_int32),
                                   P(d["first_child"], C.c_int32), P(d["next_brother"], C.c_int32),
                                   P(d["num_children"], C.c_int32), P(d["leaf_count"], C.c_uint32),
                                   P(d["leaf_idx"], C.c_uint32), P(d["areas"], C.c_float), P(d["moments"], C.c_float)))
        return d


class APSS:
    """The projection plugin interface (GPUUtils.h:22-67) on the GPU octree."""

    MODE = _lib.MODE_MLOD

    def __init__(self, apssOctree: APSSOctree, kernel: Kernel | None = None, nbOfVectors: int = 0):
        self.m_apssOctree = apssOctree
        self.m_scalingProtectionSphere = apssOctree.m_scalingProtectionSphere
        self.m_minDepth = apssOctree.m_minDepth
        self.semantics = _lib.SEM_REF_GPU
        self.precision = _lib.PREC_FP32
        self.traversal = 0
        self._kernel = None
        self._m = 0
        if kernel is not None:
            self.copyToGPU(kernel, nbOfVectors)

    # -- configuration ------------------------------------------------------------------
    def copyToGPU(self, kernel: Kernel, nbOfVectors: int = 0):
        """Fastapss::copyToGPU (Fastapss.cu:457-484): the tree is already resident; installs the kernel."""
        self.updateKernel(kernel)

    def eraseFromGPU(self):
        self.m_apssOctree.free()

    def updateKernel(self, kernel: Kernel):
        """APSS::updateKernel (GPUUtils.cu:466-472)."""
        self._kernel = kernel.copy()
        check(lib().mlod_set_kernel(self.m_apssOctree.handle, C.byref(kernel.to_c())))

    def setScalingProtectionSphere(self, val):
        self.m_scalingProtectionSphere = float(val)
        self.m_apssOctree.setScalingProtectionSphere(val)

    def getScalingProtectionSphere(self):
        return self.m_scalingProtectionSphere

    def setMinDepth(self, val):
        self.m_minDepth = int(val)
        self.m_apssOctree.setMinDepth(val)

    def getMinDepth(self):
        return self.m_minDepth

    def set_semantics(self, semantics):
        """'ref_gpu' (Fastapss.cu) or 'ref_cpu' (APSSOctree::projectCPU)."""
        self.semantics = {"ref_gpu": 0, "ref_cpu": 1}.get(semantics, semantics)

    def set_precision(self, precision):
        """'fp32' (plain), 'fp32c' (compensated, FP64 fit) or 'fp64'."""
        self.precision = {"fp32": 0, "fp32c": 1, "fp64": 2}.get(precision, precision)

    def set_traversal(self, traversal):
        """'lane' (one walk per query, the default) or 'wave' (one walk per wavefront over the
        union of its queries' walks); no counterpart in the reference."""
        self.traversal = {"lane": 0, "wave": 1}.get(traversal, traversal)

    def _mode(self):
        t = self.m_apssOctree
        if self.MODE == _lib.MODE_MLOD and t.m_knnMode:
            return _lib.MODE_KNN
        if self.MODE == _lib.MODE_MLOD and t.m_ballMode:
            return _lib.MODE_BALL
        return self.MODE

    def _sync_params(self):
        h = self.m_apssOctree.handle
        check(lib().mlod_set_params(h, self.m_scalingProtectionSphere, self.m_minDepth, int(self.semantics),
                                    int(self.precision)))
        t = self.m_apssOctree
        check(lib().mlod_set_mode(h, self._mode(), t.m_knn, t.m_ballRadius))
        check(lib().mlod_set_traversal(h, int(self.traversal)))

    # -- queries --------------------------------------------------------------------------
    def copyPointsToGPU(self, nbOfVectors, outputPoints, stream=None):
        """APSS::copyPointsToGPU (GPUUtils.cu:406-411)."""
        q = _f32(outputPoints, "points")[: int(nbOfVectors)]
        check(lib().mlod_queries_upload(self.m_apssOctree.handle, _fp(q), len(q), stream))
        self._m = len(q)

    def copyPointsToGPU_device(self, d_points, nbOfVectors, stream=None):
        check(lib().mlod_queries_set_device(self.m_apssOctree.handle, C.c_void_p(d_points), int(nbOfVectors), stream))
        self._m = int(nbOfVectors)

    def erasePointsFromGPU(self):
        if self.m_apssOctree._h is not None:
            check(lib().mlod_queries_erase(self.m_apssOctree.handle))
        self._m = 0

    def project(self, nbOfVectors, outputPoints=None, outputNormals=None, n_iterations=1, kernel=None, stepMaxSize=-1.0,
                stream=None):
        """APSS::project (GPUUtils.h:26-27): n_iterations steps in place on the uploaded
        queries; the ``kernel`` argument is ignored as in Fastapss::project (the installed
        kernel is used).  Fills outputPoints / outputNormals ((n, 3) float32) when given and
        returns them."""
        n = int(nbOfVectors)
        self._sync_params()
        op = on = None
        if outputPoints is not None:
            op = outputPoints if (isinstance(outputPoints, np.ndarray) and outputPoints.dtype == np.float32 and
                                  outputPoints.flags.c_contiguous) else np.zeros((n, 3), np.float32)
        if outputNormals is not None:
            on = outputNormals if (isinstance(outputNormals, np.ndarray) and outputNormals.dtype == np.float32 and
                                   outputNormals.flags.c_contiguous) else np.zeros((n, 3), np.float32)
        check(lib().mlod_project(self.m_apssOctree.handle, n, int(n_iterations), float(stepMaxSize),
                                 _fp(op) if op is not None else None, _fp(on) if on is not None else None, stream))
        return op, on

    def project_points(self, points, n_iterations=1, stepMaxSize=-1.0, stream=None):
        """Projects host points (m, 3) with the current settings without touching the uploaded
        queries; returns (points, normals)."""
        q = _f32(points, "points")
        self._sync_params()
        op = np.zeros((len(q), 3), np.float32)
        on = np.zeros((len(q), 3), np.float32)
        check(lib().mlod_project_points(self.m_apssOctree.handle, _fp(q), len(q), int(n_iterations), float(stepMaxSize),
                                        _fp(op), _fp(on), stream))
        return op, on

    def projectCPU(self, qStart, n_iterations=1, kernel=None, stepMaxSize=-1.0, stream=None):
        """APSS::projectCPU (GPUUtils.h:33-34): the reference's CPU function (REF_CPU semantics,
        APSSOctree::projectCPU, CPUUtils.cpp:621-636), computed on the GPU for one point (3,) or
        a batch (m, 3); returns (points, normals) of the same shape."""
        q = np.asarray(qStart, np.float32)
        single = q.ndim == 1
        sem = self.semantics
        self.semantics = _lib.SEM_REF_CPU
        try:
            op, on = self.project_points(q.reshape(-1, 3), n_iterations, stepMaxSize, stream)
        finally:
            self.semantics = sem
        return (op[0], on[0]) if single else (op, on)

    def reorganizePoints(self, invalidPts, endValid, stream=None):
        """APSS::reorganizePoints (GPUUtils.cu:436-450): uploaded queries invalidPts[i] and
        endValid - 1 - i exchange positions, in order."""
        inv = np.ascontiguousarray(invalidPts, np.uint32)
        check(lib().mlod_queries_reorganize(self.m_apssOctree.handle, inv.ctypes.data_as(C.POINTER(C.c_uint32)),
                                            len(inv), int(endValid), stream))

    def fit_kinds(self, n=None, stream=None):
        """Branch of the last step's sphere fit per query: 0 sphere, 1 plane, 2 invalid."""
        n = self._m if n is None else int(n)
        out = np.zeros(n, np.uint8)
        check(lib().mlod_queries_fit_kinds(self.m_apssOctree.handle, n, out.ctypes.data_as(C.c_void_p), stream))
        return out

    def download(self, n=None, stream=None):
        n = self._m if n is None else int(n)
        op = np.zeros((n, 3), np.float32)
        on = np.zeros((n, 3), np.float32)
        check(lib().mlod_queries_download(self.m_apssOctree.handle, n, _fp(op), _fp(on), stream))
        return op, on

    def count_visits(self, n=None, stream=None):
        """Traversal counters of one step: far nodes, leaf points, children tested, near nodes'
        own shares, lane loop iterations, and the lane slots they were issued in (64 x wavefront iterations)."""
        n = self._m if n is None else int(n)
        self._sync_params()
        out = (C.c_uint64 * 6)()
        check(lib().mlod_count_visits(self.m_apssOctree.handle, n, out, stream))
        return dict(far=out[0], leaf_points=out[1], tested=out[2], near=out[3], iterations=out[4], lane_slots=out[5])


class Fastapss(APSS):
    """MLoD projection (Fastapss, Fastapss.h:32-51)."""

    MODE = _lib.MODE_MLOD


class GlobalAPSS(APSS):
    """Brute-force APSS over all points (GlobalAPSS, GlobalAPSS.h:19-43)."""

    MODE = _lib.MODE_GLOBAL


class OctreeGenerator:
    """The adaptive-octree driver (OctreeGenerator, upsample.h:26-52) on the GPU.

    Cells are returned as (n, 4) int32 rows (depth, x, y, z) of the [-1, 1]^3 octree in the
    reference's node_comparator order.  ``generate`` projects with ``apss``'s tree and settings
    (its installed kernel, cutoff, min depth, semantics, form and neighbourhood mode); the
    ``kernel`` argument is kept for the reference's signature, which passes it to a project call
    that ignores it (Fastapss.cu:439-455)."""

    def __init__(self, apss: APSS | None = None, kernel: Kernel | None = None):
        self.apss = apss
        self.kernel = kernel
        h = C.c_void_p()
        check(lib().mlod_generator_create(apss.m_apssOctree.handle if apss is not None else None, C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().mlod_generator_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def generate(self, max_depth, total_nodes: list | None = None, stream=None):
        """OctreeGenerator::generate (upsample.cpp:228-251).  When ``total_nodes`` is a list it
        receives the crossed cells (total_tree) as one (m, 4) array."""
        if self.apss is None:
            raise MlodError(_lib.ERR_STATE, "generate needs an APSS")
        self.apss._sync_params()
        check(lib().mlod_generate(self._h, int(max_depth), int(total_nodes is not None), stream))
        if total_nodes is not None:
            total_nodes.append(self.cells(1, stream))
        return self.cells(0, stream)

    def generateFromPoints(self, max_depth, pts, box=None, stream=None):
        """OctreeGenerator::generateFromPoints (upsample.cpp:128-144); box = (min xyz, max xyz)."""
        p = _f32(pts, "points")
        b = None if box is None else np.ascontiguousarray(np.asarray(box, np.float32).reshape(6))
        check(lib().mlod_generate_from_points(self._h, _fp(p), len(p), 0, int(max_depth),
                                              _fp(b) if b is not None else None, stream))
        return self.cells(0, stream)

    def info(self):
        o = (C.c_uint64 * 6)()
        check(lib().mlod_generator_info(self._h, o))
        return dict(leaves=o[0], total=o[1], max_rel_depth=o[2], projected=o[3], calls=o[4], nonfinite=o[5])

    def cells(self, which=0, stream=None):
        n = self.info()["total" if which else "leaves"]
        out = np.zeros((n, 4), np.int32)
        if n:
            check(lib().mlod_generator_cells(self._h, int(which), out.ctypes.data_as(C.c_void_p), stream))
        return out

    def centres(self, stream=None):
        """Centres of the generated cells (aabb_octree::center), (n, 3) float32."""
        n = self.info()["leaves"]
        out = np.zeros((n, 3), np.float32)
        d = C.c_void_p()
        check(lib().mlod_generator_centres(self._h, _fp(out), C.byref(d), stream))
        return out

    def centres_device(self, stream=None):
        """Device address of the centres (n x 3 float32), valid until the next call."""
        d = C.c_void_p()
        check(lib().mlod_generator_centres(self._h, None, C.byref(d), stream))
        return d.value
