"""mlod_project_points_noalloc, the entry point that allocates nothing, called through the C ABI:
two calls in a row, the second with fewer points, each giving the bits of mlod_project_points
on the same points (MLoD FP32C REF_GPU, and kNN APSS); the uploaded queries are left as they
were; more points than MLOD_NOALLOC_MAX_QUERIES are refused."""
import ctypes as C

import numpy as np
import pytest

from mlodsurfaces_amd import _lib
from mlodsurfaces_amd import apss as A
from mlodsurfaces_amd import datasets as ds
from mlodsurfaces_amd import hiprt

pytestmark = pytest.mark.gpu

MAXQ = 262144


def _download(buf, shape):
    out = np.zeros(shape, np.float32)
    hiprt.check(hiprt.hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), buf.ptr, out.nbytes, 2), "hipMemcpy D2H")
    return out


def _noalloc(ap, q, n_iter):
    """The test allocates the device inputs and outputs; the entry point allocates nothing."""
    L = _lib.lib()
    ap._sync_params()
    dq = hiprt.DeviceBuffer(q)
    dp = hiprt.DeviceBuffer(np.full(q.shape, np.nan, np.float32))
    dn = hiprt.DeviceBuffer(np.full(q.shape, np.nan, np.float32))
    s = hiprt.Stream()
    r = L.mlod_project_points_noalloc(ap.m_apssOctree.handle, dq.ptr, len(q), n_iter, -1.0, dp.ptr, dn.ptr, s.ptr)
    assert r == 0, L.mlod_last_error()
    s.synchronize()
    return _download(dp, q.shape), _download(dn, q.shape)


@pytest.fixture(scope="module")
def cloud():
    p, n = ds.torus(40000, seed=2)
    return p, n, A.APSSOctree().build(p, n)


@pytest.mark.parametrize("mode", ["mlod", "knn"])
def test_noalloc_equals_project_points(cloud, mode):
    p, n, t = cloud
    R = t.radius()
    K = A.Kernel(type=2, sigmas=[0.01 * 2.0 ** i * R for i in range(4)])
    if mode == "knn":
        t.knnMode(True, 20)
    try:
        ap = A.Fastapss(t, K)
        ap.setScalingProtectionSphere(3.0)
        ap.set_semantics("ref_gpu")
        ap.set_precision("fp32c")
        q0 = ds.jitter_queries(p[:3000], 0.01, seed=5)
        ap.copyPointsToGPU(len(q0), q0)
        before = ap.download()
        rng = np.random.default_rng(9)
        for m in (30000, 7001):  # the second call reuses the scratch with fewer points
            q = ds.jitter_queries(p[rng.choice(len(p), m, replace=False)], 0.02, seed=m)
            gp, gn = _noalloc(ap, q, 3)
            ep, en = ap.project_points(q, 3)
            assert np.array_equal(gp, ep) and np.array_equal(gn, en), m
            assert np.isfinite(gp).all()
        after = ap.download()
        assert np.array_equal(before[0], after[0]) and np.array_equal(before[1], after[1])
    finally:
        if mode == "knn":
            t.knnMode(False)


def test_noalloc_rejects_too_many_points(cloud):
    p, n, t = cloud
    L = _lib.lib()
    ap = A.Fastapss(t, A.Kernel(type=0, sigma=0.05 * t.radius()))
    ap._sync_params()
    r = L.mlod_project_points_noalloc(ap.m_apssOctree.handle, C.c_void_p(16), MAXQ + 1, 1, -1.0, None, None, None)
    assert r == _lib.ERR_INVALID and b"MLOD_NOALLOC_MAX_QUERIES" in L.mlod_last_error()
    assert L.mlod_project_points_noalloc(ap.m_apssOctree.handle, None, 0, 1, -1.0, None, None, None) == 0
