"""The rest of the APSS plugin interface through the C ABI: projectCPU (the reference's CPU
function for points given directly, without touching the uploaded queries) and
reorganizePoints (exchange of uploaded query positions, GPUUtils.cu:413-450)."""
import numpy as np
import pytest

from mlodsurfaces_amd import apss as A
from mlodsurfaces_amd import datasets as ds

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup():
    p, n = ds.torus(30000, seed=71)
    t = A.APSSOctree().build(p, n)
    R = t.radius()
    K = A.Kernel(type=2, sigmas=[float(np.float32(0.01 * 2 ** i) * np.float32(R)) for i in range(4)])
    q = ds.jitter_queries(p[:4000], 0.01, seed=72)
    return p, n, t, K, q


def test_project_points_leaves_uploaded_queries(setup):
    p, n, t, K, q = setup
    ap = A.Fastapss(t, K)
    ap.set_precision("fp32c")
    ap.copyPointsToGPU(len(q), q)
    ap.project(len(q), None, None, 2, K)
    before = ap.download()
    other = q[::-1].copy()
    pp, pn = ap.project_points(other, 3)
    after = ap.download()
    assert np.array_equal(before[0], after[0]) and np.array_equal(before[1], after[1])
    ap.copyPointsToGPU(len(other), other)
    ep, en = ap.project(len(other), np.zeros_like(other), np.zeros_like(other), 3, K)
    assert np.array_equal(pp, ep) and np.array_equal(pn, en)


def test_project_cpu_is_ref_cpu_function(setup):
    p, n, t, K, q = setup
    ap = A.Fastapss(t, K)
    ap.set_semantics("ref_gpu")
    one_p, one_n = ap.projectCPU(q[7], 2, K)
    batch_p, batch_n = ap.projectCPU(q[:500], 2, K)
    ref = A.Fastapss(t, K)
    ref.set_semantics("ref_cpu")
    ref.copyPointsToGPU(500, q[:500])
    rp, rn = ref.project(500, np.zeros((500, 3), np.float32), np.zeros((500, 3), np.float32), 2, K)
    assert np.array_equal(batch_p, rp) and np.array_equal(batch_n, rn)
    assert np.array_equal(one_p, rp[7]) and np.array_equal(one_n, rn[7])
    assert ap.semantics == 0  # restored


def test_reorganize_points_sequential_exchange(setup):
    p, n, t, K, q = setup
    ap = A.Fastapss(t, K)
    m = 1000
    ap.copyPointsToGPU(m, q[:m])
    ap.project(m, None, None, 1, K)
    P0, N0 = ap.download()
    rng = np.random.default_rng(73)
    invalid = np.sort(rng.choice(900, 40, replace=False)).astype(np.uint32)
    invalid = np.concatenate([invalid, [905, 990]]).astype(np.uint32)  # some inside the tail: overlapping pairs
    end_valid = 1000
    ap.reorganizePoints(invalid, end_valid)
    P1, N1 = ap.download()
    EP = P0.copy()
    for i, a in enumerate(invalid):  # GPUUtils.cu:413-424, run in order
        b = end_valid - 1 - i
        EP[[a, b]] = EP[[b, a]]
    assert np.array_equal(P1, EP)
    assert np.array_equal(N1, N0)  # normals stay with their index
    # the next step continues from the exchanged positions
    ap.project(m, None, None, 1, K)
    P2, _ = ap.download()
    fresh = A.Fastapss(t, K)
    fresh.copyPointsToGPU(m, EP)
    FP, _ = fresh.project(m, np.zeros((m, 3), np.float32), None, 1, K)
    assert np.array_equal(P2, FP)
    with pytest.raises(A.MlodError):
        ap.reorganizePoints(np.array([m + 3], np.uint32), end_valid)


def test_fit_kinds_plane_and_sphere():
    """The fit-branch diagnostic: noise-free plane data takes the plane branch, a sphere the
    sphere branch (Fastapss.cu:354-395)."""
    rng = np.random.default_rng(74)
    xy = rng.uniform(-1, 1, (20000, 2)).astype(np.float32)
    P = np.column_stack([xy, np.zeros(len(xy), np.float32)]).astype(np.float32)
    N = np.tile(np.array([[0, 0, 1]], np.float32), (len(P), 1))
    S, SN = ds.sphere(20000, seed=75, noise=0.0)
    for pts, nrm, want in ((P, N, 1), (S, SN, 0)):
        t = A.APSSOctree().build(pts, nrm)
        R = t.radius()
        K = A.Kernel(type=2, sigmas=[float(np.float32(0.02 * 2 ** i) * np.float32(R)) for i in range(3)])
        ap = A.Fastapss(t, K)
        ap.set_precision("fp32c")
        q = pts[:500] * np.float32(0.8) + nrm[:500] * np.float32(0.01)
        ap.copyPointsToGPU(len(q), q)
        ap.project(len(q), None, None, 1, K)
        kinds = ap.fit_kinds()
        assert np.all(kinds == want), np.bincount(kinds)
        with pytest.raises(A.MlodError):
            ap.fit_kinds(len(q) + 1)


def test_count_visits_consistent(setup):
    """The untimed counter pass: every tested child and every near node made current is one loop
    iteration, the wavefront slots bound the iterations, and the pass leaves the queries as
    they were."""
    p, n, t, K, q = setup
    ap = A.Fastapss(t, K)
    ap.set_precision("fp32c")
    ap.copyPointsToGPU(len(q), q)
    before = ap.download()
    c = ap.count_visits()
    after = ap.download()
    assert np.array_equal(before[0], after[0])
    assert c["far"] > 0 and c["tested"] > 0 and c["near"] > 0
    assert c["iterations"] >= c["tested"] + c["near"]
    assert c["iterations"] <= c["lane_slots"] <= 64 * c["iterations"]
