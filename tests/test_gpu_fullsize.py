"""The headline workload at its full size (8M points on the displaced superquadric, 8M offset
queries, multi-Gaussian S = 6, cutoff 2), checked through properties that do not depend on the
size: the octree holds every point once, its root moments are the FP64 sums over the points, the
compensated form stays within the 1e-6 x diag gate of the FP64 form on every query after three
steps, and a sample projected on its own gives the same bits as inside the full batch."""
import numpy as np
import pytest

from mlodsurfaces_amd import apss as A
from mlodsurfaces_amd import datasets as ds

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def workload():
    return ds.make_workload("superquadric8m")


def test_full_size_tree(workload):
    P, N = workload["points"], workload["normals"]
    t = A.APSSOctree().build(P, N)
    d = t.dump()
    n = len(P)
    assert t.info()["n_points"] == n and d["depth"].max() <= 12
    assert int(d["leaf_count"].sum()) == n
    seen = np.zeros(n, np.int8)
    seen[d["leaf_idx"]] += 1
    assert np.all(seen == 1)  # every point in exactly one leaf
    a = d["areas"].astype(np.float64)
    p = P.astype(np.float64)
    nn = N.astype(np.float64)
    root = d["moments"][0].astype(np.float64)
    expect = np.concatenate([[a.sum()], (a[:, None] * p).sum(0), (a[:, None] * nn).sum(0),
                             [(a * (p * nn).sum(1)).sum()], [(a * (p * p).sum(1)).sum()]])
    scale = np.concatenate([[a.sum()], np.full(3, (a[:, None] * np.abs(p)).sum()), np.full(3, a.sum()),
                            [(a * np.abs((p * nn).sum(1))).sum()], [(a * (p * p).sum(1)).sum()]])
    # FP32 sums of 8M terms, bottom-up over the tree: a few 1e-6 of the sum of magnitudes
    assert np.all(np.abs(root - expect) <= 2e-5 * scale), (root, expect)


def test_full_size_projection_gate_and_batch_invariance(workload):
    P, N, Q = workload["points"], workload["normals"], workload["queries"]
    t = A.APSSOctree().build(P, N)
    K = A.Kernel(**workload["kernel"])
    diag = float(np.linalg.norm(P.max(0).astype(np.float64) - P.min(0)))
    out = {}
    for prec in ("fp32c", "fp64"):
        ap = A.Fastapss(t, K)
        ap.setScalingProtectionSphere(workload["cutoff"])
        ap.set_precision(prec)
        ap.copyPointsToGPU(len(Q), Q)
        out[prec] = ap.project(len(Q), np.zeros_like(Q), np.zeros_like(Q), 3, K)
    err = np.linalg.norm(out["fp32c"][0].astype(np.float64) - out["fp64"][0], axis=1) / diag
    print(f"8M queries, 3 steps: FP32C vs FP64 max {err.max():.3e} x diag")
    assert np.isfinite(err).all() and err.max() < 1e-6
    # a sample of the queries projected on its own: the same bits as in the full batch
    ap = A.Fastapss(t, K)
    ap.setScalingProtectionSphere(workload["cutoff"])
    ap.set_precision("fp32c")
    idx = np.arange(0, len(Q), 997)
    sp, sn = ap.project_points(Q[idx], 3)
    assert np.array_equal(sp, out["fp32c"][0][idx]) and np.array_equal(sn, out["fp32c"][1][idx])


@pytest.mark.timeout(600)
def test_full_size_headline_vs_fp64_oracle(workload, oracle_built):
    """The headline configuration as the bench runs it (REF_GPU semantics, the compensated FP32C
    form) against the FP64 oracle's REF_GPU function on the same 8M-point octree (same tie
    policy): 2000 sampled queries, 3 steps, every one within 1e-6 x diag.  The plain FP32 form on
    the same sample is reported beside it."""
    P, N, Q = workload["points"], workload["normals"], workload["queries"]
    t = A.APSSOctree().build(P, N)
    K = A.Kernel(**workload["kernel"])
    diag = float(np.linalg.norm(P.max(0).astype(np.float64) - P.min(0)))
    idx = np.sort(np.random.default_rng(17).choice(len(Q), 2000, replace=False))
    got = {}
    for prec in ("fp32c", "fp32"):
        ap = A.Fastapss(t, K)
        ap.setScalingProtectionSphere(workload["cutoff"])
        ap.set_semantics("ref_gpu")
        ap.set_precision(prec)
        got[prec], _ = ap.project_points(Q[idx], 3)
    o = oracle_built.OracleTree(P, N, tie_policy=0)
    k = oracle_built.make_kernel(type=K.type, sigma=K.sigma, exponent=K.exponent, constant_shift=K.constant_shift,
                                 sigmas=K.sigmas)
    par = oracle_built.make_params(cutoff=workload["cutoff"], min_depth=0, semantics="ref_gpu")
    op, _ = o.project(Q[idx], k, par, n_iter=3, precision=1, nthreads=16)
    err = {f: np.linalg.norm(g.astype(np.float64) - op, axis=1) / diag for f, g in got.items()}
    print(f"8M headline, 2000 queries x 3 steps vs the FP64 oracle (REF_GPU): FP32C max {err['fp32c'].max():.3e} "
          f"RMS {np.sqrt(np.mean(err['fp32c'] ** 2)):.3e}; plain FP32 max {err['fp32'].max():.3e} x diag")
    assert np.isfinite(err["fp32c"]).all() and err["fp32c"].max() < 1e-6


def test_full_size_surface_mesh_is_closed(workload):
    """The surface mesh of the full cloud at depth 8 (the bench's mesh): the superquadric is closed,
    so dual contouring of its projected corner field gives a closed, consistently oriented mesh of
    genus 0 (every directed edge once, its reverse once, V - E + F = 2), and the refined vertices lie
    on the cloud."""
    from mlodsurfaces_amd import dc
    p = ds.pn_normalize(workload["points"])
    t = A.APSSOctree().build(p, workload["normals"])
    ap = A.Fastapss(t, A.Kernel(**workload["kernel"]))
    ap.setScalingProtectionSphere(workload["cutoff"])
    ap.set_precision("fp32c")
    V, N, F, info = dc.surface_mesh(ap, 8, refine=0)
    assert len(F) > 100000
    e = np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]]).astype(np.int64)
    key = e[:, 0] * len(V) + e[:, 1]
    rev = e[:, 1] * len(V) + e[:, 0]
    assert len(np.unique(key)) == len(key)  # no directed edge twice
    assert np.isin(rev, key).all()  # every edge has its opposite
    assert len(V) - len(key) // 2 + len(F) == 2
    from scipy.spatial import cKDTree
    d, _ = cKDTree(p[::4]).query(V)
    print(f"depth-8 mesh: {len(V)} vertices, {len(F)} triangles, vertex-to-cloud median {np.median(d):.2e}")
    assert np.median(d) < 2e-3


@pytest.mark.parametrize("name", ["sphere100k", "torus1m"])
def test_full_size_small_configs(oracle_built, name):
    """The 100k-point sphere (multi-Gaussian S = 4, cutoff 3) and the 1M-point torus (rational kernel
    e = 4, cutoff 3) at full size with the reference CPU function (REF_CPU): the compensated form
    within 1e-6 x diag of the FP64 form on every query after 3 steps, and of the FP64 oracle (same
    tie policy) on a sample of 2000 queries."""
    w = ds.make_workload(name)
    P, N, Q = w["points"], w["normals"], w["queries"]
    t = A.APSSOctree().build(P, N)
    K = A.Kernel(**w["kernel"])
    diag = float(np.linalg.norm(P.max(0).astype(np.float64) - P.min(0)))
    out = {}
    for prec in ("fp32c", "fp64"):
        ap = A.Fastapss(t, K)
        ap.setScalingProtectionSphere(w["cutoff"])
        ap.set_semantics("ref_cpu")
        ap.set_precision(prec)
        ap.copyPointsToGPU(len(Q), Q)
        out[prec] = ap.project(len(Q), np.zeros_like(Q), np.zeros_like(Q), 3, K)
    err = np.linalg.norm(out["fp32c"][0].astype(np.float64) - out["fp64"][0], axis=1) / diag
    idx = np.random.default_rng(7).choice(len(Q), 2000, replace=False)
    o = oracle_built.OracleTree(P, N)
    k = oracle_built.make_kernel(type=K.type, sigma=K.sigma, exponent=K.exponent, constant_shift=K.constant_shift,
                                 sigmas=K.sigmas)
    par = oracle_built.make_params(cutoff=w["cutoff"], min_depth=0, semantics="ref_cpu")
    op, _ = o.project(Q[idx], k, par, n_iter=3, precision=1, nthreads=16)
    oerr = np.linalg.norm(out["fp32c"][0][idx].astype(np.float64) - op, axis=1) / diag
    print(f"{name}: FP32C vs FP64 max {err.max():.3e}, vs the FP64 oracle (2000 queries) max {oerr.max():.3e} x diag")
    assert err.max() < 1e-6 and oerr.max() < 1e-6


@pytest.mark.timeout(600)
def test_full_size_c5_neighbourhoods(oracle_built):
    """kNN (k = 32) and ball (r = 3 sigma_0) APSS at the C5 size (20M points in the reference's
    normalised frame, queries from the depth-11 dual-contouring grid): for a sample of queries, one
    step on the GPU equals the reference's fit (the oracle's FP64 projectOnAPSSStats) over the
    neighbourhood an exact k-d tree finds, with unit-area weights of the reference's
    multi-Gaussian kernel at the raw points (CPUUtils.cpp:690-737): the FP64 form to the float32
    rounding of its output (~1e-8 x diag; one neighbour more or less moves a fit by ~1e-5 x diag),
    the compensated FP32C form (the configuration's arithmetic) within 1e-6 x diag."""
    from scipy.spatial import cKDTree
    from mlodsurfaces_amd import dc
    w = ds.make_workload("c5_knn")
    P, N = w["points"], w["normals"]
    t = A.APSSOctree().build(P, N)
    K = A.Kernel(**w["kernel"])
    ap = A.Fastapss(t, K)
    ap.setScalingProtectionSphere(w["cutoff"])
    Q = dc.grid_centres(ap, 11)
    print(f"C5 DC grid at depth 11: {len(Q)} cell centres", flush=True)
    assert len(Q) > 8_000_000
    q = Q[np.random.default_rng(11).choice(len(Q), 150, replace=False)]
    P64, N64 = P.astype(np.float64), N.astype(np.float64)
    kd = cKDTree(P64)
    diag = float(np.linalg.norm(P64.max(0) - P64.min(0)))
    sig = np.asarray(K.sigmas, np.float32).astype(np.float64)  # the kernel as installed (float sigmas)
    r = 3.0 * sig[0]
    R = float(t.radius())

    def fit(qi, j):
        # sums in the query's frame (the neighbourhood spans ~1e-3 of the cloud, so global-frame
        # sums would lose five digits to cancellation before the fit)
        p, n = P64[j] - qi, N64[j]
        d2 = (p * p).sum(1)
        wgt = (np.exp(-d2[:, None] / (2 * sig[None] ** 2)) / sig[None] ** 3).sum(1)
        s = np.concatenate([[wgt.sum()], wgt @ p, wgt @ n, [wgt @ (p * n).sum(1)], [wgt @ d2]])
        return oracle_built.oracle_fit(s, np.zeros(3), R, ref_gpu=False)[0] + qi

    for mode in ("knn", "ball"):
        if mode == "knn":
            t.knnMode(True, 32)
        else:
            t.ballMode(True, -1.0)
        got = {}
        for prec in ("fp64", "fp32c"):
            g = A.Fastapss(t, K)
            g.setScalingProtectionSphere(w["cutoff"])
            g.set_semantics("ref_cpu")
            g.set_precision(prec)
            got[prec], _ = g.project_points(q, 1)
        errs, errc = [], []
        for i, qi in enumerate(q.astype(np.float64)):
            if mode == "knn":
                dist, j = kd.query(qi, 33)
                if dist[32] - dist[31] < 1e-9:  # a tie at the k-th neighbour: either set is right
                    continue
                j = j[:32]
            else:
                j = np.asarray(kd.query_ball_point(qi, r), np.int64)
                if len(j) < 10:  # a fit on a few points is decided by rounding
                    continue
            ref = fit(qi, j)
            errs.append(np.linalg.norm(got["fp64"][i] - ref) / diag)
            errc.append(np.linalg.norm(got["fp32c"][i] - ref) / diag)
        errs, errc = np.array(errs), np.array(errc)
        print(f"C5 {mode}: {len(errs)} queries, GPU vs the exact neighbourhood's fit: FP64 median {np.median(errs):.1e} "
              f"max {errs.max():.1e}, FP32C median {np.median(errc):.1e} max {errc.max():.1e} x diag", flush=True)
        assert len(errs) > 100 and errs.max() < 1e-7 and errc.max() < 1e-6
