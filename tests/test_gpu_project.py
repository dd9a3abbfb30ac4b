"""GPU MLoD / global APSS projection parity.

Tolerances (x bbox diagonal): the reference CPU path (golden vectors) 1e-5 for every form.
Against the FP64 oracle with the same semantics: FP32C 1e-6 (the precision form's target);
plain FP32 1e-5 (the global-frame FP32 sums cancel as in the reference's own FP32 path,
whose error is of the same order); FP64 1e-7 against the oracle rounded to float32 (the
output positions are float32, so one unit in the last place is the floor)."""
import numpy as np
import pytest

from conftest import MODE_NAMES, golden_kernel, golden_names, gpu_tree_from_fixture, load_golden, oracle_tree_opts
from mlodsurfaces_amd import datasets as ds

pytestmark = pytest.mark.gpu

TOL_REF = 1e-5
TOL_ORACLE = {"fp32": 1e-5, "fp32c": 1e-6, "fp64": 1e-7}

_trees = {}


def tree_for(name, p, n):
    from gpu_util import gpu_tree
    if name not in _trees:
        _trees[name] = gpu_tree(p, n)
    return _trees[name]


def golden_tree(name, t):
    """The octree of golden tree fixture `name`, built with the fixture's options."""
    key = "golden:" + name
    if key not in _trees:
        _trees[key] = gpu_tree_from_fixture(t)
    return _trees[key]


MLOD_GOLDEN = [n for n in golden_names("proj_") if "knn" not in n and "ball" not in n]


@pytest.mark.parametrize("precision", ["fp32", "fp32c", "fp64"])
@pytest.mark.parametrize("name", MLOD_GOLDEN)
def test_matches_reference_cpu_golden(oracle_built, name, precision):
    from gpu_util import gpu_project, rel_err
    g = load_golden(name)
    tn = str(g["tree"])
    t = load_golden(f"tree_{tn}.npz")
    print("case", name, precision, flush=True)
    tree = golden_tree(tn, t)
    mode = MODE_NAMES[int(g["mode"])]
    kern = golden_kernel(g)
    op, on, _ = gpu_project(tree, g["queries"], kern, cutoff=float(g["cutoff"]), min_depth=int(g["min_depth"]),
                            semantics="ref_cpu", precision=precision, n_iter=int(g["n_iter"]), mode=mode)
    diag = 2 * float(t["radius"][0])
    e = rel_err(op, g["out_points"], diag)
    assert e.max() < TOL_REF, e.max()
    assert np.all(np.abs(np.linalg.norm(on, axis=1) - 1) < 1e-4)
    # and the FP64 oracle with the same semantics
    o = oracle_built.OracleTree(t["points"], t["normals"], **oracle_tree_opts(t))
    prm = oracle_built.make_params(cutoff=float(g["cutoff"]), min_depth=int(g["min_depth"]), semantics="ref_cpu",
                                   mode=mode)
    o64, _ = o.project(g["queries"], oracle_built.make_kernel(**kern), prm, n_iter=int(g["n_iter"]), precision=1)
    e = rel_err(op, o64.astype(np.float32), diag)
    assert e.max() < TOL_ORACLE[precision], e.max()


KERNELS = {
    "mg4": lambda R: dict(type=2, sigmas=[0.01 * 2.0 ** i * R for i in range(4)]),
    "mg6": lambda R: dict(type=2, sigmas=[0.01 * 2.0 ** i * R for i in range(6)]),
    # the most sigmas a kernel holds (the sample models' .params files go up to 30)
    "mg32": lambda R: dict(type=2, sigmas=[0.004 * 1.2 ** i * R for i in range(32)]),
    "rational": lambda R: dict(type=1, exponent=4.0, constant_shift=1e-3),
    "rational_e2.5": lambda R: dict(type=1, exponent=2.5, constant_shift=2e-3),
    "gauss": lambda R: dict(type=0, sigma=0.03 * R, constant_shift=0.0),
}


@pytest.mark.parametrize("semantics", ["ref_gpu", "ref_cpu"])
@pytest.mark.parametrize("kname", list(KERNELS))
@pytest.mark.parametrize("precision", ["fp32", "fp32c", "fp64"])
def test_matches_fp64_oracle(oracle_built, semantics, kname, precision):
    from gpu_util import gpu_project, rel_err
    print("case", semantics, kname, precision, flush=True)
    p, n = ds.torus(40000, seed=31)
    tree = tree_for("torus40k", p, n)
    R = ds.bbox_radius(p)
    kern = KERNELS[kname](R)
    q = ds.jitter_queries(p[:3000], 0.01, seed=5)
    op, on, _ = gpu_project(tree, q, kern, cutoff=2.5, semantics=semantics, precision=precision, n_iter=2)
    o = oracle_built.OracleTree(p, n)
    o64, n64 = o.project(q, oracle_built.make_kernel(**kern), oracle_built.make_params(cutoff=2.5, semantics=semantics),
                         n_iter=2, precision=1)
    e = rel_err(op, o64.astype(np.float32), 2 * R)
    assert e.max() < TOL_ORACLE[precision], (e.max(), np.sqrt(np.mean(e ** 2)))
    assert np.all(np.sum(on * n64, 1) > 0.9999)


@pytest.mark.parametrize("max_pts,max_depth", [(4, 12), (8, 8), (32, 10)])
@pytest.mark.parametrize("semantics", ["ref_gpu", "ref_cpu"])
@pytest.mark.parametrize("precision", ["fp32", "fp32c", "fp64"])
def test_multi_point_leaves_match_fp64_oracle(oracle_built, max_pts, max_depth, semantics, precision):
    """Octrees with several points per leaf (maxNumberOfPointsPerLeaf, CPUUtils.cpp:279-283): the
    GPU function sums every point of a leaf (Fastapss.cu:189-200), the CPU one keeps the last
    (CPUUtils.cpp:448-457); both against the FP64 oracle with the same semantics and tree."""
    from gpu_util import gpu_project, gpu_tree, rel_err
    print("case", max_pts, max_depth, semantics, precision, flush=True)
    p, n = ds.torus(40000, seed=31)
    key = f"torus40k_m{max_pts}_d{max_depth}"
    if key not in _trees:
        _trees[key] = gpu_tree(p, n, maxDepth=max_depth, maxNumberOfPointsPerLeaf=max_pts)
    tree = _trees[key]
    R = ds.bbox_radius(p)
    kern = KERNELS["mg4"](R)
    q = ds.jitter_queries(p[:2000], 0.01, seed=7)
    op, on, _ = gpu_project(tree, q, kern, cutoff=2.5, semantics=semantics, precision=precision, n_iter=2)
    o = oracle_built.OracleTree(p, n, max_depth=max_depth, max_pts_per_leaf=max_pts)
    o64, n64 = o.project(q, oracle_built.make_kernel(**kern), oracle_built.make_params(cutoff=2.5, semantics=semantics),
                         n_iter=2, precision=1)
    e = rel_err(op, o64.astype(np.float32), 2 * R)
    assert e.max() < TOL_ORACLE[precision], (e.max(), np.sqrt(np.mean(e ** 2)))


@pytest.mark.parametrize("precision", ["fp32", "fp32c", "fp64"])
def test_global_apss_matches_oracle(oracle_built, precision):
    from gpu_util import gpu_project, rel_err
    p, n = ds.sphere(6000, seed=32)
    tree = tree_for("sphere6k", p, n)
    R = ds.bbox_radius(p)
    kern = KERNELS["mg4"](R)
    q = ds.jitter_queries(p[:700], 0.02, seed=6)
    op, _, _ = gpu_project(tree, q, kern, precision=precision, n_iter=2, mode="global")
    o = oracle_built.OracleTree(p, n)
    o64, _ = o.project(q, oracle_built.make_kernel(**kern), oracle_built.make_params(mode="global"), n_iter=2,
                       precision=1)
    assert rel_err(op, o64.astype(np.float32), 2 * R).max() < TOL_ORACLE[precision]


def test_in_place_iterations_and_subset():
    """project(n_iter=3) once equals three project(n_iter=1) calls (the queries stay on the
    device between calls, Fastapss.cu:450); project(n < m) touches only the first n."""
    from mlodsurfaces_amd import apss as A
    p, n = ds.sphere(20000, seed=33)
    tree = tree_for("sphere20k", p, n)
    R = ds.bbox_radius(p)
    K = A.Kernel(**KERNELS["mg4"](R))
    q = ds.jitter_queries(p[:5000], 0.01, seed=7)
    ap = A.Fastapss(tree, K)
    ap.setScalingProtectionSphere(3.0)
    ap.copyPointsToGPU(len(q), q)
    a3, _ = ap.project(len(q), np.zeros_like(q), None, 3, K)
    ap.copyPointsToGPU(len(q), q)
    for _ in range(3):
        b3, _ = ap.project(len(q), np.zeros_like(q), None, 1, K)
    assert np.array_equal(a3, b3)
    ap.copyPointsToGPU(len(q), q)
    part, _ = ap.project(1000, np.zeros((1000, 3), np.float32), None, 1, K)
    full, _ = ap.download()
    assert np.array_equal(full[1000:], q[1000:])
    assert np.array_equal(full[:1000], part)
    ap.copyPointsToGPU(len(q), q)
    one, _ = ap.project(len(q), np.zeros_like(q), None, 1, K)
    assert np.array_equal(one[:1000], part)


@pytest.mark.parametrize("precision", ["fp32", "fp32c", "fp64"])
def test_query_order_invariance(precision):
    """Each query's result is independent of the batch it is in and of its position."""
    from gpu_util import gpu_project
    p, n = ds.sphere(20000, seed=33)
    tree = tree_for("sphere20k", p, n)
    kern = KERNELS["mg4"](ds.bbox_radius(p))
    q = ds.jitter_queries(p[:4000], 0.01, seed=8)
    perm = np.random.default_rng(0).permutation(len(q))
    a, an, _ = gpu_project(tree, q, kern, cutoff=3.0, precision=precision)
    b, bn, _ = gpu_project(tree, q[perm], kern, cutoff=3.0, precision=precision)
    c, cn, _ = gpu_project(tree, q[:777], kern, cutoff=3.0, precision=precision)
    assert np.array_equal(a[perm], b) and np.array_equal(an[perm], bn)
    assert np.array_equal(a[:777], c)


def test_step_clamp_and_plane_closed_form():
    """Planar samples with constant normals: u4 = 0, so the plane branch projects orthogonally
    (CPUUtils.cpp:751-757); a step bound caps the displacement (CPUUtils.cpp:778-781)."""
    from gpu_util import gpu_project, gpu_tree
    rng = np.random.default_rng(3)
    p = np.concatenate([rng.uniform(-1, 1, (8000, 2)), np.zeros((8000, 1))], 1).astype(np.float32)
    n = np.tile(np.array([0, 0, 1], np.float32), (8000, 1))
    tree = gpu_tree(p, n)
    kern = dict(type=2, sigmas=[0.05, 0.1])
    q = np.concatenate([rng.uniform(-0.5, 0.5, (500, 2)), rng.uniform(-0.05, 0.05, (500, 1))], 1).astype(np.float32)
    for sem in ("ref_gpu", "ref_cpu"):
        for prec in ("fp32", "fp32c", "fp64"):
            op, on, _ = gpu_project(tree, q, kern, cutoff=2.0, semantics=sem, precision=prec)
            tol = 1e-5 if prec == "fp32" else 1e-7  # plain FP32: global-frame cancellation
            assert np.abs(op[:, 2]).max() < tol, (sem, prec, np.abs(op[:, 2]).max())
            assert np.allclose(op[:, :2], q[:, :2], atol=tol), (sem, prec)
            assert np.allclose(on, [0, 0, 1], atol=1e-5), (sem, prec)
            op, _, _ = gpu_project(tree, q, kern, cutoff=2.0, semantics=sem, precision=prec, step_max=0.01)
            step = np.linalg.norm(op - q, axis=1)
            assert step.max() <= 0.01 * (1 + 1e-5)


def test_sphere_closed_form():
    from gpu_util import gpu_project, gpu_tree
    p, n = ds.sphere(50000, seed=34, noise=0.0)
    tree = gpu_tree(p, n)
    R = ds.bbox_radius(p)
    u = np.random.default_rng(1).standard_normal((3000, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    q = (u * np.random.default_rng(2).uniform(0.97, 1.03, (3000, 1))).astype(np.float32)
    op, on, _ = gpu_project(tree, q, KERNELS["mg4"](R), cutoff=3.0, precision="fp32c")
    assert np.abs(np.linalg.norm(op.astype(np.float64), axis=1) - 1).max() < 2e-6
    assert np.abs(np.sum(on * op, 1) / np.linalg.norm(op, axis=1) - 1).max() < 1e-5


def test_state_errors():
    from mlodsurfaces_amd import apss as A
    from mlodsurfaces_amd.apss import MlodError
    p, n = ds.sphere(500)
    t = A.APSSOctree().build(p, n)
    ap = A.Fastapss(t)
    ap.copyPointsToGPU(10, p[:10])
    with pytest.raises(MlodError):
        ap.project(10, None, None, 1)  # no kernel installed yet
    ap.updateKernel(A.Kernel(type=0, sigma=0.1))
    ap.project(0, None, None, 1)  # empty batch is a no-op (Fastapss.cu:441-445)
    with pytest.raises(MlodError):
        ap.project(11, None, None, 1)  # more than uploaded


NB_GOLDEN = [n for n in golden_names("proj_") if "knn" in n or "ball" in n]


def _ball_counts(P, Q, r):
    P = P.astype(np.float64)
    return np.array([(np.sum((P - q) ** 2, 1) <= r * r).sum() for q in np.asarray(Q, np.float64)])


@pytest.mark.parametrize("precision", ["fp32", "fp32c", "fp64"])
@pytest.mark.parametrize("name", NB_GOLDEN)
def test_knn_ball_match_reference_golden(oracle_built, name, precision):
    """kNN / ball APSS (the reference's CPU-only projectAPSSCPU) against its golden vectors
    and the FP64 oracle.  Balls holding fewer than 4 points give degenerate fits (a single
    point's sphere fit is 0/0) whose output is rounding noise in every form: they are left out.
    Every other query must come out finite and within tolerance."""
    from gpu_util import gpu_project, rel_err
    g = load_golden(name)
    tn = str(g["tree"])
    t = load_golden(f"tree_{tn}.npz")
    print("case", name, precision, flush=True)
    tree = golden_tree(tn, t)
    mode = MODE_NAMES[int(g["mode"])]
    kern = golden_kernel(g)
    op, on, _ = gpu_project(tree, g["queries"], kern, semantics="ref_cpu", precision=precision,
                            n_iter=int(g["n_iter"]), mode=mode, knn=int(g["knn"]), ball_radius=float(g["ball_radius"]))
    diag = 2 * float(t["radius"][0])
    ref = g["out_points"]
    ok = np.isfinite(ref).all(1)
    if mode == "ball":
        r = 3 * float(g["kernel_sigmas"][0])
        ok &= _ball_counts(t["points"], np.where(np.isfinite(ref), ref, 0), r) >= 4
        ok &= _ball_counts(t["points"], g["queries"], r) >= 4
    assert ok.sum() >= len(ok) // 2
    assert np.isfinite(op[ok]).all()
    e = rel_err(op[ok], ref[ok], diag)
    assert np.median(e) < 1e-7 and e.max() < TOL_REF, e.max()
    o = oracle_built.OracleTree(t["points"], t["normals"], **oracle_tree_opts(t))
    prm = oracle_built.make_params(semantics="ref_cpu", mode=mode, knn=int(g["knn"]), ball_radius=float(g["ball_radius"]))
    o64, _ = o.project(g["queries"], oracle_built.make_kernel(**kern), prm, n_iter=int(g["n_iter"]), precision=1)
    ok &= np.isfinite(o64).all(1)
    e = rel_err(op[ok], o64[ok].astype(np.float32), diag)
    assert e.max() < TOL_ORACLE[precision], e.max()


@pytest.mark.parametrize("k", [5, 20, 32, 50])
@pytest.mark.parametrize("precision", ["fp32", "fp32c"])
def test_knn_matches_oracle(oracle_built, k, precision):
    from gpu_util import gpu_project, rel_err
    p, n = ds.torus(30000, seed=51)
    tree = tree_for("torus30k", p, n)
    R = ds.bbox_radius(p)
    kern = KERNELS["mg4"](R)
    q = ds.jitter_queries(p[:2000], 0.01, seed=52)
    op, _, _ = gpu_project(tree, q, kern, semantics="ref_cpu", precision=precision, n_iter=2, mode="knn", knn=k)
    o = oracle_built.OracleTree(p, n)
    o64, _ = o.project(q, oracle_built.make_kernel(**kern), oracle_built.make_params(mode="knn", knn=k), n_iter=2,
                       precision=1)
    ok = np.isfinite(o64).all(1)
    assert ok.mean() > 0.99
    tol = TOL_ORACLE[precision] if k >= 20 else 1e-4  # few neighbours: ill-conditioned fits
    if precision == "fp32":
        # the plain FP32 form is held to the reference's own FP32 arithmetic on the same queries
        # (at k = 20 this torus set has fits whose FP32 cancellation reaches ~1.7e-5 diag)
        o32, _ = o.project(q, oracle_built.make_kernel(**kern), oracle_built.make_params(mode="knn", knn=k), n_iter=2,
                           precision=0)
        e32 = rel_err(o32[ok], o64[ok], 2 * R)
        tol = max(tol, 2 * e32.max())
    e = rel_err(op[ok], o64[ok].astype(np.float32), 2 * R)
    assert np.median(e) < 1e-7 and e.max() < tol, e.max()


def test_ball_empty_returns_start():
    """A query whose ball is empty comes back unchanged with a zero normal (CPUUtils.cpp:711-716)."""
    from gpu_util import gpu_project
    p, n = ds.sphere(5000, seed=53)
    tree = tree_for("sphere5k_b", p, n)
    q = np.array([[3.0, 0, 0], [0, 2.5, 0]], np.float32)
    op, on, _ = gpu_project(tree, q, dict(type=2, sigmas=[0.01, 0.02]), mode="ball", ball_radius=-1.0, n_iter=3)
    assert np.array_equal(op, q) and np.all(on == 0)
