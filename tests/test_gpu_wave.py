"""The wave-shared MLoD walk (MLOD_TRAVERSAL_WAVE) against the per-lane walk and the reference.

The wave walk visits the union of its lanes' paths depth first, in child order, and each lane
takes part only in the nodes its own walk visits, so every lane adds its terms in the order of its
own walk.  The FP64 form must give the per-lane walk's bits exactly.  The FP32 forms run the same
operations, but the compiler may fuse a multiply-add differently in the two loops (a last-place
difference in a sum, which the plain form's global-frame fit can magnify to a few 1e-6 x diag),
so they are held to the per-lane walk within their budgets, 1e-5 x diag (plain FP32) and 1e-7 x
diag (compensated), and to the same tolerances as the per-lane walk against the
reference's CPU outputs (1e-5 x diag) and the FP64 oracle (1e-5 / 1e-6 / 1e-7)."""
import numpy as np
import pytest

from conftest import MODE_NAMES, golden_kernel, golden_names, gpu_tree_from_fixture, load_golden
from mlodsurfaces_amd import apss as A
from mlodsurfaces_amd import datasets as ds

pytestmark = pytest.mark.gpu

_trees = {}


def golden_tree(name, t):
    if name not in _trees:
        _trees[name] = gpu_tree_from_fixture(t)
    return _trees[name]


MLOD_GOLDEN = [n for n in golden_names("proj_") if "knn" not in n and "ball" not in n and "global" not in n]


TOL_LANE = {"fp32": 1e-5, "fp32c": 1e-7}
TOL_ORACLE = {"fp32": 1e-5, "fp32c": 1e-6, "fp64": 1e-7}


@pytest.mark.parametrize("semantics", ["ref_gpu", "ref_cpu"])
@pytest.mark.parametrize("name", MLOD_GOLDEN)
def test_wave_walk_matches_lane_walk_on_golden_sets(oracle_built, name, semantics):
    from gpu_util import gpu_project, rel_err
    from conftest import oracle_tree_opts
    g = load_golden(name)
    tn = str(g["tree"])
    t = load_golden(f"tree_{tn}.npz")
    tree = golden_tree(tn, t)
    assert MODE_NAMES[int(g["mode"])] == "mlod"
    kern = golden_kernel(g)
    diag = 2 * float(t["radius"][0])
    o = oracle_built.OracleTree(t["points"], t["normals"], **oracle_tree_opts(t))
    prm = oracle_built.make_params(cutoff=float(g["cutoff"]), min_depth=int(g["min_depth"]), semantics=semantics,
                                   mode="mlod")
    o64, _ = o.project(g["queries"], oracle_built.make_kernel(**kern), prm, n_iter=int(g["n_iter"]), precision=1)
    for precision in ("fp32", "fp32c", "fp64"):
        print("case", name, semantics, precision, flush=True)
        out = {}
        for trav in ("lane", "wave"):
            out[trav] = gpu_project(tree, g["queries"], kern, cutoff=float(g["cutoff"]), min_depth=int(g["min_depth"]),
                                    semantics=semantics, precision=precision, n_iter=int(g["n_iter"]),
                                    traversal=trav)[:2]
        (lp, ln), (wp, wn) = out["lane"], out["wave"]
        if precision == "fp64":
            np.testing.assert_array_equal(wp.view(np.uint32), lp.view(np.uint32))
            np.testing.assert_array_equal(wn.view(np.uint32), ln.view(np.uint32))
        else:
            same = float(np.mean((wp.view(np.uint32) == lp.view(np.uint32)).all(1)))
            print(f"  {precision}: {same:.4f} of the points bit-equal to the per-lane walk", flush=True)
            assert rel_err(wp, lp, diag).max() < TOL_LANE[precision]
            assert np.abs(wn - ln).max() < 1e-4
        assert rel_err(wp, o64.astype(np.float32), diag).max() < TOL_ORACLE[precision]
        if semantics == "ref_cpu":  # the golden outputs are the reference's projectCPU
            assert rel_err(wp, g["out_points"], diag).max() < 1e-5


def test_wave_walk_counters_and_batch_invariance():
    """On the headline workload at reduced size: the per-lane counters (nodes tested, far, near,
    leaf points, lane iterations) are the same for both walks, the wave walk's lane slots are its
    wavefront iterations, and a query's compensated result is the same bits alone (any other
    wavefront composition) as in the full batch."""
    w = ds.make_workload("superquadric8m", 0.02)
    P, N, Q = w["points"], w["normals"], w["queries"]
    tree = A.APSSOctree().build(P, N)
    K = A.Kernel(**w["kernel"])
    counts, full = {}, {}
    for trav in ("lane", "wave"):
        ap = A.Fastapss(tree, K)
        ap.setScalingProtectionSphere(w["cutoff"])
        ap.set_semantics("ref_gpu")
        ap.set_precision("fp32c")
        ap.set_traversal(trav)
        ap.copyPointsToGPU(len(Q), Q)
        counts[trav] = ap.count_visits()
        full[trav] = ap.project_points(Q, 2)[0]
    print("counters", counts, flush=True)
    for k in ("far", "leaf_points", "tested", "near", "iterations"):
        assert counts["wave"][k] == counts["lane"][k], k
    assert counts["wave"]["lane_slots"] >= counts["wave"]["iterations"]
    ap = A.Fastapss(tree, K)
    ap.setScalingProtectionSphere(w["cutoff"])
    ap.set_semantics("ref_gpu")
    ap.set_precision("fp32c")
    ap.set_traversal("wave")
    sub = np.random.default_rng(4).choice(len(Q), 1000, replace=False)
    alone = ap.project_points(Q[sub], 2)[0]
    np.testing.assert_array_equal(alone.view(np.uint32), full["wave"][sub].view(np.uint32))


def test_traversal_argument_checked():
    from mlodsurfaces_amd._lib import lib
    tree = A.APSSOctree().build(*ds.sphere(500, seed=2))
    assert lib().mlod_set_traversal(tree.handle, 2) != 0
    assert lib().mlod_set_traversal(tree.handle, 1) == 0
