"""GPU octree build: bit-exact cell assignment, topology, boxes, areas and FP32 node moments
against the reference's own CPU build (golden vectors) and the oracle."""
import numpy as np
import pytest

from conftest import golden_names, gpu_tree_from_fixture, load_golden
from mlodsurfaces_amd import datasets as ds
from mlodsurfaces_amd.apss import MlodError

pytestmark = pytest.mark.gpu

KEYS = ["depth", "bb", "center", "radius", "father", "first_child", "next_brother", "num_children", "leaf_count"]


def compare(g, ref, moments_ref=None, areas_ref=None):
    for k in KEYS:
        assert np.array_equal(g[k], ref[k]), k
    assert np.array_equal(g["leaf_idx"], ref["leaf_idx"])
    if areas_ref is not None:
        assert np.array_equal(g["areas"], areas_ref)
    if moments_ref is not None:
        assert np.array_equal(g["moments"], moments_ref)


@pytest.mark.parametrize("name", golden_names("tree_"))
def test_build_matches_reference_golden(name):
    from gpu_util import gpu_tree
    r = load_golden(name)
    print("case", name, flush=True)
    t = gpu_tree_from_fixture(r)
    assert t.info()["ties"] == 0
    compare(t.dump(), r, r["moments"], r["areas"])
    assert np.allclose(t.info()["bbox"], r["bbox"])


CASES = [("n1", 1), ("n2", 2), ("n3", 3), ("sphere50k", 50000), ("torus120k", 120000), ("sq60k", 60000)]


@pytest.mark.parametrize("name,n", CASES)
def test_build_matches_oracle(oracle_built, name, n):
    from gpu_util import gpu_tree
    print("case", name, flush=True)
    gen = ds.torus if name.startswith("torus") else ds.superquadric if name.startswith("sq") else ds.sphere
    p, nr = gen(n)
    o = oracle_built.OracleTree(p, nr)
    t = gpu_tree(p, nr)
    inf = t.info()
    assert inf["n_nodes"] == o.num_nodes
    assert inf["ties"] == o.ties
    mf, md = o.moments()
    compare(t.dump(), o.nodes(), mf, o.areas())


def test_build_duplicates_and_max_depth(oracle_built):
    from gpu_util import gpu_tree
    p, nr = ds.sphere(20000, seed=21, noise=0.0)
    p = np.concatenate([p, p[:300], p[:100]])
    nr = np.concatenate([nr, nr[:300], nr[:100]])
    for md in (12, 8, 3):
        print("case dups maxDepth", md, flush=True)
        o = oracle_built.OracleTree(p, nr, max_depth=md)
        t = gpu_tree(p, nr, maxDepth=md)
        mf, _ = o.moments()
        compare(t.dump(), o.nodes(), mf, o.areas())


@pytest.mark.parametrize("policy", [0, 1])
def test_build_ties_follow_policy(oracle_built, policy):
    """Points on exact split planes (a grid) meet ties at every level."""
    from gpu_util import gpu_tree
    g = np.linspace(-1, 1, 17, dtype=np.float32)
    p = np.stack(np.meshgrid(g, g, g[:5]), -1).reshape(-1, 3).astype(np.float32)
    nr = np.tile(np.array([0, 0, 1], np.float32), (len(p), 1))
    o = oracle_built.OracleTree(p, nr, tie_policy=policy)
    t = gpu_tree(p, nr, tie_policy=policy)
    assert o.ties > 0 and t.info()["ties"] == o.ties
    mf, _ = o.moments()
    compare(t.dump(), o.nodes(), mf, o.areas())


def test_build_rejects_bad_options():
    from gpu_util import gpu_tree
    from mlodsurfaces_amd.apss import APSSOctree
    p, nr = ds.sphere(100)
    with pytest.raises(MlodError):
        gpu_tree(p, nr, maxDepth=13)
    with pytest.raises(MlodError):
        gpu_tree(p[:0], nr[:0])
    with pytest.raises(MlodError):
        APSSOctree().build(p, nr, updateRootBoundingBox=False)  # no root box to keep
    for bb, BB in [((0, 0, 0), (1, -1, 1)), ((0, np.nan, 0), (1, 1, 1)), ((0, 0, 0), (1, 1, np.inf))]:
        t = APSSOctree()
        t.setRootBoundingBox(np.array(bb, np.float32), np.array(BB, np.float32))
        with pytest.raises(MlodError):
            t.build(p, nr, updateRootBoundingBox=False)


def _dup_cloud(n, seed):
    p, nr = ds.sphere(n, seed=seed, noise=0.0)
    return np.concatenate([p, p[:n // 20], p[:n // 60]]), np.concatenate([nr, nr[:n // 20], nr[:n // 60]])


@pytest.mark.parametrize("max_pts,max_depth", [(2, 12), (4, 8), (3, 12), (16, 10), (64, 12), (0, 7), (0, 12),
                                               (10 ** 6, 12), (5, 1)])
def test_build_points_per_leaf_matches_oracle(oracle_built, max_pts, max_depth):
    """maxNumberOfPointsPerLeaf (CPUUtils.cpp:279-283): a node with at most that many points is a
    leaf (0: only at maxDepth; more than the cloud: the root alone).  Topology, boxes, leaf point
    lists, areas and FP32 moments equal the oracle's bit for bit, on a cloud with duplicated
    points (leaves at maxDepth holding more than max_pts points)."""
    from gpu_util import gpu_tree
    print("case max_pts", max_pts, "max_depth", max_depth, flush=True)
    for gen, n, seed in ((_dup_cloud, 30000, 31), (lambda n, seed: ds.torus(n, seed=seed), 50000, 32)):
        p, nr = gen(n, seed)
        o = oracle_built.OracleTree(p, nr, max_depth=max_depth, max_pts_per_leaf=max_pts)
        t = gpu_tree(p, nr, maxDepth=max_depth, maxNumberOfPointsPerLeaf=max_pts)
        assert t.info()["n_nodes"] == o.num_nodes and t.info()["ties"] == o.ties
        mf, _ = o.moments()
        compare(t.dump(), o.nodes(), mf, o.areas())
        lc = t.dump()["leaf_count"]
        if 0 < max_pts < len(p):
            assert lc.max() > 1
        if max_pts >= len(p):
            assert o.num_nodes == 1


@pytest.mark.parametrize("grow,max_pts", [((0.1, 0.2, 0.05, 0.3, 0.0, 0.15), 1), ((0.0, 0.0, 0.0, 0.0, 0.0, 0.0), 2),
                                          ((-0.2, 0.1, 0.0, -0.1, 0.05, 0.3), 1)])
def test_build_root_box_matches_oracle(oracle_built, grow, max_pts):
    """updateRootBoundingBox = false (CPUUtils.cpp:547): the root keeps a box given beforehand,
    here the points' box grown (or, with negative factors, shrunk so that points lie outside the
    root) by the given fractions of its extent; splits, boxes, radii, areas and moments follow
    that box and equal the oracle's."""
    from mlodsurfaces_amd.apss import APSSOctree
    print("case root box", grow, max_pts, flush=True)
    p, nr = ds.superquadric(40000, seed=33)
    lo, hi = p.min(0).astype(np.float64), p.max(0).astype(np.float64)
    ext = hi - lo
    g = np.asarray(grow)
    box = np.concatenate([lo - g[:3] * ext, hi + g[3:] * ext]).astype(np.float32)
    o = oracle_built.OracleTree(p, nr, max_pts_per_leaf=max_pts, root_box=box)
    t = APSSOctree()
    t.setRootBoundingBox(box[:3], box[3:])
    t.build(p, nr, updateRootBoundingBox=False, maxNumberOfPointsPerLeaf=max_pts)
    assert np.array_equal(np.asarray(t.info()["bbox"], np.float32), box)
    assert t.info()["ties"] == o.ties
    mf, _ = o.moments()
    compare(t.dump(), o.nodes(), mf, o.areas())
    # a rebuild with updateRootBoundingBox = false keeps the box of the previous build
    t.build(p[:20000], nr[:20000], updateRootBoundingBox=False, maxNumberOfPointsPerLeaf=max_pts)
    assert np.array_equal(np.asarray(t.info()["bbox"], np.float32), box)
