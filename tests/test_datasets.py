"""Synthetic workload generators: deterministic, well-formed."""
import numpy as np

from mlodsurfaces_amd import datasets as ds


def test_generators_deterministic_and_unit_normals():
    for gen in (ds.sphere, ds.torus, ds.superquadric):
        p1, n1 = gen(2000)
        p2, n2 = gen(2000)
        assert p1.dtype == np.float32 and n1.dtype == np.float32 and p1.shape == (2000, 3)
        assert np.array_equal(p1, p2) and np.array_equal(n1, n2)
        assert np.all(np.isfinite(p1)) and np.allclose(np.linalg.norm(n1, axis=1), 1, atol=1e-5)


def test_torus_holes_removed():
    p, _ = ds.torus(20000, seed=2, holes=3, hole_radius=0.1)
    rng = np.random.default_rng(2)
    th = rng.uniform(0, 2 * np.pi, 3)
    ph = rng.uniform(0, 2 * np.pi, 3)
    hc = np.stack([(1 + 0.35 * np.cos(ph)) * np.cos(th), (1 + 0.35 * np.cos(ph)) * np.sin(th), 0.35 * np.sin(ph)], 1)
    for c in hc:
        assert np.min(np.linalg.norm(p - c, axis=1)) > 0.1 - 1e-6


def test_workloads_small():
    w = ds.make_workload("sphere100k", scale=0.01)
    assert len(w["points"]) == 1000 and w["kernel"]["type"] == 2 and len(w["kernel"]["sigmas"]) == 4
    w = ds.make_workload("torus1m", scale=0.001)
    assert w["kernel"]["type"] == 1 and w["queries"].shape == w["points"].shape


def test_cell_centres_of_points_match_generator_oracle(oracle_built):
    """The C5 query set (centres of the depth-d cells enclosing the points) equals the oracle's
    generateFromPoints + aabb_octree::center."""
    p, _ = ds.superquadric(20000, seed=9, holes=3)
    q = ds.cell_centres_of_points(p, 7, cap=10 ** 9, seed=1)
    og = oracle_built.OracleGenerator(None)
    cells = og.generate_from_points(p, 7)
    oc = np.array([oracle_built.oracle_cell_center(c) for c in cells])
    key = lambda a: np.lexsort((a[:, 0], a[:, 1], a[:, 2]))  # noqa: E731
    assert len(q) == len(oc)
    assert np.array_equal(q[key(q)], oc[key(oc)])
    capped = ds.cell_centres_of_points(p, 7, cap=100, seed=1)
    assert len(capped) == 100 and len(np.unique(capped, axis=0)) == 100


def test_pn_normalize_frame():
    """Pn::normalize (Pn.cpp:262-290): centroid at the origin, largest |coordinate| of the box
    grown by 0.5 % of its diagonal equal to 1."""
    p, _ = ds.torus(20000, seed=4)
    q = ds.pn_normalize(p * 3.0 + 5.0)
    assert np.allclose(q.astype(np.float64).mean(0), 0, atol=1e-6)
    ext = np.abs(q).max()
    lo, hi = q.min(0).astype(np.float64), q.max(0).astype(np.float64)
    small = np.linalg.norm(hi - lo) * 0.005
    assert abs(max(np.abs(hi + small).max(), np.abs(lo - small).max()) - 1.0) < 1e-5
    assert ext < 1.0
