"""Pins the oracle's adaptive-octree driver (oracle/adaptive_oracle.cpp, a restatement of
OctreeGenerator::generate / generateFromPoints, upsample.cpp:128-144 and 228-251).

The reference's octree.h / upsample.cpp need Eigen, which is absent here, so they cannot be run.
The oracle is checked instead against (1) an independent set-based restatement written below
from the reference's statements, driven by analytic sphere projections (including failed
projections and surfaces that leave the root cube, which exercise the root fallback, the outer
depth-0 loop and the relative-depth cut), and (2) closed forms of the octree-id arithmetic."""
import math

import numpy as np
import pytest

ob = pytest.importorskip("oracle.bind")


def _sphere_projection(c, r, fail=None, fail_scale=0.0, scale=1.0):
    """Closest-point projection on a sphere; queries where fail(q) holds get their normal scaled
    by fail_scale (0: a failed projection; 0.6: |n|^2 = 0.36, still below the 0.5 threshold)
    and every other normal by `scale`."""
    c = np.asarray(c, np.float64)

    def project(q):
        v = q.astype(np.float64) - c
        L = np.linalg.norm(v, axis=1, keepdims=True)
        n = np.where(L > 0, v / np.where(L > 0, L, 1), 0.0)
        p = c + r * n
        n = n * scale
        if fail is not None:
            bad = fail(q)
            n[bad] *= fail_scale / scale
            p[bad] = q[bad]
        return p.astype(np.float32), n.astype(np.float32)

    return project


def _enclosing(p, d):
    # aabb_octree::enclosing over [-1, 1]^3 in float32: floor(((p + 1) / 2) / 2^-d)
    t = (np.float32(p) - np.float32(-1.0)) / np.float32(2.0) / np.float32(2.0 ** -d)
    return int(math.floor(float(t)))


def _py_generate(max_depth, project):
    """OctreeGenerator::generate restated over Python sets of (depth, x, y, z)."""
    def centre(cell):
        d = cell[0]
        return [-1.0 + (2 * p + 1) * 2.0 ** -d for p in cell[1:]]

    def find_stable(cells):
        cells = sorted(cells)
        C = np.array([centre(c) for c in cells], np.float32)
        P, N = project(C)
        out = set()
        for c, p, n in zip(cells, P, N):
            n = n.astype(np.float32)
            l2 = np.float32(np.float32(n[0] * n[0]) + np.float32(n[1] * n[1])) + np.float32(n[2] * n[2])
            if l2 < np.float32(0.5):
                out.add((0, 0, 0, 0))
            else:
                out.add((c[0],) + tuple(_enclosing(p[a], c[0]) for a in range(3)))
        return out

    def neighbours(c):
        return {(c[0], c[1] + i, c[2] + j, c[3] + k) for i in (-1, 0, 1) for j in (-1, 0, 1) for k in (-1, 0, 1)}

    def compute_leaves(cur, maxd, total):
        for depth in range(maxd + 1):
            sub = find_stable(cur) if cur else set()
            cur = {(d + 1, 2 * x + i, 2 * y + j, 2 * z + k) for (d, x, y, z) in sub
                   for i in (0, 1) for j in (0, 1) for k in (0, 1)}
            for c in list(cur):
                if c[0] == depth + 1:
                    cur |= neighbours(c)
            total |= cur
        return cur

    def anc(c, d):
        s = c[0] - d
        return (d, c[1] >> s, c[2] >> s, c[3] >> s) if s > 0 else c

    depth0, old0, nodes, total = {(0, 0, 0, 0)}, set(), set(), set()
    maxrel = max_depth
    mn, mx = [2 ** 31 - 1] * 3, [-(2 ** 31 - 1)] * 3
    while depth0:
        total |= depth0
        cur = compute_leaves(depth0, maxrel, total)
        rd = 0
        a0 = {anc(c, 0) for c in cur}
        if cur:
            for c in a0:
                for k in range(3):
                    mn[k], mx[k] = min(mn[k], c[k + 1]), max(mx[k], c[k + 1])
            rd = max(int(math.ceil(float(np.float32(math.log2(mx[k] - mn[k] + 1))))) for k in range(3))
        maxrel = max(0, max_depth - rd)
        nodes |= cur
        old0 |= depth0
        nxt = set()
        for c in a0:
            nxt |= neighbours(c)
        depth0 = nxt - old0
    key = lambda c: (c[0], c[3], c[2], c[1])  # noqa: E731  node_comparator: depth, z, y, x
    leaves = sorted({anc(c, maxrel) for c in nodes}, key=key)
    tot = sorted({anc(c, maxrel) for c in total}, key=key)
    return np.array(leaves, np.int32).reshape(-1, 4), np.array(tot, np.int32).reshape(-1, 4), maxrel


CASES = {
    "inner_sphere": (4, _sphere_projection([0.1, -0.05, 0.02], 0.55)),
    "sphere_leaving_root": (4, _sphere_projection([0.1, 0.0, 0.0], 0.97)),
    "failed_projections": (4, _sphere_projection([0.0, 0.0, 0.0], 0.6, fail=lambda q: q[:, 0] > 0.45)),
    "deep_small_sphere": (6, _sphere_projection([-0.3, 0.2, 0.1], 0.12)),
    "short_normals": (4, _sphere_projection([0.0, 0.1, 0.0], 0.6, fail=lambda q: q[:, 1] < -0.3, fail_scale=0.6,
                                            scale=0.8)),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_generate_matches_set_restatement(case):
    max_depth, proj = CASES[case]
    g = ob.OracleGenerator(proj)
    leaves = g.generate(max_depth)
    info = g.info()
    ref_leaves, ref_total, maxrel = _py_generate(max_depth, proj)
    assert info["max_rel_depth"] == maxrel
    assert np.array_equal(leaves, ref_leaves)
    assert np.array_equal(g.total(), ref_total)
    assert info["bad"] == 0
    if case == "sphere_leaving_root":
        assert maxrel < max_depth  # the depth-0 extent cut the depth
        assert info["calls"] > max_depth + 1  # the outer loop ran again from the depth-0 neighbours
    if case in ("failed_projections", "short_normals"):
        assert len(np.unique(leaves[:, 0])) > 1  # root fallback re-enters shallow cells


def test_generate_invariants():
    proj = _sphere_projection([0.0, 0.0, 0.0], 0.5)
    g = ob.OracleGenerator(proj)
    leaves = g.generate(5)
    # sorted by node_comparator and unique
    k = leaves[:, 0].astype(np.int64) << 60 | (leaves[:, 3].astype(np.int64) + 2 ** 19) << 40 | \
        (leaves[:, 2].astype(np.int64) + 2 ** 19) << 20 | (leaves[:, 1].astype(np.int64) + 2 ** 19)
    assert np.all(np.diff(k) > 0)
    # every cell of the final depth that a surface point falls in is generated (for a sphere
    # this smooth at depth 5 the 64-block expansion covers the surface)
    rng = np.random.default_rng(0)
    v = rng.standard_normal((20000, 3))
    s = (0.5 * v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)
    d = int(leaves[:, 0].max())
    cells = {tuple(r) for r in leaves.tolist()}
    for p in s[:2000]:
        assert (d,) + tuple(_enclosing(p[a], d) for a in range(3)) in cells


def test_cell_centre_closed_form():
    rng = np.random.default_rng(1)
    for _ in range(500):
        d = int(rng.integers(0, 14))
        p = rng.integers(-2 ** d - 3, 2 ** d + 3, 3)
        c = ob.oracle_cell_center([d, *p])
        exact = np.array([-1.0 + (2 * v + 1) * 2.0 ** -d for v in p], np.float64)
        assert np.array_equal(c.astype(np.float64), exact)
        # the centre's enclosing cell is the cell itself
        assert [_enclosing(c[a], d) for a in range(3)] == list(p)


def test_cell_ancestor_is_floor_division():
    rng = np.random.default_rng(2)
    for _ in range(500):
        d = int(rng.integers(1, 14))
        p = rng.integers(-3 * 2 ** d, 3 * 2 ** d, 3)
        t = int(rng.integers(0, d))
        a = ob.oracle_cell_ancestor([d, *p], t)
        assert list(a) == [t] + [int(math.floor(v / 2 ** (d - t))) for v in p]


def test_generate_from_points_closed_form():
    rng = np.random.default_rng(3)
    P = rng.uniform(-1.2, 1.2, (5000, 3)).astype(np.float32)
    g = ob.OracleGenerator(None)
    got = g.generate_from_points(P, 7)
    exp = {(7,) + tuple(_enclosing(p[a], 7) for a in range(3)) for p in P}
    assert {tuple(r) for r in got.tolist()} == exp and len(got) == len(exp)
    box = np.array([-0.5, -2.0, 0.0, 1.5, 2.0, 1.0], np.float32)
    got = g.generate_from_points(P, 5, box)
    mn, dg = box[:3], box[3:] - box[:3]
    exp = {(5,) + tuple(int(math.floor(float((p[a] - mn[a]) / dg[a] / np.float32(2.0 ** -5)))) for a in range(3))
           for p in P}
    assert {tuple(r) for r in got.tolist()} == exp
