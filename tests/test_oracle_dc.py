"""Pins the oracle's dual contouring (oracle/dc_oracle.cpp, a restatement of uniformize,
minimal_edges::traverse and dual_contour::extract; the reference's own code needs Eigen and cannot
be built here) with properties that a wrong index, order or sign breaks:
  * uniformize equals the same pipeline written as set operations (shift under the new root,
    graded closure to a fixed point, ancestors, sibling sets); internal nodes have all eight
    children; re-rooting over the depth-0 cells' extent;
  * extraction on a uniform grid equals an independent grid formulation (every interior grid edge
    with a sign change gives the quad of its four cells, oriented by the sign of its first corner);
  * closed surfaces give closed, consistently oriented meshes (every edge in exactly two triangles,
    each direction once) with the sphere's Euler characteristic, on uniform and graded octrees;
  * vertices are the mean of their cell's edge crossings, inside the cell."""
import itertools
from collections import Counter

import numpy as np
import pytest


def uniform_ids(d):
    g = np.arange(2 ** d)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    return np.stack([np.full(X.size, d), X.ravel(), Y.ravel(), Z.ravel()], 1)


def leaf_corners(nodes, leaf, D):
    L = nodes[leaf]
    sc = (1 << (D - L[:, 0]))[:, None]
    return np.unique(np.concatenate([(L[:, 1:] + np.array([c & 1, (c >> 1) & 1, (c >> 2) & 1])) * sc
                                     for c in range(8)]), axis=0)


def sphere_sdf(c=(0.5, 0.5, 0.5), r=0.3):
    return lambda p: np.linalg.norm(p - np.asarray(c), axis=1) - r


def run(oracle_built, ids, f, cut=False, max_depth=0):
    dc = oracle_built.OracleDC(ids, max_depth, cut)
    nodes, leaf = dc.nodes()
    diag, origin, D = dc.frame()
    cs = leaf_corners(nodes, leaf, D)
    vals = f(cs / 2.0 ** D).astype(np.float32)
    dc.set_corners(cs, vals)
    v, vn, t = dc.extract()
    return dc, nodes, leaf, D, cs, vals, v, vn, t


def manifold(t):
    E = Counter()
    for a, b, c in t:
        for e in ((a, b), (b, c), (c, a)):
            E[e] += 1
    und = Counter()
    for (a, b), k in E.items():
        und[(min(a, b), max(a, b))] += k
    return max(E.values()), set(und.values()), len(und)


def test_uniformize_invariants(oracle_built):
    # cells near a sphere at depth 6 of the [-1, 1]^3 octree, reaching a neighbouring depth-0 cell
    d = 6
    g = (np.arange(-2 ** d, 2 * 2 ** d) + 0.5) / 2 ** d
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    P = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    near = np.abs(np.linalg.norm(P - 0.8, axis=1) - 0.5) < 0.03
    ids = np.concatenate([np.full((near.sum(), 1), d), np.floor(P[near] * 2 ** d).astype(int)], 1)
    dc = oracle_built.OracleDC(ids)
    nodes, leaf = dc.nodes()
    diag, origin, D = dc.frame()
    d0 = np.unique(np.floor(P[near]).astype(int), axis=0)
    ext = d0.max(0) - d0.min(0) + 1
    delta = int(max(np.ceil(np.log2(e)) for e in ext))
    assert diag == 2.0 ** delta and np.array_equal(origin, d0.min(0).astype(np.float32)) and D == d + delta
    S = set(map(tuple, nodes))
    inside = lambda n: all(0 <= c < 2 ** n[0] for c in n[1:])  # noqa: E731
    # the same pipeline as set operations: shift under the new root, graded closure (parents of the
    # diagonal neighbours, to a fixed point), all ancestors, full sibling sets
    sh = [(int(r[0]) + delta, *(int(c) - int(m) * 2 ** int(r[0]) for c, m in zip(r[1:], d0.min(0))))
          for r in np.concatenate([ids, np.concatenate([np.zeros((len(d0), 1), int), d0], 1)])]
    G, todo = set(sh), list(sh)
    while todo:
        dd, x, y, z = todo.pop()
        if dd == 0:
            continue
        for sx, sy, sz in itertools.product((-1, 1), repeat=3):
            p = (dd - 1, (x + sx) >> 1, (y + sy) >> 1, (z + sz) >> 1)
            if p not in G:
                G.add(p)
                todo.append(p)
    A = set(G)
    for n in G:
        dd, x, y, z = n
        while dd > 0:
            dd, x, y, z = dd - 1, x >> 1, y >> 1, z >> 1
            A.add((dd, x, y, z))
    full = set(A) | {(0, 0, 0, 0)}
    for dd, x, y, z in A:
        if dd > 0:
            for c in range(8):
                full.add((dd, 2 * (x >> 1) + (c & 1), 2 * (y >> 1) + ((c >> 1) & 1), 2 * (z >> 1) + ((c >> 2) & 1)))
    assert {n for n in S if inside(n)} == {n for n in full if inside(n)}
    # internal nodes have all eight children; leaves none
    for n, isleaf in zip(map(tuple, nodes), leaf):
        if not inside(n):
            continue
        dd, x, y, z = n
        kids = [(dd + 1, 2 * x + (c & 1), 2 * y + ((c >> 1) & 1), 2 * z + ((c >> 2) & 1)) in S for c in range(8)]
        assert (not any(kids)) if isleaf else all(kids)


def grid_triangles(d, f):
    """Independent grid formulation on the uniform depth-d grid: interior edge along axis a with
    a sign change between its ends -> the four cells around it in the traversal's order
    (b-, c-), (b+, c-), (b-, c+), (b+, c+) for the other axes (X: y, z; Y: x, z; Z: x, y), two
    triangles (0, 1, 3), (0, 3, 2), flipped when the first corner (low end along X and Z, high
    end along Y) is negative."""
    n = 2 ** d
    tris = set()
    perp = {0: (1, 2), 1: (0, 2), 2: (0, 1)}
    for a in range(3):
        b, c = perp[a]
        for pos in itertools.product(range(n + 1), repeat=3):
            p = list(pos)
            if p[a] >= n or not (0 < p[b] < n and 0 < p[c] < n):
                continue
            lo = np.array(p, float)
            hi = lo.copy()
            hi[a] += 1
            first, second = (hi, lo) if a == 1 else (lo, hi)
            f0, f1 = f(first[None] / n)[0], f(second[None] / n)[0]
            f0, f1 = np.float32(f0), np.float32(f1)
            if (f0 < 0) == (f1 < 0):
                continue
            cells = []
            for qc in (0, 1):
                for qb in (0, 1):
                    cc = list(p)
                    cc[b] = p[b] - 1 + qb
                    cc[c] = p[c] - 1 + qc
                    cells.append((d, *cc))
            m0 = f0 < 0
            t1 = (cells[0], cells[3] if m0 else cells[1], cells[1] if m0 else cells[3])
            t2 = (cells[0], cells[2] if m0 else cells[3], cells[3] if m0 else cells[2])
            tris.add(t1)
            tris.add(t2)
    return tris


def test_uniform_grid_matches_grid_formulation(oracle_built):
    d = 4
    f = sphere_sdf((0.47, 0.52, 0.5), 0.31)
    dc, nodes, leaf, D, cs, vals, v, vn, t = run(oracle_built, uniform_ids(d), f)
    # the re-rooted grid: one depth-0 cell, so the cells keep their depth
    got = set(tuple(tuple(int(q) for q in nodes[vn[i]]) for i in tri) for tri in t)
    assert got == grid_triangles(d, f)
    assert len(t) > 100


@pytest.mark.parametrize("kind", ["uniform", "graded", "cut"])
def test_closed_meshes(oracle_built, kind):
    f = sphere_sdf((0.5, 0.45, 0.55), 0.3)
    if kind == "uniform":
        ids = uniform_ids(4)
    else:
        d = 6
        g = (np.arange(2 ** d) + 0.5) / 2 ** d
        X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
        P = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
        near = np.abs(f(P)) < 0.04
        ids = np.concatenate([np.full((near.sum(), 1), d), np.floor(P[near] * 2 ** d).astype(int)], 1)
    dc, nodes, leaf, D, cs, vals, v, vn, t = run(oracle_built, ids, f, cut=kind == "cut", max_depth=5)
    if kind == "cut":
        # cut (upsample.cpp:95-100) before the sibling sets: the uncut octree's nodes down to depth 5,
        # whose deepest nodes become leaves
        full, _ = oracle_built.OracleDC(ids).nodes()
        assert D == 5 and np.array_equal(nodes, full[full[:, 0] <= 5])
        assert leaf[nodes[:, 0] == 5].all()
    dup, counts, ne = manifold(t)
    assert dup == 1 and counts == {2}
    assert len(v) - ne + len(t) == 2
    # outward orientation: the signed volume of the closed mesh is positive and near (4/3) pi r^3
    a, b, c = v[t[:, 0]].astype(np.float64), v[t[:, 1]].astype(np.float64), v[t[:, 2]].astype(np.float64)
    vol = np.einsum("ij,ij->i", a, np.cross(b, c)).sum() / 6
    assert abs(abs(vol) / (4 / 3 * np.pi * 0.3 ** 3) - 1) < 0.05
    # vertices: the mean of the cell's edge crossings, inside the cell
    cmap = {tuple(k): float(x) for k, x in zip(cs, vals)}
    E = [(0, 1), (0, 2), (1, 3), (2, 3), (4, 5), (4, 6), (5, 7), (6, 7), (0, 4), (1, 5), (2, 6), (3, 7)]
    for i in range(0, len(v), max(1, len(v) // 200)):
        dd, x, y, z = nodes[vn[i]]
        s = 1 << (D - dd)
        pts = []
        for e0, e1 in E:
            u = np.array([x + (e0 & 1), y + ((e0 >> 1) & 1), z + ((e0 >> 2) & 1)]) * s
            w = np.array([x + (e1 & 1), y + ((e1 >> 1) & 1), z + ((e1 >> 2) & 1)]) * s
            fu, fw = cmap[tuple(u)], cmap[tuple(w)]
            if (fu < 0) != (fw < 0):
                tt = -fu / (fw - fu)
                pts.append((tt * w + (1 - tt) * u) / 2.0 ** D)
        lo = np.array([x, y, z]) / 2.0 ** dd
        hi = (np.array([x, y, z]) + 1) / 2.0 ** dd
        if pts:
            assert np.allclose(v[i], np.mean(pts, 0), atol=1e-6)
        assert np.all(v[i] >= lo - 1e-7) and np.all(v[i] <= hi + 1e-7)


def test_mesh_passes(oracle_built):
    """smoothMesh / subdivideMesh restated: smoothing = mean of the other two vertices of every
    incident triangle (an independent float64 scatter), subdivision = a midpoint per edge, numbered
    by first use, four triangles per triangle, the closed mesh staying closed."""
    f = sphere_sdf((0.5, 0.45, 0.55), 0.3)
    dc, nodes, leaf, D, cs, vals, v, vn, t = run(oracle_built, uniform_ids(4), f)
    S = oracle_built.oracle_mesh_smooth(v, t)
    acc = np.zeros((len(v), 3))
    cnt = np.zeros(len(v))
    for a, b, c in t.astype(int):
        acc[a] += v[b].astype(np.float64) + v[c]
        acc[b] += v[a].astype(np.float64) + v[c]
        acc[c] += v[a].astype(np.float64) + v[b]
        cnt[[a, b, c]] += 2
    assert np.allclose(S, acc / cnt[:, None], atol=1e-6)
    P, T = oracle_built.oracle_mesh_subdivide(S, t)
    edges = {}
    for a, b, c in t.astype(int):
        for u, w in ((a, b), (c, b), (c, a)):
            edges.setdefault((min(u, w), max(u, w)), len(v) + len(edges))
    assert len(P) == len(v) + len(edges) and len(T) == 4 * len(t)
    for (u, w), i in edges.items():
        assert np.array_equal(P[i], ((S[u] + S[w]) / np.float32(2)).astype(np.float32))
    dup, counts, ne = manifold(T)
    assert dup == 1 and counts == {2} and len(P) - ne + len(T) == 2


def test_uniformize_from_points(oracle_built):
    """uniformizeFromPoints (upsample.cpp:253-259): for cells inside the root it equals uniformize
    (whose only depth-0 ancestor is then the root and whose re-rooting is the identity); for cells
    reaching outside it equals the set pipeline without a shift (graded closure, ancestors, sibling
    sets), in the [-1, 1]^3 frame (diag 1, origin 0)."""
    d = 5
    g = (np.arange(2 ** d) + 0.5) / 2 ** d
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    P = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    near = np.abs(np.linalg.norm(P - 0.5, axis=1) - 0.3) < 0.05
    ids = np.concatenate([np.full((near.sum(), 1), d), np.floor(P[near] * 2 ** d).astype(int)], 1)
    a, la = oracle_built.OracleDC(ids, from_points=True).nodes()
    b, lb = oracle_built.OracleDC(ids).nodes()
    assert np.array_equal(a, b) and np.array_equal(la, lb)
    # reaching outside the root: no re-rooting
    g = (np.arange(-4, 2 ** d + 4) + 0.5) / 2 ** d
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    P = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    near = np.abs(np.linalg.norm(P - 0.6, axis=1) - 0.5) < 0.04
    ids = np.concatenate([np.full((near.sum(), 1), d), np.floor(P[near] * 2 ** d).astype(int)], 1)
    dc = oracle_built.OracleDC(ids, from_points=True)
    nodes, leaf = dc.nodes()
    diag, origin, D = dc.frame()
    assert diag == 1.0 and np.array_equal(origin, np.zeros(3, np.float32)) and D == d
    G, todo = set(map(tuple, ids.tolist())), [tuple(r) for r in ids.tolist()]
    while todo:
        dd, x, y, z = todo.pop()
        if dd == 0:
            continue
        for sx, sy, sz in itertools.product((-1, 1), repeat=3):
            q = (dd - 1, (x + sx) >> 1, (y + sy) >> 1, (z + sz) >> 1)
            if q not in G:
                G.add(q)
                todo.append(q)
    A = set(G)
    for n in G:
        dd, x, y, z = n
        while dd > 0:
            dd, x, y, z = dd - 1, x >> 1, y >> 1, z >> 1
            A.add((dd, x, y, z))
    full = set(A) | {(0, 0, 0, 0)}
    for dd, x, y, z in A:
        if dd > 0:
            for c in range(8):
                full.add((dd, 2 * (x >> 1) + (c & 1), 2 * (y >> 1) + ((c >> 1) & 1), 2 * (z >> 1) + ((c >> 2) & 1)))
    inside = lambda n: all(0 <= c < 2 ** n[0] for c in n[1:])  # noqa: E731
    assert {n for n in map(tuple, nodes.tolist()) if inside(n)} == {n for n in full if inside(n)}
