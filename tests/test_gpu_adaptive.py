"""The GPU adaptive-octree driver (mlod_generate / mlod_generate_from_points) against the oracle's
restatement of OctreeGenerator, with the same projection: the oracle's driver calls the GPU
projection of the same tree for its find_stable step, so the two cell sets must agree bit for bit
(the projection itself is covered by the projection parity tests)."""
import numpy as np
import pytest

from mlodsurfaces_amd import apss as A
from mlodsurfaces_amd import datasets as ds

pytestmark = pytest.mark.gpu


def _gpu_projector(ap):
    def project(c):
        ap.copyPointsToGPU(len(c), c)
        p = np.zeros((len(c), 3), np.float32)
        n = np.zeros((len(c), 3), np.float32)
        ap.project(len(c), p, n, 1, None, np.finfo(np.float32).max)
        return p, n
    return project


def _setup(shape, n, seed, scale=1.0, shift=(0.0, 0.0, 0.0)):
    p, nm = {"sphere": ds.sphere, "torus": ds.torus, "sq": ds.superquadric}[shape](n, seed=seed)
    p = (p * np.float32(scale) + np.asarray(shift, np.float32)).astype(np.float32)
    t = A.APSSOctree().build(p, nm)
    return p, t


@pytest.mark.parametrize("shape,max_depth,scale,shift,semantics,precision,mode", [
    ("sphere", 5, 0.6, (0.1, 0.0, -0.05), "ref_gpu", "fp32", "mlod"),
    ("torus", 6, 0.8, (0.0, 0.0, 0.0), "ref_gpu", "fp32c", "mlod"),
    ("sq", 5, 1.0, (0.05, 0.0, 0.0), "ref_gpu", "fp32", "mlod"),      # reaches outside [-1, 1]^3
    ("sphere", 4, 0.7, (0.0, 0.0, 0.0), "ref_cpu", "fp32", "knn"),
])
def test_generate_matches_oracle(oracle_built, shape, max_depth, scale, shift, semantics, precision, mode):
    print("case", shape, max_depth, semantics, precision, mode, flush=True)
    p, t = _setup(shape, 20000, 61, scale, shift)
    R = t.radius()
    if mode == "knn":
        t.knnMode(True, 16)
    k = A.Kernel(type=2, sigmas=[float(np.float32(0.01 * 2 ** i) * np.float32(R)) for i in range(4)])
    ap = A.Fastapss(t, k)
    ap.set_semantics(semantics)
    ap.set_precision(precision)
    gen = A.OctreeGenerator(ap, k)
    total = []
    leaves = gen.generate(max_depth, total)
    info = gen.info()
    og = oracle_built.OracleGenerator(_gpu_projector(ap))
    o_leaves = og.generate(max_depth)
    oi = og.info()
    assert info["max_rel_depth"] == oi["max_rel_depth"]
    assert info["projected"] == oi["projected"] and info["calls"] == oi["calls"]
    assert np.array_equal(leaves, o_leaves)
    assert np.array_equal(total[0], og.total())
    assert len(leaves) > 100
    # the query points of the upsampling: the leaves' centres
    c = gen.centres()
    oc = np.array([oracle_built.oracle_cell_center(r) for r in o_leaves[:: max(1, len(o_leaves) // 500)]])
    assert np.array_equal(c[:: max(1, len(o_leaves) // 500)], oc)


def test_generate_from_points_matches_oracle(oracle_built):
    rng = np.random.default_rng(62)
    P = rng.uniform(-1.3, 1.3, (200000, 3)).astype(np.float32)
    gen = A.OctreeGenerator(None)
    og = oracle_built.OracleGenerator(None)
    for depth, box in [(8, None), (11, None), (6, [-0.5, -2.0, 0.0, 1.5, 2.0, 1.0])]:
        got = gen.generateFromPoints(depth, P, box)
        exp = og.generate_from_points(P, depth, None if box is None else np.asarray(box, np.float32))
        assert np.array_equal(got, exp), (depth, box)
        c = gen.centres()
        oc = np.array([oracle_built.oracle_cell_center(r, None if box is None else np.asarray(box, np.float32))
                       for r in exp[:300]])
        assert np.array_equal(c[:300], oc)


def test_generator_errors():
    gen = A.OctreeGenerator(None)
    with pytest.raises(A.MlodError):
        gen.generate(4)
    p, t = _setup("sphere", 2000, 63)
    g2 = A.OctreeGenerator(A.Fastapss(t))
    with pytest.raises(A.MlodError):
        g2.generate(4)  # no kernel installed
    with pytest.raises(A.MlodError):
        A.OctreeGenerator(A.Fastapss(t, A.Kernel(type=0, sigma=0.1))).generate(13)


def _keys(ids):
    """node_comparator order key of (depth, x, y, z) rows: depth, then z, y, x (signed)."""
    ids = np.asarray(ids, np.int64)
    b = 1 << 20
    return (ids[:, 0] << 60) | ((ids[:, 3] + b // 2) << 40) | ((ids[:, 2] + b // 2) << 20) | (ids[:, 1] + b // 2)


def _enclosing(p, depth):
    """aabb_octree::enclosing in [-1, 1]^3 (octree.h:956-965), float32 arithmetic."""
    t = ((np.asarray(p, np.float32) - np.float32(-1.0)) / np.float32(2.0)) / np.float32(2.0 ** -depth)
    return np.floor(t).astype(np.int64)


def _centres(ids):
    d = ids[:, :1].astype(np.float32)
    return ((2 * ids[:, 1:] + 1).astype(np.float32) * np.exp2(-d).astype(np.float32) - np.float32(1.0)).astype(np.float32)


@pytest.mark.timeout(600)
def test_generate_depth10_torus1m(oracle_built):
    """The adaptive octree of the 1M-point torus configuration (rational kernel e = 4, cutoff 3,
    compensated form) at max depth 10, on the cloud in the reference's normalised frame
    (Pn::normalize).  Invariants: cells unique and in node_comparator order, depths <= 10, the
    relative depth unchanged (the cloud is inside the root cell), every point's depth-10 cell among
    the cells; one level transition of find_stable_nodesGPU / add_children / expand
    (upsample.cpp:274-358) checked on samples in both directions with the GPU projection; and the
    same driver at depth 6 equal to the oracle's driver cell for cell."""
    w = ds.make_workload("torus1m")
    p = ds.pn_normalize(w["points"])
    t = A.APSSOctree().build(p, w["normals"])
    K = A.Kernel(**w["kernel"])
    ap = A.Fastapss(t, K)
    ap.setScalingProtectionSphere(w["cutoff"])
    ap.set_precision("fp32c")
    gen = A.OctreeGenerator(ap, K)
    total = []
    leaves = gen.generate(10, total)
    total = total[0]
    info = gen.info()
    print(f"torus 1M depth 10: {len(leaves)} cells, {len(total)} crossed, {info['projected']} centres projected in "
          f"{info['calls']} calls", flush=True)
    assert info["max_rel_depth"] == 10 and info["nonfinite"] == 0
    for ids in (leaves, total):
        k = _keys(ids)
        assert np.all(np.diff(k) > 0)  # unique, node_comparator order
        assert ids[:, 0].max() == 10
    assert len(leaves) > 1_000_000
    # covering: the generated cells are a band around the surface, so the depth-10 cell of each
    # input point's one-step projection is a generated cell, or (rarely: the band is built from
    # projections of cell centres, not of these points) a neighbour of one.  The raw points
    # carry normal noise of several depth-10 cells and are only reported.
    proj = _gpu_projector(ap)
    kl = _keys(leaves)
    pp, pn = proj(p)
    ok = np.sum(pn.astype(np.float64) ** 2, 1) >= 0.5
    cell = _enclosing(pp[ok], 10)
    inside = np.isin(_keys(np.concatenate([np.full((len(cell), 1), 10), cell], 1)), kl)
    raw = np.isin(_keys(np.concatenate([np.full((len(p), 1), 10), _enclosing(p, 10)], 1)), kl)
    print(f"projected points: {ok.mean():.6f} valid, {inside.mean():.6f} in generated cells; raw points "
          f"{raw.mean():.4f}", flush=True)
    assert ok.mean() > 0.999 and inside.mean() > 0.999
    miss = cell[~inside]
    near = np.zeros(len(miss), bool)
    for dd in np.stack(np.meshgrid(*[np.arange(-1, 2)] * 3, indexing="ij"), -1).reshape(-1, 3):
        near |= np.isin(_keys(np.concatenate([np.full((len(miss), 1), 10), miss + dd], 1)), kl)
    assert near.all()
    # one level transition, depth 8 -> 9, on samples: S9 = expand_9(children(stable(S8)))
    S8 = total[total[:, 0] == 8]
    S9k = _keys(total[total[:, 0] == 9])
    pp, pn = proj(_centres(S8))
    ok = np.sum(pn.astype(np.float32) ** 2, 1) >= 0.5
    stable = np.unique(_keys(np.concatenate([np.full((int(ok.sum()), 1), 8), _enclosing(pp[ok], 8)], 1)))
    rng = np.random.default_rng(5)
    # (a) children of stable cells and their 26 neighbours are crossed cells at depth 9
    sk = stable[rng.choice(len(stable), min(3000, len(stable)), replace=False)]
    b = 1 << 19
    sx, sy, sz = (sk & 0xFFFFF) - b, ((sk >> 20) & 0xFFFFF) - b, ((sk >> 40) & 0xFFFFF) - b
    d = np.arange(-1, 3)
    gx, gy, gz = np.meshgrid(d, d, d, indexing="ij")
    off = np.stack([gx.ravel(), gy.ravel(), gz.ravel()], 1)  # 4x4x4 block: children and their neighbours
    blk = np.stack([2 * sx, 2 * sy, 2 * sz], 1)[:, None, :] + off[None]
    bk = _keys(np.concatenate([np.full((blk.shape[0] * 64, 1), 9), blk.reshape(-1, 3)], 1))
    assert np.isin(bk, S9k).all()
    # (b) every sampled crossed cell at depth 9 has a stable depth-8 cell whose child is within one cell
    xs = total[total[:, 0] == 9][rng.choice(len(S9k), 3000, replace=False)]
    found = np.zeros(len(xs), bool)
    for dd in off[(off[:, 0] < 2) & (off[:, 1] < 2) & (off[:, 2] < 2)]:
        nb = xs[:, 1:] + dd
        anc = np.floor_divide(nb, 2)
        found |= np.isin(_keys(np.concatenate([np.full((len(xs), 1), 8), anc], 1)), stable)
    assert found.all()
    # the same driver on this cloud at depth 6 against the oracle's driver
    gen6 = A.OctreeGenerator(ap, K)
    tot6 = []
    l6 = gen6.generate(6, tot6)
    og = oracle_built.OracleGenerator(proj)
    assert np.array_equal(l6, og.generate(6)) and np.array_equal(tot6[0], og.total())
