"""Pins the oracle to the reference's own CPU path: every golden vector written from
oracle/_ref (the reference's CPUUtils.cpp + Kernel.cpp) must be reproduced."""
import numpy as np
import pytest

from conftest import MODE_NAMES, golden_kernel, golden_names, load_golden, oracle_tree_opts

TREE_KEYS = ["depth", "bb", "center", "radius", "father", "first_child", "next_brother", "num_children", "leaf_count",
             "leaf_offset"]


@pytest.mark.parametrize("name", golden_names("tree_"))
def test_oracle_tree_bit_exact(oracle_built, name):
    g = load_golden(name)
    o = oracle_built.OracleTree(g["points"], g["normals"], **oracle_tree_opts(g))
    assert o.ties == 0
    d = o.nodes()
    for k in TREE_KEYS:
        assert np.array_equal(d[k], g[k]), k
    assert np.array_equal(d["leaf_idx"], g["leaf_idx"])
    assert np.array_equal(o.bbox(), g["bbox"])
    assert np.array_equal(o.areas(), g["areas"])
    mf, md = o.moments()
    assert np.array_equal(mf, g["moments"])
    # FP64 moments of the same tree agree with the FP32 sums to float rounding
    scale = np.abs(md).max(axis=0) + 1e-30
    assert np.all(np.abs(mf - md) <= 1e-5 * scale)


@pytest.mark.parametrize("name", [n for n in golden_names("proj_")])
def test_oracle_projection_matches_reference(oracle_built, name):
    g = load_golden(name)
    t = load_golden(f"tree_{g['tree']}.npz")
    o = oracle_built.OracleTree(t["points"], t["normals"], **oracle_tree_opts(t))
    k = oracle_built.make_kernel(**golden_kernel(g))
    mode = MODE_NAMES[int(g["mode"])]
    prm = oracle_built.make_params(cutoff=float(g["cutoff"]), min_depth=int(g["min_depth"]), semantics="ref_cpu",
                                   mode=mode, knn=int(g["knn"]), ball_radius=float(g["ball_radius"]))
    op, on = o.project(g["queries"], k, prm, n_iter=int(g["n_iter"]), precision=0)
    diag = 2 * float(t["radius"][0])
    if mode in ("mlod", "global"):
        # same tree, same float arithmetic, same summation order: bit-exact
        assert np.array_equal(op.astype(np.float32), g["out_points"])
        assert np.array_equal(on.astype(np.float32), g["out_normals"])
    else:
        # kNN / ball: same neighbour sets, summed in index order instead of heap order;
        # degenerate fits leave NaN positions (aliased call, CPUUtils.cpp:735) in both
        bad = np.isnan(g["out_points"]).any(1)
        assert np.array_equal(np.isnan(op).any(1), bad)
        assert np.all(on[bad] == 0) and np.all(g["out_normals"][bad] == 0)
        err = np.linalg.norm(op[~bad] - g["out_points"][~bad], axis=1) / diag
        assert np.median(err) < 1e-7 and err.max() < 1e-5, err.max()
    # the FP64 oracle stays within the reference-CPU gate of the FP32 reference
    op64, _ = o.project(g["queries"], k, prm, n_iter=int(g["n_iter"]), precision=1)
    ok = ~np.isnan(g["out_points"]).any(1)
    if mode == "ball":
        # a ball holding 1-3 points is a degenerate fit whose FP32 result is rounding noise
        P = t["points"].astype(np.float64)
        r = 3 * float(g["kernel_sigmas"][0])
        cnt = np.array([(np.sum((P - qq) ** 2, 1) <= r * r).sum() for qq in g["out_points"].astype(np.float64)])
        ok &= cnt >= 4
    err = np.linalg.norm(op64[ok] - g["out_points"][ok], axis=1) / diag
    assert err.max() < 1e-5, err.max()


def test_oracle_knn_sets_match_kdtree(oracle_built):
    """Brute-force kNN of the oracle = the reference kd-tree's neighbour sets."""
    g = load_golden("knn_sphere5k_k20.npz")
    t = load_golden("tree_sphere5k.npz")
    P = t["points"].astype(np.float32)
    for qi, q in enumerate(g["queries"]):
        d = P - q[None, :]
        d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        mine = P[np.argsort(d2.astype(np.float64), kind="stable")[: int(g["k"])]]
        ref = g["neighbours"][qi]
        key = lambda a: np.lexsort(a.T[::-1])  # noqa: E731
        assert np.array_equal(mine[key(mine)], ref[key(ref)])
