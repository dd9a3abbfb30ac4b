"""Writes the golden vectors in tests/golden/*.npz from the reference's own CPU path.

The reference's sources/apss/CPUUtils.cpp and Kernel.cpp are compiled where they lie by
``make -C oracle ref`` into oracle/_ref/libmlodref.so (this needs /root/reference, so it
runs in the build container only).  This script feeds that library small seeded inputs and
stores the inputs and the reference's outputs (no reference code) as compressed .npz:

  tree_<name>.npz   octree in BFS order (CPUUtils.cpp:540-619): depth, boxes, centres,
                    radii, relations, leaf point lists, areas, FP32 node moments
  proj_<name>.npz   APSSOctree::projectCPU results (CPUUtils.cpp:621-636 / 690-737) for a
                    kernel, cutoff, minDepth, mode and iteration count
  knn_<name>.npz    kd-tree k-nearest neighbour positions (kdTreeNN.h:590-602)

Inputs are chosen free of exact-midpoint ties (the reference resolves those with rand(),
CPUUtils.cpp:299-308); the tie count is checked with the oracle.  Re-run with
    make -C oracle all ref && python tests/golden/make_golden.py [tree names ...]
(names: only those trees and their projections).
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from mlodsurfaces_amd import datasets as ds  # noqa: E402
from oracle import bind  # noqa: E402

SEED = 12345


def clouds():
    out = {}
    p, n = ds.sphere(5000, seed=11, noise=0.002)
    out["sphere5k"] = (p, n)
    p, n = ds.torus(6000, seed=12)
    out["torus6k"] = (p, n)
    p, n = ds.superquadric(6000, seed=13)
    out["sq6k"] = (p, n)
    # duplicated points: leaves at depth 12 with several points
    p, n = ds.sphere(1500, seed=14, noise=0.0)
    out["dups"] = (np.concatenate([p, p[:40], p[:10]]), np.concatenate([n, n[:40], n[:10]]))
    return out


# Trees built with other APSSOctree::build options (CPUUtils.h:176-180): several points per leaf
# at a lower maximum depth, and a fixed root box (updateRootBoundingBox = false).  The fixture
# records the options (max_depth, max_pts, root_box) next to the tree.
def option_trees():
    cl = clouds()
    p, n = cl["sphere5k"]
    out = {"sphere5k_m4d8": (p, n, dict(max_depth=8, max_pts_per_leaf=4))}
    p, n = cl["torus6k"]
    lo, hi = p.min(0).astype(np.float64), p.max(0).astype(np.float64)
    ext = hi - lo
    box = np.concatenate([lo - [0.17, 0.05, 0.31] * ext, hi + [0.29, 0.11, 0.02] * ext]).astype(np.float32)
    out["torus6k_box"] = (p, n, dict(max_pts_per_leaf=2, root_box=box))
    return out


def tree_fixture(name, p, n, opts=None):
    opts = opts or {}
    r = bind.RefTree(p, n, seed=SEED, **opts)
    d = r.nodes()
    d["areas"] = r.areas()
    d["bbox"] = r.bbox()
    d["points"] = p
    d["normals"] = n
    d["max_depth"] = np.int32(opts.get("max_depth", 12))
    d["max_pts"] = np.int32(opts.get("max_pts_per_leaf", 1))
    if "root_box" in opts:
        d["root_box"] = np.asarray(opts["root_box"], np.float32)
    np.savez_compressed(os.path.join(HERE, f"tree_{name}.npz"), **d)
    return r


def main(only=None):
    cases = []
    trees = {name: (p, n, {}) for name, (p, n) in clouds().items()}
    trees.update(option_trees())
    for name, (p, n, opts) in trees.items():
        if only and name not in only:
            continue
        o = bind.OracleTree(p, n, **opts)
        if o.ties:
            raise SystemExit(f"{name}: {o.ties} exact-midpoint ties; pick another seed")
        r = tree_fixture(name, p, n, opts)
        R = ds.bbox_radius(p)
        rng = np.random.default_rng(7)
        qj = (p[rng.choice(len(p), 600, replace=False)] + rng.uniform(-0.01, 0.01, (600, 3))).astype(np.float32)
        mg4 = dict(type=2, sigmas=[0.01 * 2.0 ** i * R for i in range(4)])
        cases.append((name, r, p, qj, "mg4_c3", mg4, dict(cutoff=3.0), 1))
        if name in ("sphere5k", "torus6k"):
            cases.append((name, r, p, qj, "mg4_c3_it3", mg4, dict(cutoff=3.0), 3))
            cases.append((name, r, p, qj, "mg4_c2_md2", mg4, dict(cutoff=2.0, min_depth=2), 1))
            cases.append((name, r, p, qj, "rat4_c3", dict(type=1, exponent=4.0, constant_shift=1e-3),
                          dict(cutoff=3.0), 1))
            cases.append((name, r, p, qj, "gauss_c2", dict(type=0, sigma=0.04 * R, constant_shift=0.0),
                          dict(cutoff=2.0), 1))
            cases.append((name, r, p, qj[:200], "global_mg4", mg4, dict(mode="global"), 1))
            cases.append((name, r, p, qj[:300], "knn20_mg4", mg4, dict(mode="knn", knn=20), 2))
            cases.append((name, r, p, qj[:300], "ball3s_mg4", mg4, dict(mode="ball", ball_radius=-1.0), 2))
        if opts:
            cases.append((name, r, p, qj, "mg4_c3_it3", mg4, dict(cutoff=3.0), 3))
        if name == "sq6k":
            mg6 = dict(type=2, sigmas=[0.01 * 2.0 ** i * R for i in range(6)])
            cases.append((name, r, p, qj, "mg6_c2", mg6, dict(cutoff=2.0), 1))
    for name, r, p, q, tag, kern, prm, iters in cases:
        k = bind.make_kernel(**kern)
        pr = bind.make_params(semantics="ref_cpu", **prm)
        op, on = r.project(q, k, pr, n_iter=iters, step_max=-1.0)
        kd = dict(type=k.type, sigma=k.sigma, exponent=k.exponent, constant_shift=k.constant_shift,
                  sigmas=np.array(list(k.sigmas)[: k.n_sigmas], np.float32))
        np.savez_compressed(os.path.join(HERE, f"proj_{name}_{tag}.npz"), tree=name, queries=q, out_points=op,
                            out_normals=on, n_iter=iters, cutoff=pr.cutoff, min_depth=pr.min_depth, mode=pr.mode,
                            knn=pr.knn, ball_radius=pr.ball_radius, **{f"kernel_{a}": v for a, v in kd.items()})
        print(f"proj_{name}_{tag}: {len(q)} queries")
    if only:
        print("done")
        return
    # kd-tree neighbour sets
    p, n = clouds()["sphere5k"]
    r = bind.RefTree(p, n, seed=SEED)
    q = (p[:200] + 0.003).astype(np.float32)
    np.savez_compressed(os.path.join(HERE, "knn_sphere5k_k20.npz"), queries=q, k=20, neighbours=r.knn(q, 20))
    print("done")


if __name__ == "__main__":
    main(set(sys.argv[1:]) or None)
