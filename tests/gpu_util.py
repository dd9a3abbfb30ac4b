"""Helpers for the GPU parity tests: run the HIP path through the C ABI."""
import numpy as np

from mlodsurfaces_amd import apss as A

PREC = {"fp32": 0, "fp32c": 1, "fp64": 2}


def gpu_tree(points, normals, **kw):
    return A.APSSOctree().build(points, normals, **kw)


def gpu_project(tree, queries, kernel, cutoff=2.0, min_depth=0, semantics="ref_cpu", precision="fp32", n_iter=1,
                mode="mlod", step_max=-1.0, knn=20, ball_radius=0.1, traversal="lane"):
    K = A.Kernel(**kernel) if isinstance(kernel, dict) else kernel
    cls = A.GlobalAPSS if mode == "global" else A.Fastapss
    if mode == "knn":
        tree.knnMode(True, knn)
    elif mode == "ball":
        tree.ballMode(True, ball_radius)
    else:
        tree.knnMode(False, knn)
        tree.ballMode(False, ball_radius)
    ap = cls(tree, K)
    ap.setScalingProtectionSphere(cutoff)
    ap.setMinDepth(min_depth)
    ap.set_semantics(semantics)
    ap.set_precision(precision)
    ap.set_traversal(traversal)
    q = np.ascontiguousarray(queries, np.float32)
    ap.copyPointsToGPU(len(q), q)
    op = np.zeros((len(q), 3), np.float32)
    on = np.zeros((len(q), 3), np.float32)
    ap.project(len(q), op, on, n_iter, K, step_max)
    return op, on, ap


def rel_err(a, b, diag):
    return np.linalg.norm(np.asarray(a, np.float64) - np.asarray(b, np.float64), axis=1) / diag
