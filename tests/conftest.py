"""Shared test plumbing.  Tests marked ``gpu`` need an MI355X (gfx950) and libmlod.so; the rest
run on the CPU (oracle pinning, host logic, C-ABI exports)."""
import glob
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU and the built libmlod.so")


def load_golden(name):
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def golden_names(prefix):
    return sorted(os.path.basename(f) for f in glob.glob(os.path.join(GOLDEN, prefix + "*.npz")))


def golden_kernel(g):
    """Kernel dict stored in a proj_* fixture."""
    return dict(type=int(g["kernel_type"]), sigma=float(g["kernel_sigma"]), exponent=float(g["kernel_exponent"]),
                constant_shift=float(g["kernel_constant_shift"]), sigmas=[float(s) for s in g["kernel_sigmas"]])


MODE_NAMES = {0: "mlod", 1: "global", 2: "knn", 3: "ball"}


def oracle_tree_opts(t):
    """APSSOctree::build options recorded in a tree_* fixture, as OracleTree / RefTree keywords."""
    o = dict(max_depth=int(t.get("max_depth", 12)), max_pts_per_leaf=int(t.get("max_pts", 1)))
    if "root_box" in t:
        o["root_box"] = t["root_box"]
    return o


def gpu_tree_opts(t):
    """The same options as keywords of the library's APSSOctree.build (a fixed root box is set
    on the octree first, as setRootBoundingBox)."""
    o = dict(maxDepth=int(t.get("max_depth", 12)), maxNumberOfPointsPerLeaf=int(t.get("max_pts", 1)))
    if "root_box" in t:
        o["updateRootBoundingBox"] = False
    return o


def gpu_tree_from_fixture(t):
    from mlodsurfaces_amd.apss import APSSOctree
    tr = APSSOctree()
    if "root_box" in t:
        tr.setRootBoundingBox(t["root_box"][:3], t["root_box"][3:])
    return tr.build(t["points"], t["normals"], **gpu_tree_opts(t))


@pytest.fixture(scope="session")
def oracle_built():
    from oracle import bind
    if not os.path.exists(bind.ORACLE_SO):
        import subprocess
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "all"], check=True, capture_output=True)
    return bind
