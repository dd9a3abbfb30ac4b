"""Writes tests/golden/real_<model>.npz from the reference's own data and CPU path.

For three of the reference's sample models (data/Face.ply with Face.params: 8 sigmas, cutoff 1.4,
min depth 1; data/half-igea.ply with half-igea.params: 15 sigmas, cutoff 2; data/gargoyle_0.ply with
gargoyle_0.params: 10 sigmas, cutoff 1.4; the last two kept to seeded 24k-point subsets) it stores the raw points and normals as read from the PLY file, the 134-byte
.params file, seeded queries (input points jittered in the normalised frame) and the results of
the reference's APSSOctree::projectCPU (REF_CPU; oracle/_ref/libmlodref.so, compiled from
sources/apss/CPUUtils.cpp and Kernel.cpp where they lie) on the cloud normalised as
Pn::normalize does, with the kernel of the .params file (Kernel::updateSigmas), after 1 and 3
steps, plus the reference octree's node count.  No reference code is stored.  Re-run with
    make -C oracle all ref && python tests/golden/make_realdata.py [model ...]
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from mlodsurfaces_amd import pointio  # noqa: E402
from oracle import bind  # noqa: E402

DATA = "/root/reference/data"
SEED = 12345
MODELS = {"Face": None, "half-igea": 24000, "gargoyle_0": 24000}


def main(names=None):
    for name, keep in MODELS.items():
        if names and name not in names:
            continue
        P, N = pointio.read_points(os.path.join(DATA, name + ".ply"))
        if keep is not None and keep < len(P):
            idx = np.sort(np.random.default_rng(7).choice(len(P), keep, replace=False))
            P, N = P[idx], N[idx]
        raw = open(os.path.join(DATA, name + ".params"), "rb").read()
        prm = pointio.read_params(raw)
        Pn, frame = pointio.normalize(P)
        ties = bind.OracleTree(Pn, N).ties
        ref = bind.RefTree(Pn, N, seed=SEED)
        radius = ref_radius(ref)
        K = prm.kernel(radius)
        rng = np.random.default_rng(21)
        j = rng.choice(len(Pn), 4096, replace=False)
        q = (Pn[j] + rng.uniform(-0.005, 0.005, (len(j), 3))).astype(np.float32)
        kern = bind.make_kernel(type=K.type, sigma=K.sigma, exponent=K.exponent, constant_shift=K.constant_shift,
                                sigmas=K.sigmas)
        par = bind.make_params(cutoff=prm.cutoff, min_depth=prm.min_depth)
        out = {}
        for it in (1, 3):
            p, n = ref.project(q, kern, par, n_iter=it, nthreads=8)
            out[f"p{it}"], out[f"n{it}"] = p, n
        np.savez_compressed(os.path.join(HERE, f"real_{name}.npz"), points=P, normals=N,
                            params=np.frombuffer(raw, np.uint8), queries=q, ref_nodes=np.int64(ref.num_nodes),
                            ties=np.int64(ties), radius=np.float64(radius), **out)
        print(f"{name}: {len(P)} points, {len(K.sigmas)} sigmas, cutoff {prm.cutoff:g}, min depth {prm.min_depth}, "
              f"{ref.num_nodes} reference nodes, {ties} ties")


def ref_radius(ref):
    """BBOX::radius of the reference octree's root box (CPUUtils.h:84-85), in float32."""
    b = np.asarray(ref.bbox(), np.float32)
    d = (b[:3] - b[3:]).astype(np.float32)
    d2 = np.float32(np.float32(d[0] * d[0]) + np.float32(d[1] * d[1])) + np.float32(d[2] * d[2])
    return float(np.float32(np.sqrt(np.float32(d2))) / np.float32(2.0))


if __name__ == "__main__":
    main(sys.argv[1:])
