#!/usr/bin/env python3
"""Projected points/s of the MLoD projection hot path on one MI355X.

Workload (default): the 8M-point precision form — a superquadric (exponents 0.4/0.4, axes
1/0.7/0.5) with multi-octave sinusoidal displacement (amplitude 0.01) and Gaussian noise
(sigma 0.001), 8M off-surface queries at uniform offsets +-0.02 diag, multi-Gaussian kernel
with 6 sigmas (sigma_1 = 0.01 R, factor 2, scaled by the bbox radius R), cutoff 2,
minDepth 0, octree depth <= 12 with one point per leaf.

A "step" is one projection step of all M queries (APSS::project(M, ..., 1, ...), the
reference harness's unit, scene.cpp:929-933): one kernel launch on device-resident queries,
projected in place.  value = M * steps / time (projected points/s), timed with HIP events on
the launch stream.  The same K steps are run in the plain FP32 and plain FP64 forms; the
FP64 run is the error reference (max / RMS |dp| / diag).  The CPU baseline is the
reference's own OpenMP CPU path (its CPUUtils.cpp compiled into oracle/_ref) on a bounded
query sample on this host.

    python bench.py [--gpus 1] [--steps 10] [--warmup 2] [--scale 1.0]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3  # MI355X published FP32 vector peak
PEAK_HBM_GBS = 8000.0
SERVE_BATCH, SERVE_CHUNK, SERVE_INFLIGHT = 131072, 32768, 3


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def weight_cost(kern):
    """(flops, transcendentals) of one weight evaluation."""
    t = kern["type"]
    if t == 2:
        s = len(kern["sigmas"])
        return 3 * s, s
    if t == 0:
        return 3, 1
    return 6, 1


def algorithmic_work(counts, m, kern):
    """FP32 flops (FMA = 2) and transcendentals of one step over m queries, from the
    traversal counters: each far node, leaf point and near node's own share costs the
    offset and squared distance (8), the weight, f = m w (1) and 9 moment FMAs (18); each
    tested child its distance (8), sqrt, far test (2) and blend factor (10 flops, 2
    transcendentals); each query the sphere fit (70)."""
    wf, wt = weight_cost(kern)
    contrib = counts["far"] + counts["leaf_points"] + counts["near"]
    flops = contrib * (27 + wf) + counts["tested"] * 20 + 70 * m
    trans = contrib * wt + counts["tested"] * 3
    return float(flops), float(trans)


def kd_tree(w):
    """A k-d tree over the workload's points (built once per workload), for ball statistics."""
    if "_kd" not in w:
        from scipy.spatial import cKDTree
        w["_kd"] = cKDTree(w["points"].astype(np.float64))
    return w["_kd"]


def ball_population(w, sample=4000):
    """Mean number of points in a query's ball (radius 3 sigma_0), from a seeded query sample."""
    r = 3.0 * w["kernel"]["sigmas"][0] if w["ball_radius"] < 0 else w["ball_radius"]
    q = w["queries"][np.random.default_rng(1).choice(len(w["queries"]), min(sample, len(w["queries"])), replace=False)]
    return float(np.mean(kd_tree(w).query_ball_point(q.astype(np.float64), r, return_length=True)))


def cpu_baseline(w, kern, cutoff, budget_s=15.0):
    """The reference's CPU projection (APSSOctree::projectCPU with OpenMP over queries,
    scene.cpp:982-986), REF_CPU semantics, one step per query, on a bounded sample."""
    from oracle import bind
    if not bind.have_ref():
        return None
    release = os.path.exists(bind.REF_SO_RELEASE)
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    t0 = time.time()
    log(f"cpu baseline: building the reference octree over {len(w['points'])} points")
    tree = bind.RefTree(w["points"], w["normals"], release=release)
    build_s = time.time() - t0
    log(f"cpu baseline: reference octree built in {build_s:.1f} s")
    k = bind.make_kernel(**kern)
    prm = bind.make_params(cutoff=cutoff, semantics="ref_cpu", mode=w.get("mode", "mlod"), knn=w.get("knn", 20),
                           ball_radius=w.get("ball_radius", 0.1))
    q = w["queries"]
    n0 = min(len(q), 4000)
    t0 = time.time()
    tree.project(q[:n0], k, prm, n_iter=1, nthreads=cores)
    rate = n0 / max(time.time() - t0, 1e-6)
    log(f"cpu baseline: calibration sample {n0} queries at {rate:.0f} /s")
    ns = int(min(len(q), max(n0, rate * budget_s)))
    t0 = time.time()
    tree.project(q[:ns], k, prm, n_iter=1, nthreads=cores)
    dt = time.time() - t0
    return {"value": ns / dt, "unit": "projected pts/s", "cores": cores, "kind": "reference",
            "sample": f"first {ns} of {len(q)} queries, 1 step each, reference APSSOctree::projectCPU (REF_CPU, "
                      f"{w.get('mode', 'mlod')}) with OpenMP on {cores} host threads, built "
                      f"{'with the reference release flags (-O3, x86-64-v4)' if release else '-O2'}; reference octree "
                      f"build {build_s:.1f} s untimed"}


def adaptive_measure(A, ds, w, depth, semantics, precision):
    """OctreeGenerator::generate at `depth` on the workload's cloud in the reference's normalised
    frame (Pn::normalize), projecting with the workload's kernel: wall time, cells, projections."""
    from mlodsurfaces_amd import hiprt
    p = ds.pn_normalize(w["points"])
    t = A.APSSOctree().build(p, w["normals"])
    k = A.Kernel(**w["kernel"])
    ap = A.Fastapss(t, k)
    ap.setScalingProtectionSphere(w["cutoff"])
    ap.set_semantics(semantics)
    ap.set_precision(precision)
    gen = A.OctreeGenerator(ap, k)
    gen.generate(max(1, depth - 2))  # warm-up
    hiprt.device_synchronize()
    t0 = time.perf_counter()
    total = []
    leaves = gen.generate(depth, total)
    dt = time.perf_counter() - t0
    inf = gen.info()
    log(f"adaptive octree depth {depth}: {len(leaves)} cells, {len(total[0])} crossed, {inf['calls']} projection "
        f"calls of {inf['projected']} centres, {dt * 1e3:.1f} ms")
    return {"max_depth": depth, "max_relative_depth": inf["max_rel_depth"], "cells": len(leaves),
            "crossed_cells": len(total[0]), "projected_centres": inf["projected"], "projection_calls": inf["calls"],
            "ms": dt * 1e3}


def mesh_measure(ap, depth, label):
    """scene::computeDualContouring's sequence (dc.surface_mesh: adaptive octree, uniformize, corner
    field, extraction, 3-step vertex projection, two rounds of smoothing, subdivision and projection)
    at `depth` with `ap`'s tree and settings, after one untimed run; wall times of its stages, each
    ending at a host synchronisation."""
    from mlodsurfaces_amd import dc, hiprt
    dc.surface_mesh(ap, depth)
    hiprt.device_synchronize()
    t0 = time.perf_counter()
    V, N, F, info = dc.surface_mesh(ap, depth)
    ms = (time.perf_counter() - t0) * 1e3
    out = {"depth": depth, "ms": ms, **{k: (round(v, 3) if isinstance(v, float) else v) for k, v in info.items()}}
    log(f"surface mesh {label} depth {depth}: {info['cells']} crossed cells -> {info['nodes']} nodes, "
        f"{info['corners']} corners, {info['dc_triangles']} triangles -> {info['triangles']} refined, {ms:.1f} ms "
        f"(generate {info['ms_generate']:.1f}, uniformize {info['ms_uniformize']:.1f}, corners "
        f"{info['ms_corners']:.1f}, extract {info['ms_extract']:.1f}, refine {info['ms_refine']:.1f})")
    return out


def dc_cpu_baseline(A, ap, depth, gpu):
    """CPU baseline of the dual-contouring stages (kind "port": the oracle's sequential restatement
    of the reference's uniformize, minimal-edge extraction and mesh passes, one thread), on the same
    crossed cells and the GPU's corner values; the reference's own code needs Eigen and is not built."""
    from mlodsurfaces_amd.dc import DualContouring
    from oracle import bind
    total = []
    gen = A.OctreeGenerator(ap, None)
    gen.generate(depth, total)
    gen.close()
    cells = total[0]
    g = DualContouring(cells, depth)
    ap._sync_params()
    g.project_corners(ap.m_apssOctree)
    xyz, _ = g.corners()
    f = g.corner_values()
    g.close()
    t0 = time.perf_counter()
    o = bind.OracleDC(cells, depth)
    t1 = time.perf_counter()
    o.set_corners(xyz, f)
    v, vn, tri = o.extract()
    t2 = time.perf_counter()
    for _ in range(2):
        v = bind.oracle_mesh_smooth(v, tri)
        v, tri = bind.oracle_mesh_subdivide(v, tri)
    t3 = time.perf_counter()
    out = {"kind": "port", "cores": 1, "uniformize_s": t1 - t0, "extract_s": t2 - t1, "mesh_passes_s": t3 - t2,
           "triangles": int(len(tri)),
           "sample": f"the oracle's uniformize, extraction and two smoothing + subdivision rounds (no projection) "
                     f"on the same {len(cells)} cells and corner values, single-threaded"}
    log(f"surface mesh CPU port: uniformize {out['uniformize_s']:.2f} s (GPU {gpu['ms_uniformize']:.1f} ms), "
        f"extract {out['extract_s']:.2f} s (GPU {gpu['ms_extract']:.1f} ms), smoothing + subdivision "
        f"{out['mesh_passes_s']:.2f} s")
    return out


def workload_mesh_measure(A, ds, w, depth, semantics, precision, cpu=True):
    """The surface mesh of the workload's cloud in the reference's normalised frame."""
    p = ds.pn_normalize(w["points"])
    t = A.APSSOctree().build(p, w["normals"])
    k = A.Kernel(**w["kernel"])
    ap = A.Fastapss(t, k)
    ap.setScalingProtectionSphere(w["cutoff"])
    ap.set_semantics(semantics)
    ap.set_precision(precision)
    out = mesh_measure(ap, depth, "workload")
    if cpu:
        try:
            out["cpu_baseline"] = dc_cpu_baseline(A, ap, depth, out)
        except Exception as ex:  # the baseline is reported, never the measured product
            log(f"surface mesh cpu baseline failed: {ex}")
    return out


def real_models_measure(A):
    """The reference's own sample models (Face; a 24k-point subset of half-igea) with their own
    .params settings: each model's points projected maxNbIterations steps (FP32C, REF_GPU) and
    timed with HIP events; error against the FP64 form; and, with REF_CPU semantics, the distance
    to the reference's projectCPU results stored with the model (4096 queries, 3 steps)."""
    from mlodsurfaces_amd import hiprt, pointio
    out = {}
    for name in ("Face", "half-igea", "gargoyle_0"):
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden", f"real_{name}.npz")
        if not os.path.exists(path):
            continue
        with np.load(path, allow_pickle=False) as z:
            g = {k: z[k] for k in z.files}
        prm = pointio.read_params(g["params"].tobytes())
        Pn, _ = pointio.normalize(g["points"])
        tree = A.APSSOctree().build(Pn, g["normals"])
        K = prm.kernel(tree.radius())
        diag = 2.0 * tree.radius()
        steps = max(prm.max_iterations, 1)
        res = {}
        for prec in ("fp32c", "fp64"):
            ap = A.Fastapss(tree, K)
            ap.setScalingProtectionSphere(prm.cutoff)
            ap.setMinDepth(prm.min_depth)
            ap.set_precision(prec)
            ap.copyPointsToGPU(len(Pn), Pn)
            ap.project(len(Pn), None, None, 1, K)  # warm-up step
            ap.copyPointsToGPU(len(Pn), Pn)
            s = hiprt.Stream()
            e0, e1 = hiprt.Event(), hiprt.Event()
            hiprt.device_synchronize()
            e0.record(s)
            ap.project(len(Pn), None, None, steps, K, stream=s.ptr)
            e1.record(s)
            e1.synchronize()
            op, _ = ap.download(len(Pn))
            res[prec] = (op, e0.elapsed_ms(e1))
        err = np.linalg.norm(res["fp32c"][0].astype(np.float64) - res["fp64"][0], axis=1) / diag
        ap = A.Fastapss(tree, K)
        ap.setScalingProtectionSphere(prm.cutoff)
        ap.setMinDepth(prm.min_depth)
        ap.set_semantics("ref_cpu")
        ap.set_precision("fp32c")
        ap.copyPointsToGPU(len(g["queries"]), g["queries"])
        rp, _ = ap.project(len(g["queries"]), np.zeros_like(g["queries"]), np.zeros_like(g["queries"]), 3, K)
        rerr = np.linalg.norm(rp.astype(np.float64) - g["p3"], axis=1) / diag
        ap.set_semantics("ref_gpu")
        mesh = mesh_measure(ap, prm.max_depth, name)
        ms = res["fp32c"][1]
        out[name] = {"points": len(Pn), "sigmas": len(K.sigmas), "cutoff": prm.cutoff, "min_depth": prm.min_depth,
                     "steps": steps, "ms": ms, "pts_per_s": len(Pn) * steps / (ms * 1e-3),
                     "fp64_ms": res["fp64"][1], "fp32c_vs_fp64_max": float(err.max()),
                     "ref_cpu_vs_reference_projectCPU_max": float(rerr.max()), "surface_mesh": mesh}
        log(f"real model {name}: {len(Pn)} points, {len(K.sigmas)} sigmas, {steps} steps in {ms:.2f} ms "
            f"({out[name]['pts_per_s'] / 1e6:.1f} M pts/s; FP64 {res['fp64'][1]:.2f} ms), FP32C vs FP64 max "
            f"{err.max():.2e}, REF_CPU FP32C vs the reference's projectCPU max {rerr.max():.2e} x diag")
    return out


def lib_handle():
    from mlodsurfaces_amd import _lib
    return _lib.lib()


def calibrate_copy():
    """A kernel of known traffic for rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE: out = 2 * in over
    128 Mi floats, i.e. 512 MiB read and 512 MiB written per dispatch, three dispatches."""
    import torch
    n = 128 << 20
    a = torch.ones(n, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        torch.mul(a, 2.0, out=b)
    torch.cuda.synchronize()
    log(f"copy calibration: 3 dispatches, {n * 4} bytes read and written each")


def kernel_source_sha():
    """SHA-256 (first 16 hex digits) of the hot kernel's source, project.hip."""
    import hashlib
    with open(os.path.join(ROOT, "mlodsurfaces_amd", "csrc", "project.hip"), "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()[:16]


def traffic_from_profile(kernel_name, workload):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3 PMC summary
    (FETCH_SIZE and WRITE_SIZE in KiB per dispatch, each scaled by the factor a plain copy
    kernel of known size reads for the same counter) and whether that summary was measured on
    the current kernel source (its `kernel_source_sha16:` line), or (None, False)."""
    path = os.path.join(ROOT, "profiles", f"{kernel_name}_{workload}.md")
    if not os.path.exists(path):
        return None, False
    traffic, sha = None, None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            traffic = float(line.split(":", 1)[1].split()[0])
        if line.startswith("kernel_source_sha16:"):
            sha = line.split(":", 1)[1].strip()
    return traffic, sha == kernel_source_sha()


def serve_measure(n_clouds=64, seed_rate=2000.0):
    """Serving (64 resident clouds, Poisson requests of 1k-500k queries, mixed kernels, 10
    steps each): capacity with all requests queued at once, then p50/p99 latency at the offered
    2000 req/s and at rates the server sustains (30%, 50% and 70% of its measured capacity)."""
    from mlodsurfaces_amd import serve as SV
    from mlodsurfaces_amd.apss import APSSOctree
    t0 = time.time()
    clouds = SV.serving_workload(n_clouds)
    trees = [APSSOctree().build(p, n) for p, n in clouds]
    npts = sum(len(c[0]) for c in clouds)
    log(f"serve: {len(clouds)} clouds, {npts / 1e6:.1f}M points resident, ready in {time.time() - t0:.1f}s")
    srv = SV.Server(max_batch=SERVE_BATCH, chunk=SERVE_CHUNK, inflight=SERVE_INFLIGHT)
    # capacity: a Poisson request mix submitted at once (offered rate far above service)
    cap, _, _ = SV.run_load(srv, trees, clouds, rate=1e6, seconds=1.0, max_requests=400, seed=11, log=log)
    mean_pts = cap["projected_pts"] / max(cap["requests"], 1)
    cap_rate = cap["pts_per_s"] / mean_pts
    spec, _, _ = SV.run_load(srv, trees, clouds, rate=seed_rate, seconds=0.5, seed=12, log=log)
    keys = ("requests", "offered_rate", "pts_per_s", "p50_ms", "p99_ms", "p50_100k_ms", "p99_100k_ms", "n_100k", "wall_s")
    out = {"clouds": len(clouds), "resident_points": npts, "max_batch": SERVE_BATCH, "chunk": SERVE_CHUNK,
           "inflight": SERVE_INFLIGHT,
           "capacity_pts_per_s": cap["pts_per_s"], "capacity_req_per_s": cap_rate,
           "at_2000_req_s": {k: spec[k] for k in keys}}
    # sustained load at fractions of the measured capacity (Poisson arrivals for 6 s each)
    under20 = None
    for frac in (0.3, 0.5, 0.7):
        sus, _, _ = SV.run_load(srv, trees, clouds, rate=frac * cap_rate, seconds=6.0, seed=13, log=log)
        out[f"at_{int(frac * 100)}pct_capacity"] = {k: sus[k] for k in keys}
        if sus["p99_100k_ms"] is not None and sus["p99_100k_ms"] < 20.0:
            under20 = frac
    srv.close()
    out["p99_100k_below_20ms_up_to_load"] = under20
    out["note"] = ("latency = submit -> results in the caller's host buffers; 100k = requests of 80k-125k queries; "
                   "loads are fractions of the measured capacity")
    return out


EXTRA_CONFIGS = {
    # the other configurations' measured lines: steps, warm-up steps
    "torus1m": (10, 2),   # 1M torus, rational kernel, with the adaptive octree at depth 10
    "c5_knn": (3, 1),     # 20M points, kNN APSS k = 32, DC-grid queries at depth 11 (16M)
    "c5_ball": (2, 1),    # the same with ball APSS r = 3 sigma_0
}


def run_extra_configs(names, timeout_s=420):
    """Each named configuration measured by this script in a child process (started before this
    process uses the GPU), its progress lines passed through; returns {name: summary}."""
    import subprocess
    out = {}
    env = dict(os.environ, MLOD_DATA_CACHE=os.environ.get("MLOD_DATA_CACHE", "/tmp/mlod_data_cache"))
    for name in names:
        steps, warm = EXTRA_CONFIGS[name]
        cmd = [sys.executable, "-u", os.path.abspath(__file__), "--workload", name, "--steps", str(steps), "--warmup",
               str(warm), "--no-cpu-baseline", "--no-serve", "--no-real-models", "--extra-configs", "none"]
        log(f"other configuration {name}: {' '.join(cmd[2:])}")
        t0 = time.time()
        try:
            p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=None, text=True, timeout=timeout_s, env=env)
        except subprocess.TimeoutExpired:
            out[name] = {"status": f"timed out after {timeout_s} s"}
            continue
        lines = [ln for ln in p.stdout.splitlines() if ln.startswith("{")]
        if p.returncode != 0 or not lines:
            out[name] = {"status": f"exit {p.returncode}"}
            continue
        r = json.loads(lines[-1])
        keep = ("value", "unit", "ms_per_step", "steps", "config", "forms", "error_vs_fp64", "counters_per_step",
                "counters_later_steps", "roofline", "adaptive_octree")
        out[name] = {k: r.get(k) for k in keep}
        out[name]["wall_s"] = time.time() - t0
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="superquadric8m",
                    choices=["superquadric8m", "sphere100k", "torus1m", "c5_knn", "c5_ball"])
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--precision", default="fp32c", choices=["fp32", "fp32c", "fp64"])
    ap.add_argument("--semantics", default="ref_gpu", choices=["ref_gpu", "ref_cpu"])
    ap.add_argument("--traversal", default="lane", choices=["lane", "wave"],
                    help="MLoD walk: one per query (default) or one per wavefront over the union of its queries'")
    ap.add_argument("--no-forms", action="store_true", help="skip the plain FP32 / FP64 comparison runs")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-serve", action="store_true", help="skip the serving measurement")
    ap.add_argument("--serve-clouds", type=int, default=64)
    ap.add_argument("--adaptive-depth", type=int, default=10,
                    help="max depth of the adaptive octree timed with the torus1m workload")
    ap.add_argument("--c5-queries", default="dc-grid", choices=["dc-grid", "point-cells"],
                    help="C5 queries: the dual-contouring grid's finest cell centres, or the depth-11 cells "
                         "that contain a point")
    ap.add_argument("--no-real-models", action="store_true", help="skip the reference's sample models")
    ap.add_argument("--mesh-depth", type=int, default=8,
                    help="depth of the surface mesh of the workload's cloud (0: none)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain elementwise copy of 512 MiB three times (to scale the HBM counters)")
    ap.add_argument("--extra-configs", default=",".join(EXTRA_CONFIGS),
                    help="with the headline workload, the other configurations measured in child processes "
                         "(comma-separated names, or 'none')")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: measures one GPU only.  The projection path would shard by "
              "replicating the octree and splitting the Morton-ordered queries across GPUs; that is not "
              "built here.  Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    other = None
    extra = [x for x in args.extra_configs.split(",") if x and x != "none"]
    if extra and args.workload == "superquadric8m" and args.scale == 1.0:
        unknown = [x for x in extra if x not in EXTRA_CONFIGS]
        if unknown:
            raise SystemExit(f"unknown --extra-configs {unknown}")
        other = run_extra_configs(extra)

    from mlodsurfaces_amd import apss as A
    from mlodsurfaces_amd import datasets as ds
    from mlodsurfaces_amd import hiprt

    t0 = time.time()
    w = ds.make_workload(args.workload, scale=args.scale)
    P, N, Q = w["points"], w["normals"], w["queries"]
    n, m = len(P), len(Q)
    lo, hi = P.min(0).astype(np.float64), P.max(0).astype(np.float64)
    diag = float(np.linalg.norm(hi - lo))
    if "dc_grid" in w and args.c5_queries == "point-cells":
        w["description"] += f", {m} depth-11 cells containing a point"
    log(f"workload {w['description']} generated in {time.time() - t0:.1f}s")

    stream = hiprt.Stream()
    dP, dN, dQ = hiprt.DeviceBuffer(P), hiprt.DeviceBuffer(N), hiprt.DeviceBuffer(Q)
    hiprt.device_synchronize()
    tb0, tb1 = hiprt.Event(), hiprt.Event()
    tb0.record(stream)
    tree = A.APSSOctree().build_device(dP.ptr.value, dN.ptr.value, n, stream=stream.ptr)
    tb1.record(stream)
    tb1.synchronize()
    build_ms = tb0.elapsed_ms(tb1)
    info = tree.info()
    log(f"octree: {info['n_nodes']} nodes, depth {info['depth']}, ties {info['ties']}, build {build_ms:.1f} ms, "
        f"{info['device_bytes'] / 2**20:.0f} MiB")

    K = A.Kernel(**w["kernel"])
    mode = w.get("mode", "mlod")
    if "dc_grid" in w and args.c5_queries == "dc-grid":
        # the queries: the dual-contouring grid's finest cell centres, from the adaptive octree
        # generated with the MLoD projection of the same tree (before the kNN / ball mode is set)
        from mlodsurfaces_amd import dc
        g = w["dc_grid"]
        ap_ = A.Fastapss(tree, K)
        ap_.setScalingProtectionSphere(w["cutoff"])
        ap_.set_semantics(args.semantics)
        ap_.set_precision(args.precision)
        tg = time.perf_counter()
        Q = dc.grid_centres(ap_, g["depth"], g["cap"], g["seed"])
        del ap_
        w["queries"] = Q
        m = len(Q)
        dQ = hiprt.DeviceBuffer(Q)
        hiprt.device_synchronize()
        w["description"] += f", {m} DC-grid cell centres (adaptive octree depth {g['depth']}, uniformized)"
        log(f"DC-grid queries: {m} cell centres in {time.perf_counter() - tg:.1f} s")
    if mode == "knn":
        tree.knnMode(True, w["knn"])
    elif mode == "ball":
        tree.ballMode(True, w["ball_radius"])
    forms = [args.precision] + ([] if args.no_forms else [f for f in ("fp32", "fp64") if f != args.precision])
    if "fp64" in forms:
        forms = [f for f in forms if f != "fp64"] + ["fp64"]
    results = {}
    counts = None
    counts_later = None
    for form in forms:
        ap_ = A.Fastapss(tree, K)
        ap_.setScalingProtectionSphere(w["cutoff"])
        ap_.setMinDepth(0)
        ap_.set_semantics(args.semantics)
        ap_.set_precision(form)
        ap_.set_traversal(args.traversal)
        ap_.copyPointsToGPU_device(dQ.ptr.value, m, stream=stream.ptr)
        if counts is None and mode == "mlod":
            counts = ap_.count_visits(stream=stream.ptr)
            counts["lane_efficiency"] = counts["iterations"] / max(counts["lane_slots"], 1)
            log(f"counters (one step): {counts}")
        slow = False
        for wi in range(args.warmup):
            w0 = time.perf_counter()
            ap_.project(m, None, None, 1, K, stream=stream.ptr)
            stream.synchronize()
            if time.perf_counter() - w0 > 2.0:
                slow = True  # long steps: report each one as it ends
                log(f"{form}: warm-up step {wi + 1} took {time.perf_counter() - w0:.1f} s")
        ap_.copyPointsToGPU_device(dQ.ptr.value, m, stream=stream.ptr)
        stream.synchronize()
        hiprt.device_synchronize()
        ev = [hiprt.Event() for _ in range(args.steps + 1)]
        wall0 = time.perf_counter()
        ev[0].record(stream)
        for s in range(args.steps):
            ap_.project(m, None, None, 1, K, stream=stream.ptr)
            ev[s + 1].record(stream)
            if slow:
                ev[s + 1].synchronize()
                log(f"{form}: step {s + 1} done")
        ev[-1].synchronize()
        hiprt.device_synchronize()
        wall = time.perf_counter() - wall0
        launch_ms = [ev[s].elapsed_ms(ev[s + 1]) for s in range(args.steps)]
        total_ms = ev[0].elapsed_ms(ev[-1])
        if form == args.precision and mode == "mlod":
            # the same counters on the positions the timed steps reached: the walks of the steps after
            # the first (queries on the surface) are longer and more alike than those of the first
            # step from the offset positions; the counter pass leaves the positions as they are
            counts_later = ap_.count_visits(stream=stream.ptr)
            counts_later["lane_efficiency"] = counts_later["iterations"] / max(counts_later["lane_slots"], 1)
            log(f"counters (a step after the timed steps): {counts_later}")
        pts, nrm = ap_.download(m, stream=stream.ptr)
        kinds = ap_.fit_kinds(m, stream=stream.ptr) if hasattr(lib_handle(), "mlod_queries_fit_kinds") else None
        results[form] = dict(total_ms=total_ms, wall_s=wall, launch_ms=launch_ms, pts=pts, nrm=nrm, kinds=kinds,
                             pts_per_s=m * args.steps / (total_ms / 1e3))
        log(f"{form}: {args.steps} steps in {total_ms:.1f} ms (wall {wall * 1e3:.1f} ms), "
            f"{results[form]['pts_per_s'] / 1e6:.1f} M projected pts/s, mean launch {np.mean(launch_ms):.2f} ms")

    ref64 = results.get("fp64")
    errs = {}
    for form, r in results.items():
        if ref64 is not None and form != "fp64":
            e = np.linalg.norm(r["pts"].astype(np.float64) - ref64["pts"].astype(np.float64), axis=1) / diag
            ok = np.isfinite(e)
            errs[form] = dict(max=float(e[ok].max()), rms=float(np.sqrt(np.mean(e[ok] ** 2))), nonfinite=int((~ok).sum()),
                              p9999=float(np.quantile(e[ok], 0.9999)), above_1e6=int(np.sum(e[ok] > 1e-6)))
            if r["kinds"] is not None and ref64["kinds"] is not None:
                # plane / sphere switch flips of the last step against the FP64 form
                errs[form]["plane_sphere_flips"] = int(np.sum((r["kinds"] == 1) != (ref64["kinds"] == 1)))
                errs[form]["plane_fits"] = int(np.sum(r["kinds"] == 1))
            if mode == "ball":
                # a ball holding 1-3 points gives a degenerate fit whose result is rounding noise in
                # every arithmetic (as in the reference): the outliers are classified by their ball
                # population (start positions; exact counts from a k-d tree over the points)
                out_i = np.where(~ok | (np.nan_to_num(e, nan=1.0) > 1e-6))[0]
                if len(out_i):
                    rb = 3.0 * w["kernel"]["sigmas"][0] if w["ball_radius"] < 0 else w["ball_radius"]
                    # fewer than 4 points in the ball <=> the 4th nearest point lies beyond r
                    d4, _ = kd_tree(w).query(Q[out_i].astype(np.float64), k=4, distance_upper_bound=rb)
                    deg = ~np.isfinite(d4[:, 3])
                    keep = np.ones(len(e), bool)
                    keep[out_i[deg]] = False
                    e2 = e[keep & ok]
                    errs[form].update(degenerate_balls=int(deg.sum()), max_nondegenerate=float(e2.max()),
                                      above_1e6_nondegenerate=int(np.sum(e2 > 1e-6)),
                                      nonfinite_nondegenerate=int(np.sum(~ok & keep)))
            log(f"{form} vs fp64 after {args.steps} steps: max {errs[form]['max']:.3e} rms {errs[form]['rms']:.3e} "
                f"p99.99 {errs[form]['p9999']:.3e} diag, {errs[form]['above_1e6']} queries above 1e-6 diag, "
                f"{errs[form].get('plane_sphere_flips', '-')} plane/sphere flips"
                + (f"; outliers in balls of < 4 points: {errs[form]['degenerate_balls']}, otherwise max "
                   f"{errs[form]['max_nondegenerate']:.3e}, {errs[form]['above_1e6_nondegenerate']} above 1e-6, "
                   f"{errs[form]['nonfinite_nondegenerate']} non-finite" if "degenerate_balls" in errs[form] else ""))

    head = results[args.precision]
    mean_launch_s = float(np.mean(head["launch_ms"])) / 1e3
    if mode == "mlod":
        # the first timed step walks from the uploaded positions, the others from projected ones:
        # the work per launch is the step-weighted mean of the two counter passes
        flops, trans = algorithmic_work(counts, m, w["kernel"])
        if counts_later is not None and args.steps > 1:
            f2, t2 = algorithmic_work(counts_later, m, w["kernel"])
            flops = (flops + (args.steps - 1) * f2) / args.steps
            trans = (trans + (args.steps - 1) * t2) / args.steps
    else:
        nb = float(w["knn"]) if mode == "knn" else ball_population(w)
        counts = {"neighbours_per_query": nb}
        wf, wt = weight_cost(w["kernel"])
        flops, trans = float(m) * (nb * (27 + wf) + 70), float(m) * nb * wt
    achieved = (flops + 8.0 * trans) / mean_launch_s / 1e12
    comp_bytes = 64.0 * info["n_nodes"] + 32.0 * n + 48.0 * m
    roofline = {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None,
                "algorithmic_flops_per_launch": flops, "transcendentals_per_launch": trans,
                "compulsory_bytes_per_launch": comp_bytes,
                "compulsory_hbm_frac": comp_bytes / mean_launch_s / 1e9 / PEAK_HBM_GBS,
                "mean_launch_ms": mean_launch_s * 1e3}
    if mode != "mlod":
        roofline["note"] = ("algorithmic work counts the fit over the neighbours only; the neighbour search, "
                            "which dominates, is not counted (see the kernel's counter profile)")
    if args.workload == "superquadric8m" and args.scale == 1.0 and args.precision == "fp32c" and \
            args.semantics == "ref_gpu" and args.traversal == "lane":
        roofline["traffic"], roofline["traffic_profile_current"] = traffic_from_profile("k_project", args.workload)
    adaptive = None
    if args.workload == "torus1m":
        adaptive = adaptive_measure(A, ds, w, args.adaptive_depth, args.semantics, args.precision)
    del tree, dP, dN, dQ
    hiprt.device_synchronize()
    mesh = None
    if args.mesh_depth > 0 and args.workload == "superquadric8m" and args.scale == 1.0:
        mesh = workload_mesh_measure(A, ds, w, args.mesh_depth, args.semantics, args.precision,
                                     cpu=not args.no_cpu_baseline)
    real = None
    if args.workload == "superquadric8m" and args.scale == 1.0 and not args.no_real_models:
        real = real_models_measure(A)
    serving = None
    if not args.no_serve and args.workload == "superquadric8m" and args.scale == 1.0:
        try:
            serving = serve_measure(args.serve_clouds)
        except Exception as ex:
            log(f"serving measurement failed: {ex}")
            raise
    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(w, w["kernel"], w["cutoff"], args.cpu_budget)
            if cpu:
                log(f"cpu baseline: {cpu['value'] / 1e6:.3f} M pts/s on {cpu['cores']} threads")
        except Exception as ex:  # the baseline is reported, never the measured product
            log(f"cpu baseline failed: {ex}")
    out = {
        "metric": "projected pts/s (8M-pt MLoD, \u22641e-6\u00b7diag vs FP64) vs plain; serve p50/p99 ms",
        "value": head["pts_per_s"],
        "unit": "projected pts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["total_ms"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp32c": "fp32-compensated+fp64-fit", "fp32": "fp32", "fp64": "fp64"}[args.precision],
        "data": "synthetic (seeded superquadric cloud and offset queries)",
        "config": {"workload": args.workload if args.scale == 1.0 else f"{args.workload}@{args.scale}", "mode": mode,
                   "points": n, "queries": m, "kernel": {0: "Gaussian", 1: "rational", 2: "multi-Gaussian S=%d" % len(w["kernel"].get("sigmas", []))}[
                       w["kernel"]["type"]],
                   "cutoff": w["cutoff"], "min_depth": 0, "semantics": args.semantics, "precision": args.precision,
                   "traversal": args.traversal,
                   "octree_nodes": info["n_nodes"], "octree_depth": info["depth"], "build_ms": build_ms,
                   "parallelism": "single GPU"},
        "error_vs_fp64": errs,
        "forms": {f: {"pts_per_s": r["pts_per_s"], "ms_per_step": r["total_ms"] / args.steps} for f, r in results.items()},
        "counters_per_step": counts,
        "counters_later_steps": counts_later,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "serving": serving,
        "adaptive_octree": adaptive,
        "real_models": real,
        "surface_mesh": mesh,
        "other_configs": other,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
