"""Continuous-batching projection serving over resident octrees (libmlod.so's mlod_server_*).

``Server`` wraps the C ABI.  ``serving_workload`` and ``run_load`` build the serving
configuration — 64 resident clouds (sphere / torus / superquadric, sizes log-uniform in
[50k, 4M]) and Poisson-arriving requests of log-uniform size in [1k, 500k] queries with mixed
kernels (multi-Gaussian S=4, rational e=4, Gaussian sigma=0.01, sigmas times the cloud's bbox
radius), queries = jittered surface samples of the target cloud, 10 steps each — and measure
latency (submit -> results in host memory) and aggregate projected points/s.

    python -m mlodsurfaces_amd.serve --rate 2000 --seconds 2
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import sys
import threading
import time

import numpy as np

from . import _lib
from . import datasets as ds
from .apss import APSSOctree, Kernel
from ._lib import check, lib

__all__ = ["Server", "serving_workload", "run_load"]


class Server:
    """A serving instance: resident trees, a worker thread and ``inflight`` streams, each with
    one step launch of at most ``max_batch`` queries on the GPU at a time; requests are cut into
    chunks of ``chunk`` queries (0: max_batch / 4) that stay in a device pool between steps.
    ``stats()['batches']`` counts step launches.  ``submit`` and ``poll`` may run on different
    threads: a request's buffers are recorded under a lock before its ticket can complete."""

    def __init__(self, max_batch=131072, inflight=3, semantics="ref_gpu", precision="fp32c", chunk=0):
        sem = {"ref_gpu": 0, "ref_cpu": 1}[semantics]
        prec = {"fp32": 0, "fp32c": 1, "fp64": 2}[precision]
        h = C.c_void_p()
        check(lib().mlod_server_create(int(max_batch), int(chunk), int(inflight), sem, prec, C.byref(h)))
        self.h = h
        self.max_batch = int(max_batch)
        self._keep = {}
        self._trees = []
        self._lock = threading.Lock()

    def add_tree(self, tree: APSSOctree) -> int:
        tid = C.c_uint32()
        check(lib().mlod_server_add_tree(self.h, tree.handle, C.byref(tid)))
        self._trees.append(tree)
        return int(tid.value)

    def set_budget(self, hbm_bytes: int):
        """HBM budget over the registered trees, staging and query pool (0: none)."""
        check(lib().mlod_server_set_budget(self.h, int(hbm_bytes)))

    def memory(self):
        o = (C.c_uint64 * 5)()
        check(lib().mlod_server_memory(self.h, o))
        return dict(budget=o[0], trees=o[1], staging=o[2], pool=o[3], pool_slots=o[4])

    def cancel(self, ticket: int):
        check(lib().mlod_server_cancel(self.h, int(ticket)))

    def submit(self, tree_id, queries, kernel: Kernel, cutoff=2.0, min_depth=0, n_iter=10, step_max=-1.0, out=True):
        q = np.ascontiguousarray(queries, np.float32)
        m = len(q)
        op = np.empty((m, 3), np.float32) if out else None
        on = np.empty((m, 3), np.float32) if out else None
        t = C.c_uint64()
        kc = kernel.to_c()
        # the ticket can complete (and be polled on another thread) as soon as the call returns,
        # so its buffers are recorded under the lock the poller takes
        with self._lock:
            check(lib().mlod_server_submit(self.h, int(tree_id), C.byref(kc), float(cutoff), int(min_depth),
                                           int(n_iter), float(step_max), q.ctypes.data_as(C.c_void_p), m,
                                           op.ctypes.data_as(C.c_void_p) if out else None,
                                           on.ctypes.data_as(C.c_void_p) if out else None, C.byref(t)))
            self._keep[t.value] = (q, op, on)
        return int(t.value)

    def poll(self, timeout=0.0, cap=4096, with_status=False):
        """Completed requests: list of (ticket, latency_ms, queue_ms, completed_s, points, normals)
        (plus the status, 0 completed or 1 cancelled, with ``with_status``)."""
        buf = (_lib.ServerDoneC * cap)()
        n = C.c_uint32()
        check(lib().mlod_server_poll(self.h, buf, cap, C.byref(n), float(timeout)))
        out = []
        with self._lock:
            for i in range(n.value):
                d = buf[i]
                q, op, on = self._keep.pop(d.ticket)
                rec = (d.ticket, d.latency_ms, d.queue_ms, d.completed_s, op, on)
                out.append(rec + (d.status,) if with_status else rec)
        return out

    def stats(self):
        o = (C.c_uint64 * 4)()
        check(lib().mlod_server_stats(self.h, o))
        return dict(batches=o[0], requests=o[1], queries=o[2], pending=o[3])

    def close(self):
        if getattr(self, "h", None):
            lib().mlod_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _rotation(rng):
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def serving_workload(n_clouds=64, seed=4, min_pts=50_000, max_pts=4_000_000, base_pts=4_000_000):
    """The resident clouds: shape i mod 3 (sphere, torus, superquadric), size log-uniform in
    [min_pts, max_pts]; each is a random subset of a base cloud of that shape under a random
    rotation, scale in [0.5, 1] and translation in [-0.2, 0.2]^3."""
    rng = np.random.default_rng(seed)
    bases = [ds.sphere(base_pts, seed=41), ds.torus(base_pts, seed=42), ds.superquadric(base_pts, seed=43)]
    sizes = np.exp(rng.uniform(np.log(min_pts), np.log(max_pts), n_clouds)).astype(int)
    clouds = []
    for i in range(n_clouds):
        P, N = bases[i % 3]
        idx = rng.choice(len(P), size=min(int(sizes[i]), len(P)), replace=False)
        Rm = _rotation(rng)
        s = rng.uniform(0.5, 1.0)
        t = rng.uniform(-0.2, 0.2, 3)
        p = (P[idx].astype(np.float64) @ Rm.T * s + t).astype(np.float32)
        n = (N[idx].astype(np.float64) @ Rm.T).astype(np.float32)
        clouds.append((p, n))
    return clouds


KERNELS = ("mg4", "rational", "gauss")


def make_kernel(name, radius):
    if name == "mg4":
        return Kernel(type=2, sigmas=[float(np.float32(0.01 * 2.0 ** i) * np.float32(radius)) for i in range(4)])
    if name == "rational":
        return Kernel(type=1, exponent=4.0, constant_shift=1e-3)
    return Kernel(type=0, sigma=float(np.float32(0.01) * np.float32(radius)), constant_shift=0.0)


def run_load(server, trees, clouds, rate=2000.0, seconds=2.0, min_req=1_000, max_req=500_000, n_iter=10, cutoff=2.0,
             seed=5, pool=500_000, max_requests=None, log=None, keep_queries=False):
    """Poisson arrivals at `rate` req/s for `seconds`; returns the totals, the completion records
    and, per ticket, (queries, kernel name, cloud, submit time[, the queries themselves with
    ``keep_queries``])."""
    rng = np.random.default_rng(seed)
    tids = [server.add_tree(t) for t in trees] if not hasattr(server, "_tids") else server._tids
    server._tids = tids
    radii = [t.radius() for t in trees]
    # per-cloud query pools: jittered surface samples, requests take random slices
    pools = []
    for (p, n) in clouds:
        j = rng.integers(0, len(p), min(pool, max(len(p), 1)))
        pools.append((p[j] + rng.uniform(-0.005, 0.005, (len(j), 3)) * radii[len(pools)] * 2).astype(np.float32))
    t_arr = np.cumsum(rng.exponential(1.0 / rate, int(rate * seconds * 1.5) + 16))
    t_arr = t_arr[t_arr < seconds]
    if max_requests:
        t_arr = t_arr[:max_requests]
    nreq = len(t_arr)
    sizes = np.exp(rng.uniform(np.log(min_req), np.log(max_req), nreq)).astype(int)
    clouds_i = rng.integers(0, len(trees), nreq)
    kern_i = rng.integers(0, len(KERNELS), nreq)
    meta = {}
    done = []
    stop = threading.Event()

    def poller():
        while not stop.is_set() or len(done) < nreq:
            for rec in server.poll(timeout=0.01):
                done.append(rec)
            if stop.is_set() and len(done) >= nreq:
                break

    th = threading.Thread(target=poller, daemon=True)
    th.start()
    t0 = time.perf_counter()
    for i in range(nreq):
        now = time.perf_counter() - t0
        if t_arr[i] > now:
            time.sleep(t_arr[i] - now)
        c = int(clouds_i[i])
        qp = pools[c]
        m = int(min(sizes[i], len(qp)))
        off = int(rng.integers(0, len(qp) - m + 1))
        k = make_kernel(KERNELS[kern_i[i]], radii[c])
        tk = server.submit(tids[c], qp[off:off + m], k, cutoff=cutoff, n_iter=n_iter)
        meta[tk] = (m, KERNELS[kern_i[i]], c, time.perf_counter() - t0) + ((qp[off:off + m],) if keep_queries else ())
    submit_end = time.perf_counter() - t0
    stop.set()
    th.join()
    wall = time.perf_counter() - t0
    lat = np.array([d[1] for d in done])
    msz = np.array([meta[d[0]][0] for d in done])
    pts = float(msz.sum()) * n_iter
    big = (msz >= 80_000) & (msz <= 125_000)
    res = dict(requests=nreq, offered_rate=rate, seconds=seconds, submit_s=submit_end, wall_s=wall,
               queries=int(msz.sum()), projected_pts=pts, pts_per_s=pts / wall,
               p50_ms=float(np.percentile(lat, 50)) if len(lat) else None,
               p99_ms=float(np.percentile(lat, 99)) if len(lat) else None,
               p50_100k_ms=float(np.percentile(lat[big], 50)) if big.any() else None,
               p99_100k_ms=float(np.percentile(lat[big], 99)) if big.any() else None,
               n_100k=int(big.sum()), stats=server.stats())
    if log:
        log(f"serve: {nreq} req at {rate:g}/s over {seconds:g}s, wall {wall:.2f}s, "
            f"{res['pts_per_s'] / 1e6:.1f} M pts/s, p50 {res['p50_ms']:.1f} ms p99 {res['p99_ms']:.1f} ms, "
            f"100k: p50 {res['p50_100k_ms']} p99 {res['p99_100k_ms']} ({res['n_100k']})")
    return res, done, meta


def main():
    ap = argparse.ArgumentParser(description="continuous-batching serving load test")
    ap.add_argument("--rate", type=float, nargs="+", default=[2000.0])
    ap.add_argument("--seconds", type=float, default=2.0)
    ap.add_argument("--clouds", type=int, default=64)
    ap.add_argument("--max-batch", type=int, default=131072)
    ap.add_argument("--chunk", type=int, default=32768)
    ap.add_argument("--inflight", type=int, default=3)
    ap.add_argument("--max-pts", type=int, default=4_000_000)
    args = ap.parse_args()
    log = lambda *a: print("[serve]", *a, file=sys.stderr, flush=True)  # noqa: E731
    t0 = time.time()
    clouds = serving_workload(args.clouds, max_pts=args.max_pts, base_pts=min(args.max_pts, 4_000_000))
    log(f"{len(clouds)} clouds, {sum(len(c[0]) for c in clouds) / 1e6:.1f}M points generated in {time.time() - t0:.1f}s")
    t0 = time.time()
    trees = [APSSOctree().build(p, n) for p, n in clouds]
    log(f"octrees built in {time.time() - t0:.1f}s, {sum(t.info()['device_bytes'] for t in trees) / 2**30:.1f} GiB resident")
    srv = Server(max_batch=args.max_batch, inflight=args.inflight, chunk=args.chunk)
    out = []
    for r in args.rate:
        res, _, _ = run_load(srv, trees, clouds, rate=r, seconds=args.seconds, log=log)
        out.append(res)
    print(json.dumps(out))
    srv.close()


if __name__ == "__main__":
    main()
