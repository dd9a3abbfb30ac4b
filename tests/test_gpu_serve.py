"""Continuous-batching server: every request gives the same bits served alone, served in a mix
of trees, kernels and parameters, and projected through the one-tree entry point."""
import numpy as np
import pytest

from mlodsurfaces_amd import datasets as ds
from mlodsurfaces_amd.apss import Kernel
from mlodsurfaces_amd.serve import Server

from gpu_util import gpu_project, gpu_tree

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def trees():
    clouds = [ds.sphere(30_000, seed=1), ds.torus(50_000, seed=2), ds.superquadric(40_000, seed=3)]
    return clouds, [gpu_tree(p, n) for p, n in clouds]


def _requests(clouds, trees, n=14, seed=7):
    rng = np.random.default_rng(seed)
    reqs = []
    for r in range(n):
        c = int(rng.integers(0, len(trees)))
        P = clouds[c][0]
        rad = trees[c].radius()
        m = int(np.exp(rng.uniform(np.log(50), np.log(6000))))
        idx = rng.integers(0, len(P), m)
        q = (P[idx] + rng.uniform(-0.02, 0.02, (m, 3)) * rad).astype(np.float32)
        kt = r % 3
        if kt == 0:
            k = Kernel(type=0, sigma=float(np.float32(0.02) * np.float32(rad)), constant_shift=0.0)
        elif kt == 1:
            k = Kernel(type=1, exponent=float(rng.choice([2.0, 4.0, 2.5])), constant_shift=1e-3)
        else:
            k = Kernel(type=2, sigmas=[float(np.float32(0.005 * 2.0 ** i) * np.float32(rad)) for i in range(int(rng.integers(1, 9)))])
        reqs.append(dict(tree=c, q=q, kernel=k, cutoff=float(rng.choice([1.5, 2.0, 3.0])),
                         min_depth=int(rng.choice([0, 0, 3])), n_iter=int(rng.integers(1, 5))))
    return reqs


def _serve(srv, tids, reqs, together):
    out = {}
    if together:
        tickets = [srv.submit(tids[r["tree"]], r["q"], r["kernel"], r["cutoff"], r["min_depth"], r["n_iter"]) for r in reqs]
        while len(out) < len(reqs):
            for t, lat, qm, cs, p, n in srv.poll(timeout=5.0):
                out[t] = (p, n)
        return [out[t] for t in tickets]
    res = []
    for r in reqs:
        t = srv.submit(tids[r["tree"]], r["q"], r["kernel"], r["cutoff"], r["min_depth"], r["n_iter"])
        got = []
        while not got:
            got = srv.poll(timeout=5.0)
        assert got[0][0] == t
        res.append((got[0][4], got[0][5]))
    return res


@pytest.mark.parametrize("semantics,precision", [("ref_gpu", "fp32c"), ("ref_cpu", "fp32")])
def test_alone_equals_mixed_equals_standalone(trees, semantics, precision):
    clouds, tr = trees
    reqs = _requests(clouds, tr)
    srv_a = Server(max_batch=8192, inflight=2, semantics=semantics, precision=precision)
    tids = [srv_a.add_tree(t) for t in tr]
    alone = _serve(srv_a, tids, reqs, together=False)
    # alone, an interactive request takes all its steps in one launch
    assert srv_a.stats()["batches"] == len(reqs)
    srv_a.close()
    srv_b = Server(max_batch=8192, inflight=3, semantics=semantics, precision=precision, chunk=1024)
    tids = [srv_b.add_tree(t) for t in tr]
    mixed = _serve(srv_b, tids, reqs, together=True)
    st = srv_b.stats()
    srv_b.close()
    assert st["requests"] == len(reqs) and st["queries"] == sum(len(r["q"]) for r in reqs)
    assert st["batches"] < len(reqs)  # requests shared launches
    assert max(len(r["q"]) for r in reqs) > 1024  # and some were cut into chunks
    for r, (pa, na), (pm, nm) in zip(reqs, alone, mixed):
        assert np.array_equal(pa, pm, equal_nan=True) and np.array_equal(na, nm, equal_nan=True)
        ps, ns, _ = gpu_project(tr[r["tree"]], r["q"], r["kernel"], cutoff=r["cutoff"], min_depth=r["min_depth"],
                                semantics=semantics, precision=precision, n_iter=r["n_iter"])
        assert np.array_equal(pa, ps, equal_nan=True) and np.array_equal(na, ns, equal_nan=True)
        assert np.isfinite(pa).all()


def test_server_rejects_bad_requests(trees):
    """Requests are checked as mlod_set_kernel / mlod_set_params check a tree's settings."""
    clouds, tr = trees
    srv = Server(max_batch=1024, inflight=1)
    tid = srv.add_tree(tr[0])
    q = np.zeros((1001, 3), np.float32)
    k = Kernel(type=0, sigma=0.1)
    from mlodsurfaces_amd._lib import MlodError
    bad = [
        dict(queries=q[:0]),                                  # empty
        dict(tree_id=tid + 5),                                # unknown tree
        dict(kernel=Kernel(type=0, sigma=0.0)),               # zero sigma: -inf exponent coefficients
        dict(kernel=Kernel(type=0, sigma=float("nan"))),
        dict(kernel=Kernel(type=2, sigmas=[0.1, -0.2])),      # negative multi-Gaussian sigma
        dict(kernel=Kernel(type=2, sigmas=[0.1, 0.0])),
        dict(cutoff=-1.0),
        dict(cutoff=float("nan")),
        dict(step_max=float("nan")),
    ]
    for b in bad:
        a = dict(tree_id=tid, queries=q[:10], kernel=k, cutoff=2.0, step_max=-1.0)
        a.update(b)
        with pytest.raises(MlodError):
            srv.submit(a["tree_id"], a["queries"], a["kernel"], cutoff=a["cutoff"], step_max=a["step_max"])
    with pytest.raises(ValueError):
        Kernel(type=2, sigmas=[0.1] * 33).to_c()  # more sigmas than a kernel holds
    srv.close()
    with pytest.raises(MlodError):
        Server(max_batch=4096, chunk=8192)  # chunk larger than a batch


@pytest.mark.parametrize("precision", ["fp64", "fp32c"])
def test_many_sigmas_and_fp64_form(trees, precision):
    """Kernels of up to 32 sigmas (the reference's sample settings use up to 30) and the FP64
    form: bulk requests whose steps span several launches give the bits of one mlod_project call
    (the FP64 form carries its FP64 position from launch to launch)."""
    clouds, tr = trees
    rng = np.random.default_rng(31)
    reqs = []
    for c, m, it, ns in [(0, 4000, 30, 30), (1, 2500, 3, 32), (2, 7000, 5, 15), (1, 300, 12, 9), (0, 60_000, 40, 12)]:
        P = clouds[c][0]
        rad = tr[c].radius()
        q = (P[rng.integers(0, len(P), m)] + rng.uniform(-0.02, 0.02, (m, 3)) * rad).astype(np.float32)
        k = Kernel(type=2, sigmas=[float(np.float32(0.002 * 1.2 ** i) * np.float32(rad)) for i in range(ns)])
        reqs.append(dict(tree=c, q=q, kernel=k, cutoff=2.0, min_depth=0, n_iter=it))
    srv = Server(max_batch=16384, inflight=3, precision=precision, chunk=2048)
    tids = [srv.add_tree(t) for t in tr]
    got = _serve(srv, tids, reqs, together=True)
    st = srv.stats()
    srv.close()
    assert st["batches"] > 10  # the 60k x 40-step request is bulk: two steps per launch
    for r, (p, n) in zip(reqs, got):
        ps, ns_, _ = gpu_project(tr[r["tree"]], r["q"], r["kernel"], cutoff=r["cutoff"], min_depth=r["min_depth"],
                                 semantics="ref_gpu", precision=precision, n_iter=r["n_iter"])
        assert np.array_equal(p, ps, equal_nan=True) and np.array_equal(n, ns_, equal_nan=True)
        assert np.isfinite(p).all()


def test_budget_admission(trees):
    """Memory-footprint admission forced from the host: a budget that leaves the pool two chunks
    (a small fraction of the submitted work) keeps every request queued until blocks free up, and
    all complete with the standalone bits; a tree beyond the budget, a budget below the trees and a
    budget leaving less than two chunks of pool are refused with MLOD_ERR_OOM; the budget cannot
    change while requests are outstanding."""
    from mlodsurfaces_amd._lib import ERR_OOM, ERR_STATE, MlodError
    clouds, tr = trees
    srv = Server(max_batch=8192, inflight=2, chunk=1024)
    tids = [srv.add_tree(t) for t in tr[:2]]
    mem = srv.memory()
    assert mem["trees"] == sum(t.info()["device_bytes"] for t in tr[:2])
    with pytest.raises(MlodError) as e:
        srv.set_budget(mem["trees"] + mem["staging"] - 1)  # below what is held already
    assert e.value.code == ERR_OOM
    with pytest.raises(MlodError) as e:
        srv.set_budget(mem["trees"] + mem["staging"] + 36 * 1024)  # under two chunks of pool
    assert e.value.code == ERR_OOM
    budget = mem["trees"] + mem["staging"] + 2 * 1024 * 36 + 100
    srv.set_budget(budget)
    mem = srv.memory()
    assert mem["budget"] == budget and mem["pool_slots"] == 2048
    assert mem["trees"] + mem["staging"] + mem["pool"] <= budget
    with pytest.raises(MlodError) as e:
        srv.add_tree(tr[2])  # the third tree does not fit the budget
    assert e.value.code == ERR_OOM
    rng = np.random.default_rng(41)
    reqs = []
    for i in range(16):
        c = i % 2
        P = clouds[c][0]
        rad = tr[c].radius()
        m = int(rng.integers(500, 6000))
        q = (P[rng.integers(0, len(P), m)] + rng.uniform(-0.02, 0.02, (m, 3)) * rad).astype(np.float32)
        k = Kernel(type=2, sigmas=[float(np.float32(0.01 * 2.0 ** j) * np.float32(rad)) for j in range(4)])
        reqs.append(dict(tree=c, q=q, kernel=k, cutoff=2.0, min_depth=0, n_iter=int(rng.integers(1, 6))))
    assert sum(len(r["q"]) for r in reqs) > 10 * 2048
    tickets = [srv.submit(tids[r["tree"]], r["q"], r["kernel"], r["cutoff"], r["min_depth"], r["n_iter"]) for r in reqs]
    with pytest.raises(MlodError) as e:
        srv.set_budget(0)
    assert e.value.code == ERR_STATE
    out = {}
    while len(out) < len(reqs):
        for t, lat, qm, cs, p, n in srv.poll(timeout=5.0):
            out[t] = (p, n)
    for r, t in zip(reqs, tickets):
        ps, ns, _ = gpu_project(tr[r["tree"]], r["q"], r["kernel"], cutoff=2.0, semantics="ref_gpu", precision="fp32c",
                                n_iter=r["n_iter"])
        assert np.array_equal(out[t][0], ps) and np.array_equal(out[t][1], ns)
    srv.close()
    # the same pool is less than two chunks of 4096 queries
    srv = Server(max_batch=8192, inflight=1, chunk=4096)
    tid = srv.add_tree(tr[0])
    mem = srv.memory()
    with pytest.raises(MlodError):
        srv.set_budget(mem["trees"] + mem["staging"] + 2 * 1024 * 36 + 100)  # less than two 4096 chunks
    srv.close()


def test_cancel_leaves_other_requests_unchanged(trees):
    """A request cancelled mid-way (a bulk request of many launches) completes with status 1 and
    the requests around it give their standalone bits."""
    clouds, tr = trees
    rng = np.random.default_rng(51)
    reqs = []
    for c, m, it in [(0, 50_000, 60), (1, 3_000, 4), (2, 20_000, 3), (0, 800, 10), (1, 40_000, 50)]:
        P = clouds[c][0]
        rad = tr[c].radius()
        q = (P[rng.integers(0, len(P), m)] + rng.uniform(-0.02, 0.02, (m, 3)) * rad).astype(np.float32)
        k = Kernel(type=2, sigmas=[float(np.float32(0.01 * 2.0 ** j) * np.float32(rad)) for j in range(4)])
        reqs.append(dict(tree=c, q=q, kernel=k, n_iter=it))
    srv = Server(max_batch=16384, inflight=2, chunk=4096)
    tids = [srv.add_tree(t) for t in tr]
    tickets = [srv.submit(tids[r["tree"]], r["q"], r["kernel"], 2.0, 0, r["n_iter"]) for r in reqs]
    srv.cancel(tickets[0])
    out = {}
    while len(out) < len(reqs):
        for rec in srv.poll(timeout=5.0, with_status=True):
            out[rec[0]] = rec
    from mlodsurfaces_amd._lib import MlodError
    with pytest.raises(MlodError):
        srv.cancel(tickets[0])  # no longer outstanding
    srv.close()
    assert out[tickets[0]][6] == 1
    for r, t in list(zip(reqs, tickets))[1:]:
        assert out[t][6] == 0
        ps, ns, _ = gpu_project(tr[r["tree"]], r["q"], r["kernel"], cutoff=2.0, semantics="ref_gpu", precision="fp32c",
                                n_iter=r["n_iter"])
        assert np.array_equal(out[t][4], ps) and np.array_equal(out[t][5], ns)


def test_tree_added_while_serving(trees):
    """Trees registered while requests are in flight: requests keep their tree (resolved at
    submission) and every result is the standalone one."""
    clouds, tr = trees
    rng = np.random.default_rng(61)
    srv = Server(max_batch=8192, inflight=3, chunk=1024)
    tid0 = srv.add_tree(tr[0])
    rad = tr[0].radius()
    k = Kernel(type=2, sigmas=[float(np.float32(0.01 * 2.0 ** j) * np.float32(rad)) for j in range(4)])
    P = clouds[0][0]
    qs = [(P[rng.integers(0, len(P), m)] + rng.uniform(-0.02, 0.02, (m, 3)) * rad).astype(np.float32)
          for m in (9000, 5000, 7000)]
    tickets = [srv.submit(tid0, q, k, 2.0, 0, 20) for q in qs]
    extra = [srv.add_tree(tr[i % 3]) for i in range(40)]  # the tree list grows (and reallocates) meanwhile
    tickets.append(srv.submit(extra[-1], qs[1], k, 2.0, 0, 3))
    out = {}
    while len(out) < len(tickets):
        for t, lat, qm, cs, p, n in srv.poll(timeout=5.0):
            out[t] = (p, n)
    srv.close()
    for q, t, it, tree in zip(qs + [qs[1]], tickets, (20, 20, 20, 3), (tr[0], tr[0], tr[0], tr[39 % 3])):
        ps, ns, _ = gpu_project(tree, q, k, cutoff=2.0, semantics="ref_gpu", precision="fp32c", n_iter=it)
        assert np.array_equal(out[t][0], ps) and np.array_equal(out[t][1], ns)


def test_bulk_and_interactive_mix(trees):
    """Bulk requests (more than 2M projected points) advance two steps per launch, wait in the
    pool while interactive ones overtake them, and still give the standalone bits; a request of
    zero steps returns its queries with zero normals."""
    clouds, tr = trees
    rng = np.random.default_rng(21)
    reqs = []
    for c, m, it in [(0, 60_000, 35), (2, 45_000, 47), (1, 3_000, 3), (0, 700, 10), (2, 9_000, 1), (1, 20_000, 5)]:
        P = clouds[c][0]
        rad = tr[c].radius()
        q = (P[rng.integers(0, len(P), m)] + rng.uniform(-0.02, 0.02, (m, 3)) * rad).astype(np.float32)
        k = Kernel(type=2, sigmas=[float(np.float32(0.01 * 2.0 ** i) * np.float32(rad)) for i in range(4)])
        reqs.append(dict(tree=c, q=q, kernel=k, cutoff=2.0, min_depth=0, n_iter=it))
    srv = Server(max_batch=16384, inflight=3, chunk=4096)
    tids = [srv.add_tree(t) for t in tr]
    got = _serve(srv, tids, reqs, together=True)
    z = srv.submit(tids[0], reqs[3]["q"], reqs[3]["kernel"], 2.0, 0, 0)
    zero = srv.poll(timeout=5.0)
    st = srv.stats()
    srv.close()
    assert st["requests"] == len(reqs) + 1 and st["pending"] == 0
    assert zero[0][0] == z and np.array_equal(zero[0][4], reqs[3]["q"]) and not zero[0][5].any()
    for r, (p, n) in zip(reqs, got):
        ps, ns, _ = gpu_project(tr[r["tree"]], r["q"], r["kernel"], cutoff=r["cutoff"], min_depth=r["min_depth"],
                                semantics="ref_gpu", precision="fp32c", n_iter=r["n_iter"])
        assert np.array_equal(p, ps, equal_nan=True) and np.array_equal(n, ns, equal_nan=True)


@pytest.mark.timeout(900)
def test_serving_mix_64_clouds(oracle_built):
    """The serving configuration: 64 resident clouds (sphere / torus / superquadric, 50k-4M points)
    and a Poisson mix of requests of 1k-500k queries with the three kernel types, 10 steps each.
    Every result equals the same queries projected through mlod_project bit for bit, and a sample
    on three clouds (one of each shape) is within 1e-6 x diag of the FP64 oracle (REF_GPU)."""
    from mlodsurfaces_amd import serve as SV
    from mlodsurfaces_amd.apss import APSSOctree
    clouds = SV.serving_workload()
    trees = [APSSOctree().build(p, n) for p, n in clouds]
    srv = Server(max_batch=131072, inflight=3, chunk=32768)
    res, done, meta = SV.run_load(srv, trees, clouds, rate=120.0, seconds=1.5, seed=9, keep_queries=True)
    srv.close()
    print(f"C4 mix: {res['requests']} requests, {res['queries']} queries, p50 {res['p50_ms']:.1f} "
          f"p99 {res['p99_ms']:.1f} ms", flush=True)
    assert len(done) == res["requests"] >= 100
    assert {meta[d[0]][1] for d in done} == set(SV.KERNELS)
    by_cloud = {}
    for t, lat, qm, cs, p, n in done:
        m, kname, c, _, q = meta[t]
        k = SV.make_kernel(kname, trees[c].radius())
        ps, ns, _ = gpu_project(trees[c], q, k, cutoff=2.0, semantics="ref_gpu", precision="fp32c", n_iter=10)
        assert np.array_equal(p, ps, equal_nan=True) and np.array_equal(n, ns, equal_nan=True), (t, kname, c)
        by_cloud.setdefault(c, []).append((q, p, kname))
    # FP64 oracle on a sample of three clouds, the smallest of each shape that received requests
    rng = np.random.default_rng(3)
    checked = 0
    for shape in range(3):
        cs = sorted((c for c in by_cloud if c % 3 == shape), key=lambda c: len(clouds[c][0]))
        if not cs:
            continue
        c = cs[0]
        P, N = clouds[c]
        diag = float(np.linalg.norm(P.max(0).astype(np.float64) - P.min(0)))
        o = oracle_built.OracleTree(P, N)
        for q, p, kname in by_cloud[c][:3]:
            idx = rng.choice(len(q), min(60, len(q)), replace=False)
            k = SV.make_kernel(kname, trees[c].radius())
            ok = oracle_built.make_kernel(type=k.type, sigma=k.sigma, exponent=k.exponent,
                                          constant_shift=k.constant_shift, sigmas=k.sigmas)
            op, _ = o.project(q[idx], ok, oracle_built.make_params(cutoff=2.0, semantics="ref_gpu"), n_iter=10,
                              precision=1, nthreads=16)
            err = np.linalg.norm(p[idx].astype(np.float64) - op, axis=1) / diag
            fin = np.isfinite(op).all(1)
            assert fin.mean() > 0.9 and err[fin].max() < 1e-6, (c, kname, err[fin].max())
            checked += int(fin.sum())
    assert checked >= 300
