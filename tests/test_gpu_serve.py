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
    clouds, tr = trees
    srv = Server(max_batch=1024, inflight=1)
    tid = srv.add_tree(tr[0])
    q = np.zeros((1001, 3), np.float32)
    k = Kernel(type=0, sigma=0.1)
    from mlodsurfaces_amd._lib import MlodError
    with pytest.raises(MlodError):
        srv.submit(tid, q[:0], k)  # empty
    with pytest.raises(MlodError):
        srv.submit(tid + 5, q[:10], k)  # unknown tree
    with pytest.raises(MlodError):
        srv.submit(tid, q[:10], Kernel(type=2, sigmas=[0.1] * 9))  # more sigmas than serving takes
    srv.close()
    with pytest.raises(MlodError):
        Server(precision="fp64")
    with pytest.raises(MlodError):
        Server(max_batch=4096, chunk=8192)  # chunk larger than a batch


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
