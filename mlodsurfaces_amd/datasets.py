"""Seeded synthetic oriented point clouds and query sets for the projection workloads.

All generators are deterministic functions of their seed (numpy PCG64) and return float32
positions, float32 unit normals, and the query points.  Shapes:

* ``sphere``       unit sphere, points displaced along the normal by Gaussian noise;
* ``torus``        R = 1, r = 0.35, radial displacement 0.02 sin(6 theta) sin(4 phi) along the
                   tube normal, normal-direction noise, spherical holes removed;
* ``superquadric`` exponents (e1, e2), axes (a, b, c), multi-octave sinusoidal displacement
                   along the normal, Gaussian noise, optional spherical holes.

Normals are the analytic normals of the undisplaced surface (the reference takes normals
as given input, sources/apss/CPUUtils.cpp:540-569), renormalised in float64 before the
cast to float32.
"""
from __future__ import annotations

import os

import numpy as np

__all__ = ["sphere", "torus", "superquadric", "jitter_queries", "offset_queries", "make_workload"]


def _unit(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def sphere(n: int, seed: int = 1, noise: float = 0.002):
    """Uniform points on the unit sphere, p = u (1 + noise g), n = u."""
    rng = np.random.default_rng(seed)
    u = _unit(rng.standard_normal((n, 3)))
    g = rng.standard_normal(n)
    p = u * (1.0 + noise * g)[:, None]
    return p.astype(np.float32), u.astype(np.float32)


def _remove_holes(p, centers, radius):
    keep = np.ones(len(p), bool)
    for c in centers:
        keep &= np.sum((p - c) ** 2, axis=1) > radius * radius
    return keep


def torus(n: int, seed: int = 2, R: float = 1.0, r: float = 0.35, amp: float = 0.02,
          noise: float = 0.003, holes: int = 3, hole_radius: float = 0.1):
    """Area-uniform torus samples (rejection on the area element) with a radial
    displacement amp*sin(6 theta)*sin(4 phi), normal noise, and ``holes`` spherical holes."""
    rng = np.random.default_rng(seed)
    # hole centres on the surface
    th = rng.uniform(0, 2 * np.pi, holes)
    ph = rng.uniform(0, 2 * np.pi, holes)
    hc = np.stack([(R + r * np.cos(ph)) * np.cos(th), (R + r * np.cos(ph)) * np.sin(th), r * np.sin(ph)], 1)
    out_p, out_n, got = [], [], 0
    while got < n:
        m = int((n - got) * 1.6) + 1024
        theta = rng.uniform(0, 2 * np.pi, m)
        phi = rng.uniform(0, 2 * np.pi, m)
        acc = rng.uniform(0, 1, m) < (R + r * np.cos(phi)) / (R + r)
        theta, phi = theta[acc], phi[acc]
        nrm = np.stack([np.cos(phi) * np.cos(theta), np.cos(phi) * np.sin(theta), np.sin(phi)], 1)
        base = np.stack([R * np.cos(theta), R * np.sin(theta), np.zeros_like(theta)], 1)
        disp = r + amp * np.sin(6 * theta) * np.sin(4 * phi) + noise * rng.standard_normal(len(theta))
        p = base + nrm * disp[:, None]
        keep = _remove_holes(p, hc, hole_radius)
        out_p.append(p[keep])
        out_n.append(nrm[keep])
        got += int(keep.sum())
    p = np.concatenate(out_p)[:n]
    nn = np.concatenate(out_n)[:n]
    return p.astype(np.float32), _unit(nn).astype(np.float32)


def superquadric(n: int, seed: int = 3, e1: float = 0.4, e2: float = 0.4, axes=(1.0, 0.7, 0.5),
                 amp: float = 0.01, noise: float = 0.001, holes: int = 0, hole_radius: float = 0.1):
    """Superquadric F(x) = (|x/a|^(2/e2) + |y/b|^(2/e2))^(e2/e1) + |z/c|^(2/e1) = 1.

    Uniform directions u are mapped radially onto the surface (F(t u) = t^(2/e1) F(u), so
    t = F(u)^(-e1/2)) and accepted with probability proportional to the area the map
    covers per solid angle, t^2 / (n . u), which makes the samples area-uniform.  Points are
    displaced along the analytic normal grad F / |grad F| by a 3-octave sinusoid of the
    direction's longitude/latitude (total amplitude ``amp``) plus Gaussian noise; ``holes``
    spherical holes of radius ``hole_radius`` are removed."""
    rng = np.random.default_rng(seed)
    a, b, c = axes

    def radial(u):
        g = np.abs(u[:, 0] / a) ** (2 / e2) + np.abs(u[:, 1] / b) ** (2 / e2)
        F = g ** (e2 / e1) + np.abs(u[:, 2] / c) ** (2 / e1)
        t = F ** (-e1 / 2)
        x = u * t[:, None]
        gx = np.abs(x[:, 0] / a) ** (2 / e2) + np.abs(x[:, 1] / b) ** (2 / e2)
        k = (2 / e1) * gx ** (e2 / e1 - 1)
        nx = k * np.abs(x[:, 0] / a) ** (2 / e2 - 1) * np.sign(x[:, 0]) / a
        ny = k * np.abs(x[:, 1] / b) ** (2 / e2 - 1) * np.sign(x[:, 1]) / b
        nz = (2 / e1) * np.abs(x[:, 2] / c) ** (2 / e1 - 1) * np.sign(x[:, 2]) / c
        nr = _unit(np.stack([nx, ny, nz], 1))
        jac = t * t / np.maximum(np.sum(nr * u, 1), 1e-6)
        return x, nr, jac

    probe = _unit(rng.standard_normal((200000, 3)))
    jmax = float(np.max(radial(probe)[2])) * 1.05
    hc = None
    if holes:
        hc = radial(_unit(rng.standard_normal((holes, 3))))[0]
    out_p, out_n, got = [], [], 0
    while got < n:
        m = min(int((n - got) * 3.0) + 4096, 4_000_000)
        u = _unit(rng.standard_normal((m, 3)))
        x, nr, jac = radial(u)
        acc = rng.uniform(0, jmax, m) < jac
        u, x, nr = u[acc], x[acc], nr[acc]
        lon = np.arctan2(u[:, 1], u[:, 0])
        lat = np.arcsin(np.clip(u[:, 2], -1, 1))
        d = amp * (0.5714 * np.sin(5 * lon) * np.sin(4 * lat) + 0.2857 * np.sin(11 * lon + 1.0) * np.sin(9 * lat + 0.5)
                   + 0.1429 * np.sin(23 * lon + 2.0) * np.sin(19 * lat + 1.5))
        d = d + noise * rng.standard_normal(len(u))
        p = x + nr * d[:, None]
        ok = np.all(np.isfinite(p), axis=1) & np.all(np.isfinite(nr), axis=1)
        if hc is not None:
            ok &= _remove_holes(p, hc, hole_radius)
        out_p.append(p[ok])
        out_n.append(nr[ok])
        got += int(ok.sum())
    p = np.concatenate(out_p)[:n]
    nn = np.concatenate(out_n)[:n]
    return p.astype(np.float32), _unit(nn).astype(np.float32)


def jitter_queries(p: np.ndarray, amp: float, seed: int):
    """Queries = inputs + U(-amp, amp)^3."""
    rng = np.random.default_rng(seed)
    return (p.astype(np.float64) + rng.uniform(-amp, amp, p.shape)).astype(np.float32)


def offset_queries(p: np.ndarray, m: int, amp: float, seed: int):
    """m queries p_j + U(-amp, amp)^3 with j drawn uniformly."""
    rng = np.random.default_rng(seed)
    j = rng.integers(0, len(p), m)
    return (p[j].astype(np.float64) + rng.uniform(-amp, amp, (m, 3))).astype(np.float32)


def pn_normalize(p: np.ndarray) -> np.ndarray:
    """The reference application's normalisation of a loaded cloud (Pn::normalize,
    sources/apss/Pn.cpp:262-290): centroid to the origin, then division by the largest
    |coordinate| of the bounding box grown by 0.5 % of its diagonal, so the cloud sits inside
    [-1, 1]^3 with a margin (the frame of the adaptive octree)."""
    p64 = p.astype(np.float64)
    off = p64.mean(0)
    mx, mn = p64.max(0) - off, p64.min(0) - off
    small = np.linalg.norm(mx - mn) * 0.005
    norme = max(np.abs(mx + small).max(), np.abs(mn - small).max())
    return ((p64 - off) / norme).astype(np.float32)


def bbox_radius(p: np.ndarray) -> float:
    """Half the diagonal of the float AABB (BBOX::radius, sources/apss/CPUUtils.h:84-85)."""
    lo = p.min(0).astype(np.float32)
    hi = p.max(0).astype(np.float32)
    d = (lo - hi).astype(np.float32)
    d2 = np.float32(np.float32(d[0] * d[0]) + np.float32(d[1] * d[1])) + np.float32(d[2] * d[2])
    return float(np.float32(np.sqrt(np.float32(d2))) / np.float32(2.0))


def make_workload(name: str, scale: float = 1.0):
    """The named projection workloads (sizes scaled by ``scale`` for test runs).

    Returns dict(points, normals, queries, kernel=dict, cutoff, description)."""
    if name == "sphere100k":
        n = max(64, int(100_000 * scale))
        p, nr = sphere(n, seed=1, noise=0.002)
        R = bbox_radius(p)
        sig = [0.01 * 2.0 ** i * R for i in range(4)]
        return dict(points=p, normals=nr, queries=p.copy(), kernel=dict(type=2, sigmas=sig),
                    cutoff=3.0, description=f"sphere N=M={n}, multi-Gaussian S=4 sigma1=0.01R factor 2, cutoff 3")
    if name == "torus1m":
        n = max(64, int(1_000_000 * scale))
        p, nr = torus(n, seed=2)
        q = jitter_queries(p, 0.01, seed=22)
        return dict(points=p, normals=nr, queries=q, kernel=dict(type=1, exponent=4.0, constant_shift=1e-3),
                    cutoff=3.0, description=f"torus N=M={n}, rational e=4 c=1e-3, cutoff 3, queries jittered +-0.01")
    if name == "superquadric8m":
        n = max(64, int(8_000_000 * scale))
        p, nr = superquadric(n, seed=3)
        lo, hi = p.min(0).astype(np.float64), p.max(0).astype(np.float64)
        diag = float(np.linalg.norm(hi - lo))
        q = offset_queries(p, n, 0.02 * diag, seed=33)
        R = bbox_radius(p)
        sig = [0.01 * 2.0 ** i * R for i in range(6)]
        return dict(points=p, normals=nr, queries=q, kernel=dict(type=2, sigmas=sig), cutoff=2.0,
                    description=f"superquadric N=M={n}, multi-Gaussian S=6, cutoff 2, offsets +-0.02 diag")
    if name in ("c5_knn", "c5_ball"):
        cache = os.environ.get("MLOD_DATA_CACHE")
        path = os.path.join(cache, f"c5_{scale:g}.npz") if cache else None
        if path and os.path.exists(path):
            with np.load(path, allow_pickle=False) as z:
                p, nr, q = z["p"], z["n"], z["q"]
        else:
            p = nr = q = None
        # global-APSS variants at scale: 20M points on the displaced superquadric with 10 holes.
        # The queries are the centres of the dual-contouring grid's finest cells (the adaptive
        # octree generated to depth 11 with the MLoD projection, uniformized; dc.grid_centres),
        # capped at 16M: they need the GPU, so the caller computes them from `dc_grid` once the
        # octree is built.  `queries` holds, for callers without the GPU, the depth-11 cells of
        # [-1, 1]^3 enclosing the points (generateFromPoints, upsample.cpp:128-144).
        # The cloud is taken in the reference's normalised frame (the application normalises every
        # cloud it loads, Pn.cpp:262-290), so the adaptive octree's [-1, 1]^3 frame holds it and
        # the grid reaches depth 11.
        n = max(64, int(20_000_000 * scale))
        cap = max(64, int(16_000_000 * scale))
        if p is None:
            p, nr = superquadric(n, seed=5, holes=10)
            p = pn_normalize(p)
            q = cell_centres_of_points(p, 11, cap=cap, seed=55)
            if path:
                os.makedirs(cache, exist_ok=True)
                np.savez(path, p=p, n=nr, q=q)
        R = bbox_radius(p)
        sig = [0.01 * 2.0 ** i * R for i in range(4)]
        mode = "knn" if name == "c5_knn" else "ball"
        return dict(points=p, normals=nr, queries=q, kernel=dict(type=2, sigmas=sig), cutoff=2.0, mode=mode, knn=32,
                    ball_radius=-1.0, dc_grid=dict(depth=11, cap=cap, seed=55),
                    description=f"superquadric N={n} (10 holes), {'kNN k=32' if mode == 'knn' else 'ball r=3 sigma0'}, "
                                f"multi-Gaussian S=4")
    raise KeyError(name)


def cell_centres_of_points(p: np.ndarray, depth: int, cap: int, seed: int) -> np.ndarray:
    """Centres of the cells of the [-1, 1]^3 octree at ``depth`` that contain a point of ``p``
    (float32 arithmetic of aabb_octree::enclosing / center), at most ``cap`` of them (a seeded
    random subset, kept in cell order)."""
    t = ((p.astype(np.float32) - np.float32(-1.0)) / np.float32(2.0)) / np.float32(2.0 ** -depth)
    c = np.floor(t).astype(np.int64)
    key = (c[:, 2] << 42) | ((c[:, 1] & 0x1fffff) << 21) | (c[:, 0] & 0x1fffff)
    _, first = np.unique(key, return_index=True)
    cells = c[np.sort(first)]
    if len(cells) > cap:
        keep = np.sort(np.random.default_rng(seed).choice(len(cells), cap, replace=False))
        cells = cells[keep]
    return ((2 * cells + 1).astype(np.float32) * np.float32(2.0 ** -depth) - np.float32(1.0)).astype(np.float32)
