"""GPU: convex hulls built on the device (mfcl_convex_hulls) against qhull
(scipy.spatial.ConvexHull, the hull the reference's callers pass to
fcl::Convex).  For point sets in general position the hull's vertex set
and triangulation are unique: both must equal qhull's.  Every triangle is
oriented outward with all points on or inside it, every edge is shared by
exactly two triangles, and Euler's relation holds."""
import numpy as np
import pytest
from scipy.spatial import ConvexHull

from mfcl import workloads as W


def _sets(rng):
    sets = []
    # points on ellipsoids (every point a vertex), sizes across the convex-hull configuration's range
    for n in (4, 5, 8, 9, 16, 33, 64, 100, 128, 256, 511, 512, 700):
        u = rng.normal(size=(n, 3))
        u /= np.linalg.norm(u, axis=1, keepdims=True)
        sets.append((u * rng.uniform(0.05, 0.4, 3)).astype(np.float32))
    # points in a cube (interior points left out)
    for n in (10, 50, 200, 768):
        sets.append(rng.uniform(-1, 1, (n, 3)).astype(np.float32))
    # the configuration's own generator
    v, off, cnt, _ = W.ellipsoid_hulls(rng, 60, 8, 512)
    for i in range(60):
        sets.append(v[off[i]:off[i] + cnt[i]])
    return sets


def _check_topology(P, hull_idx, tris):
    V = P[hull_idx].astype(np.float64)
    nv, nt = len(hull_idx), len(tris)
    assert nt == 2 * nv - 4
    edges = {}
    for t in tris:
        for k in range(3):
            a, b = int(t[k]), int(t[(k + 1) % 3])
            edges[(a, b)] = edges.get((a, b), 0) + 1
    # each directed edge once, its reverse once (closed, consistently oriented)
    assert all(c == 1 for c in edges.values())
    assert all((b, a) in edges for (a, b) in edges)
    for t in tris:
        a, b, c = V[t[0]], V[t[1]], V[t[2]]
        n = np.cross(b - a, c - a)
        d = (P.astype(np.float64) - a) @ n
        assert d.max() <= 1e-9 * np.linalg.norm(n) * np.abs(P).max()


@pytest.mark.gpu
def test_device_hulls_equal_qhull(gpu_ctx):
    rng = np.random.default_rng(50)
    sets = _sets(rng)
    cnt = np.array([len(s) for s in sets], np.uint32)
    off = np.zeros(len(sets), np.uint64)
    off[1:] = np.cumsum(cnt[:-1])
    res = gpu_ctx.convex_hulls(np.concatenate(sets), off, cnt)
    n_eq = 0
    for P, (st, hidx, tris) in zip(sets, res):
        assert st == 0, (len(P), st)
        _check_topology(P, hidx, tris)
        h = ConvexHull(P.astype(np.float64))
        keep = np.unique(h.simplices)
        assert np.array_equal(hidx, keep), (len(P), len(hidx), len(keep))
        ours = {tuple(sorted(hidx[t].tolist())) for t in tris}
        theirs = {tuple(sorted(s.tolist())) for s in h.simplices}
        assert ours == theirs
        n_eq += 1
    print(f"\n[device hulls] {n_eq} sets equal to qhull, sizes {int(cnt.min())}..{int(cnt.max())}")


@pytest.mark.gpu
def test_device_hulls_degenerate(gpu_ctx):
    rng = np.random.default_rng(51)
    flat = rng.uniform(-1, 1, (40, 3)).astype(np.float32)
    flat[:, 2] = 0.25  # coplanar
    line = np.outer(rng.uniform(-1, 1, 20), [1.0, 2.0, -1.0]).astype(np.float32)
    few = rng.uniform(-1, 1, (3, 3)).astype(np.float32)
    same = np.repeat(rng.uniform(-1, 1, (1, 3)), 10, axis=0).astype(np.float32)
    big = rng.uniform(-1, 1, (800, 3)).astype(np.float32)
    sets = [flat, line, few, same, big]
    cnt = np.array([len(s) for s in sets], np.uint32)
    off = np.zeros(len(sets), np.uint64)
    off[1:] = np.cumsum(cnt[:-1])
    res = gpu_ctx.convex_hulls(np.concatenate(sets), off, cnt)
    assert [r[0] for r in res] == [1, 1, 1, 1, 2]
