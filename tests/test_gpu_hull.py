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


def _result_bytes(a):
    return np.ascontiguousarray(a).view(np.uint8)


@pytest.mark.gpu
def test_convex_points_table_equals_host_path(gpu_ctx):
    """mfcl_upload_convex_points (hull, vertex table, descriptors and
    adjacency built on the device) against the host path of the reference's
    callers: qhull's hull, mfcl_upload_shapes with its vertices and
    mfcl_upload_convex_faces with its triangles.  Every query result must be
    byte-identical (the walk tables, seeds, interior points and bounds are
    then the same), the shape descriptors equal to float rounding."""
    import mfcl

    rng = np.random.default_rng(52)
    v, off, cnt, _ = W.ellipsoid_hulls(rng, 1024, 8, 512)
    sets = [v[off[i]:off[i] + cnt[i]] for i in range(len(cnt))]
    for n in (12, 40, 300, 768):  # cube points: interior points are not hull vertices
        sets.append(rng.uniform(-0.3, 0.3, (n, 3)).astype(np.float32))
    cnt = np.array([len(s) for s in sets], np.uint32)
    off = np.zeros(len(sets), np.uint64)
    off[1:] = np.cumsum(cnt[:-1])
    pts = np.concatenate(sets)
    hc = gpu_ctx.upload_convex_points(pts, off, cnt)
    # host path
    hv, hvo, hvc, ht, hto, htc = [], [], [], [], [], []
    vo = to = 0
    for P in sets:
        keep, tris = W.hull_faces(P)
        hv.append(P[keep]); hvo.append(vo); hvc.append(len(keep)); vo += len(keep)
        ht.append(tris); hto.append(to); htc.append(len(tris)); to += len(tris)
    assert np.array_equal(hc, np.array(hvc, np.uint32))
    n = len(sets)
    host = mfcl.Context(0)
    try:
        host.upload_shapes(np.full(n, W.CONVEX, np.int32), np.zeros((n, 4), np.float32), np.array(hvo, np.uint64),
                           np.array(hvc, np.uint32), np.concatenate(hv),
                           faces=mfcl.faces_from_triangles(np.concatenate(ht), np.array(hto), np.array(htc)))
        m = 16384
        pa = rng.integers(0, n, m).astype(np.uint32)
        pb = rng.integers(0, n, m).astype(np.uint32)
        qa, qb = W.poses(rng, m, 0.0, 0.6), W.poses(rng, m, 0.0, 0.6)
        n_hit = 0
        for mode, pen in ((mfcl.MODE_SIGNED_DISTANCE, 0), (mfcl.MODE_COLLIDE, 1)):
            req = mfcl.request(mode, penetration=pen)
            a = gpu_ctx.convex_query(pa, pb, qa, qb, req)
            b = host.convex_query(pa, pb, qa, qb, req)
            assert np.array_equal(_result_bytes(a), _result_bytes(b)), mode
            if mode == mfcl.MODE_COLLIDE:
                n_hit = int(a["hit"].sum())
        feats = []
        for c in (gpu_ctx, host):
            d = [c.to_device(x) for x in (pa, pb, qa, qb)]
            f = c.alloc(m * 64 * 4)
            c.mlp_features_dev(*d, m, f)
            feats.append(f.download(np.zeros((m, 64), np.float32)))
            for x in d + [f]:
                x.free()
        ferr = np.abs(feats[0] - feats[1]).max()
        assert ferr <= 1e-6, ferr
        print(f"\n[device shape table] {n} hulls ({int(hc.sum())} vertices, {int((hc > 32).sum())} walking), "
              f"{m} pairs x 2 modes byte-identical to the host path, {n_hit} colliding; descriptor max diff {ferr:.1e}")
    finally:
        host.close()


@pytest.mark.gpu
def test_convex_points_rejects_bad_sets(gpu_ctx):
    rng = np.random.default_rng(53)
    good = rng.uniform(-1, 1, (20, 3)).astype(np.float32)
    flat = good.copy()
    flat[:, 2] = 0.5
    for bad, msg in ((flat, "point set 1 is degenerate"), (rng.uniform(-1, 1, (800, 3)).astype(np.float32),
                                                           "more than 768")):
        sets = [good, bad]
        cnt = np.array([len(s) for s in sets], np.uint32)
        off = np.array([0, len(good)], np.uint64)
        with pytest.raises(RuntimeError, match=msg):
            gpu_ctx.upload_convex_points(np.concatenate(sets), off, cnt)


@pytest.mark.gpu
def test_shared_vertices_need_shared_faces(gpu_ctx):
    """The walk data lives per vertex: two walking hulls over one vertex
    range must have one face list (the same one twice is fine)."""
    import mfcl

    rng = np.random.default_rng(54)
    v, off, cnt, _ = W.ellipsoid_hulls(rng, 2, 40, 40)
    P = v[:40]
    _, t1 = W.hull_faces(P)
    _, t2 = W.hull_faces(v[40:80])  # another 40-vertex triangulation
    types = np.full(2, W.CONVEX, np.int32)
    params = np.zeros((2, 4), np.float32)
    voff = np.zeros(2, np.uint64)
    vcnt = np.full(2, 40, np.uint32)
    tris = np.concatenate([t1, t2])
    gpu_ctx.upload_shapes(types, params, voff, vcnt, P,
                          faces=mfcl.faces_from_triangles(tris, np.array([0, 0]), np.array([len(t1), len(t1)])))
    with pytest.raises(RuntimeError, match="share vertices but not their faces"):
        gpu_ctx.upload_shapes(types, params, voff, vcnt, P,
                              faces=mfcl.faces_from_triangles(tris, np.array([0, len(t1)]),
                                                              np.array([len(t1), len(t2)])))
