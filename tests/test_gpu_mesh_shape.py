"""GPU parity: mesh-shape collision (mfcl_mesh_shape_collide and, with
penetration enabled, mfcl_mesh_shape_collide_contacts;
OrientedNodeBVHSolver::MeshShapeIntersect, bvh_solver-inl.h:6-72, leaf
tests as GJKSolver::shapeTriangleIntersect, gjk_solver-inl.h:545-563)
against the oracle, which tests every triangle (the colliding set does not
depend on the tree).  Meshes of 300-3,000 triangles against boxes, spheres,
capsules, cylinders and convex hulls (walking and scanning) near their
surfaces: every boolean equal, every complete colliding set equal when it
fits the contact cap, no stack overflow; with penetration, every contact
(triangle, position, normal, depth) bit-identical."""
import numpy as np
import pytest

import mfcl
from mfcl import workloads as W

BOX, SPHERE, CAPSULE, CYLINDER, CONVEX = 0, 1, 3, 5, 100


def _shapes(rng, k):
    types, params, hv, ho, hc, ht, hto, htc = [], [], [], [], [], [], [], []
    for t in (BOX, SPHERE, CAPSULE, CYLINDER):
        for _ in range(k):
            types.append(t)
            if t == BOX:
                params.append(list(rng.uniform(0.05, 0.4, 3)) + [0])
            else:
                params.append([rng.uniform(0.03, 0.2), rng.uniform(0.05, 0.4), 0, 0])
    v, off, cnt, _ = W.ellipsoid_hulls(rng, k, 8, 200)
    o = q = 0
    for i in range(k):
        P = v[off[i]:off[i] + cnt[i]]
        keep, tri = W.hull_faces(P)
        hv.append(P[keep]); ho.append(o); hc.append(len(keep)); o += len(keep)
        ht.append(tri); hto.append(q); htc.append(len(tri)); q += len(tri)
        types.append(CONVEX)
        params.append([0, 0, 0, 0])
    n = len(types)
    nprim = 4 * k
    vo = np.zeros(n, np.uint64)
    vc = np.zeros(n, np.uint32)
    vo[nprim:] = ho
    vc[nprim:] = hc
    to = np.zeros(n, np.int64)
    tc = np.zeros(n, np.int64)
    to[nprim:] = hto
    tc[nprim:] = htc
    hull_of_shape = np.full(n, -1, np.int32)
    hull_of_shape[nprim:] = np.arange(k)
    oracle_hulls = (np.concatenate(hv), np.array(ho, np.int64), np.array(hc, np.int32), np.concatenate(ht),
                    np.array(hto, np.int64), np.array(htc, np.int32), hull_of_shape)
    return (np.array(types, np.int32), np.array(params, np.float32), vo, vc, np.concatenate(hv),
            np.concatenate(ht), to, tc, oracle_hulls)


def _setup(oracle, gpu_ctx, seed, n):
    rng = np.random.default_rng(seed)
    pool = W.mesh_pool(8, seed + 1, 300, 3000)
    types, params, vo, vc, verts, tris, to, tc, oh = _shapes(rng, 16)
    ns = len(types)
    # the oracle's hull table holds the hulls only, mapped from their shapes
    S = oracle.ShapeSet(types, params[:, :3], *oh[:6], hull_of_shape=oh[6])
    M = oracle.MeshSet(pool["verts"], pool["vert_off"], pool["vert_cnt"], pool["tris"], pool["tri_off"],
                       pool["tri_cnt"])
    gpu_ctx.upload_shapes(types, params, vo, vc, verts, faces=mfcl.faces_from_triangles(tris, to, tc))
    gpu_ctx.upload_meshes(pool["verts"], pool["vert_off"], pool["vert_cnt"], pool["tris"], pool["tri_off"],
                          pool["tri_cnt"])
    mesh_ids = rng.integers(0, len(pool["vert_cnt"]), n).astype(np.uint32)
    shape_ids = rng.integers(0, ns, n).astype(np.uint32)
    pm = W.poses(rng, n, -1.0, 1.0)
    ps = W.poses(rng, n, 0.0, 0.0)
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    r = pool["radius"][mesh_ids] * rng.uniform(0.4, 1.8, n)
    ps[:, :3] = (pm[:, :3] + u * r[:, None]).astype(np.float32)
    return S, M, types, mesh_ids, shape_ids, pm, ps


@pytest.mark.gpu
def test_mesh_shape_bit_exact_sets(oracle, gpu_ctx):
    n = 2048
    S, M, types, mesh_ids, shape_ids, pm, ps = _setup(oracle, gpu_ctx, 90, n)
    cap = 64
    req = mfcl.request(mfcl.MODE_COLLIDE, max_contacts=cap, penetration=0)
    out, got = gpu_ctx.mesh_shape_collide(mesh_ids, shape_ids, pm, ps, req)
    tm, ts = oracle.pose12(pm), oracle.pose12(ps)
    hit, cnt = S.mesh_shape(M, mesh_ids, shape_ids, tm, ts, threads=8)
    assert np.all(out["hit"] >= 0), "traversal stack overflow"
    mism = out["hit"] != hit
    assert mism.sum() == 0, np.nonzero(mism)[0][:20]
    assert np.array_equal(out["n_contacts"], np.minimum(cnt, cap))
    n_eq = 0
    for i in np.nonzero(hit)[0]:
        if cnt[i] > cap:
            continue
        ref, _ = S.mesh_shape_all(M, int(mesh_ids[i]), int(shape_ids[i]), tm[i], ts[i])
        assert np.array_equal(got[i, :cnt[i]], ref), i
        n_eq += 1
    by_type = {t: int(hit[types[shape_ids] == t].sum()) for t in (BOX, SPHERE, CAPSULE, CYLINDER, CONVEX)}
    print(f"\n[mesh-shape] queries={n} colliding={int(hit.sum())} per shape type {by_type}; complete sets "
          f"equal on {n_eq}; bv tests mean {out['bv_tests'].mean():.0f}, leaf tests mean {out['leaf_tests'].mean():.1f}")
    assert all(v > 20 for v in by_type.values())
    assert n_eq > 300


@pytest.mark.gpu
def test_mesh_shape_rejects_penetration(gpu_ctx):
    pool = W.mesh_pool(1, 92, 300, 400)
    gpu_ctx.upload_meshes(pool["verts"], pool["vert_off"], pool["vert_cnt"], pool["tris"], pool["tri_off"],
                          pool["tri_cnt"])
    gpu_ctx.upload_shapes(np.array([SPHERE], np.int32), np.array([[0.1, 0, 0, 0]], np.float32))
    with pytest.raises(RuntimeError, match="penetration"):
        gpu_ctx.mesh_shape_collide([0], [0], W.poses(np.random.default_rng(0), 1, 0, 0),
                                   W.poses(np.random.default_rng(1), 1, 0, 0),
                                   mfcl.request(mfcl.MODE_COLLIDE, max_contacts=4, penetration=1))


@pytest.mark.gpu
def test_mesh_shape_contacts_bit_exact(oracle, gpu_ctx):
    """Penetration enabled: every query's hit flag and contact count equal
    the oracle's; when the colliding triangles fit the cap, every contact
    (b1, b2 = -1, pos, reversed normal, depth) is bit-identical to the
    oracle's (spheres by the closed form, every other shape by GJK + EPA with
    the triangle as a shape)."""
    n = 768
    S, M, types, mesh_ids, shape_ids, pm, ps = _setup(oracle, gpu_ctx, 94, n)
    cap = 128
    req = mfcl.request(mfcl.MODE_COLLIDE, max_contacts=cap, penetration=1)
    out, got = gpu_ctx.mesh_shape_collide_contacts(mesh_ids, shape_ids, pm, ps, req)
    tm, ts = oracle.pose12(pm), oracle.pose12(ps)
    assert np.all(out["hit"] >= 0), "traversal stack overflow"
    n_eq = n_hit = 0
    per_type = {}
    for i in range(n):
        ref, cnt = S.mesh_shape_contacts_all(M, int(mesh_ids[i]), int(shape_ids[i]), tm[i], ts[i])
        assert out["hit"][i] == (1 if cnt else 0), i
        assert out["n_contacts"][i] == min(cnt, cap), i
        if cnt == 0:
            continue
        n_hit += 1
        t = int(types[shape_ids[i]])
        per_type[t] = per_type.get(t, 0) + 1
        if cnt > cap:
            continue
        g = got[i, :cnt]
        assert np.array_equal(g["b1"], ref["b1"]) and np.all(g["b2"] == -1), i
        for f in ("pos", "normal", "depth"):
            assert np.array_equal(g[f], ref[f]), (i, f)
        assert np.all(got[i, cnt:]["b1"] == -1)
        n_eq += 1
    print(f"\n[mesh-shape contacts] queries={n} colliding={n_hit} per shape type {per_type}; "
          f"contacts bit-identical on {n_eq}")
    assert all(per_type.get(t, 0) > 10 for t in (BOX, SPHERE, CAPSULE, CYLINDER, CONVEX))
    assert n_eq > 300
