"""GPU: the robot-arm configuration query (mfcl_arm_query_dev) against the
CPU pipeline of oracle/arm_ref.py (numpy kinematics, the oracle's AABB tree
and signed distance).

The device composes the chain with its own sin/cos, so an FP32 link pose can
differ from numpy's in the last bit; the bars allow for that and nothing
more: the collision flag is identical outside the touching band
(|min distance| <= 1e-5), the minimum distance agrees within 1e-5 wherever
it is clearly inside the margin, the failure flag (broadphase cap, failed
pair) is identical, and the candidate pair counts agree within 0.1 %.
"""
import numpy as np
import pytest

import arm_ref as A
import mfcl
import pyoracle
from mfcl import workloads as W

MARGIN = 0.05
CAP = 64


@pytest.mark.gpu
def test_arm_query_matches_cpu_pipeline(gpu_ctx):
    scene = W.arm_scene(2048, seed=6)
    L, nb = scene["n_links"], scene["n_boxes"]
    gpu_ctx.upload_shapes(scene["types"], scene["params"], scene["vert_off"], scene["vert_cnt"], scene["verts"],
                          faces=mfcl.faces_from_triangles(scene["tris"], scene["tri_off"], scene["tri_cnt"]))
    mn, mx = W.box_aabbs(scene["box_pose"], scene["box_half"])
    gpu_ctx.aabbtree_build(mn, mx, np.arange(nb))
    arm = mfcl.make_arm(np.arange(L), scene["link_len"], scene["link_radius"], L, nb, broadphase_cap=CAP)
    n = len(scene["q"])
    dq, dbox = gpu_ctx.to_device(scene["q"]), gpu_ctx.to_device(scene["box_pose"])
    dany, dmin = gpu_ctx.alloc(4 * n), gpu_ctx.alloc(4 * n)
    gpu_ctx.reserve_arm(arm, n)
    npairs = gpu_ctx.arm_query_dev(arm, dq, n, MARGIN, dbox, dany, dmin)
    any_hit, min_d = np.zeros(n, np.int32), np.zeros(n, np.float32)
    dany.download(any_hit)
    dmin.download(min_d)

    S = A.shape_set(scene)
    tree = pyoracle.AABBTree(mn, mx, np.arange(nb))
    flags, mind, nref = A.pipeline(scene, scene["q"], MARGIN, CAP, S, tree, threads=8)
    band = np.abs(mind) <= 1e-5
    hit_mism = ((any_hit & 1) != (flags & 1)) & ~band
    inside = (mind < 0.5 * MARGIN) & (min_d < 0.5 * MARGIN)
    derr = np.abs(min_d[inside] - mind[inside])
    print(f"\n[arm] configs={n} pairs gpu={npairs} cpu={nref} colliding={int((flags & 1).sum())} "
          f"failed={int((flags & 2).sum())} hit mismatches={int(hit_mism.sum())} "
          f"max |dist err| {derr.max() if derr.size else 0:.2e}")
    assert hit_mism.sum() == 0
    assert np.array_equal(any_hit & 2, flags & 2)
    assert derr.size == 0 or derr.max() <= 1e-5
    assert abs(npairs - nref) <= 1e-3 * max(nref, 1)
    assert np.array_equal(np.isinf(min_d), np.isinf(mind))


@pytest.mark.gpu
def test_arm_query_rejects_bad_arguments(gpu_ctx):
    scene = W.arm_scene(8, seed=6)
    L, nb = scene["n_links"], scene["n_boxes"]
    gpu_ctx.upload_shapes(scene["types"], scene["params"], scene["vert_off"], scene["vert_cnt"], scene["verts"])
    mn, mx = W.box_aabbs(scene["box_pose"], scene["box_half"])
    gpu_ctx.aabbtree_build(mn, mx, np.arange(nb))
    dq, dbox = gpu_ctx.to_device(scene["q"]), gpu_ctx.to_device(scene["box_pose"])
    dany, dmin = gpu_ctx.alloc(32), gpu_ctx.alloc(32)
    bad = mfcl.make_arm(np.arange(L), scene["link_len"], scene["link_radius"], L, nb + 1)  # one box too many
    with pytest.raises(RuntimeError, match="box shape ids"):
        gpu_ctx.arm_query_dev(bad, dq, 8, MARGIN, dbox, dany, dmin)
    bad = mfcl.make_arm(np.arange(L), scene["link_len"], scene["link_radius"], L, nb, broadphase_cap=0)
    with pytest.raises(RuntimeError, match="broadphase_cap"):
        gpu_ctx.arm_query_dev(bad, dq, 8, MARGIN, dbox, dany, dmin)


@pytest.mark.gpu
def test_arm_heightmap_matches_cpu_pipeline(gpu_ctx):
    """The arm query with the 256 x 256 fBm heightmap: any-collision per
    configuration equals the CPU pipeline (box pairs by signed distance, map
    candidates by ShapeIntersect(Box, Convex) per pixel box) wherever the
    configuration is not touching (a mismatch is allowed only when some
    candidate of that configuration has |signed distance| <= 1e-5); the
    minimum distance is that of the boxes alone."""
    scene = W.arm_scene(2048, seed=6, heightmap=True)
    L, nb = scene["n_links"], scene["n_boxes"]
    gpu_ctx.upload_shapes(scene["types"], scene["params"], scene["vert_off"], scene["vert_cnt"], scene["verts"],
                          faces=mfcl.faces_from_triangles(scene["tris"], scene["tri_off"], scene["tri_cnt"]))
    gpu_ctx.upload_heightmap(scene["hm_heights"], scene["hm_res"], scene["hm_res"])
    mn, mx = W.box_aabbs(scene["box_pose"], scene["box_half"])
    gpu_ctx.aabbtree_build(mn, mx, np.arange(nb))
    hm = dict(box_shape0=scene["hm_box_shape0"], pose7=scene["hm_pose7"], cap=4096)
    arm = mfcl.make_arm(np.arange(L), scene["link_len"], scene["link_radius"], L, nb, broadphase_cap=CAP,
                        heightmap=hm)
    n = len(scene["q"])
    dq, dbox = gpu_ctx.to_device(scene["q"]), gpu_ctx.to_device(scene["box_pose"])
    dany, dmin = gpu_ctx.alloc(4 * n), gpu_ctx.alloc(4 * n)
    gpu_ctx.reserve_arm(arm, n)
    npairs = gpu_ctx.arm_query_dev(arm, dq, n, MARGIN, dbox, dany, dmin)
    n_hm = gpu_ctx.arm_last_heightmap_pairs()
    any_hit, min_d = np.zeros(n, np.int32), np.zeros(n, np.float32)
    dany.download(any_hit)
    dmin.download(min_d)

    S = A.shape_set(scene)
    tree = pyoracle.AABBTree(mn, mx, np.arange(nb))
    pose = A.fk(scene["q"], scene["link_len"])
    flags, mind, nref = A.pipeline(scene, scene["q"], MARGIN, CAP, S, tree, threads=8, pose=pose)
    hm_hit, hm_cnt, (k, x, y) = A.hm_collide(scene, S, pose, threads=8)
    want = (flags & 1).astype(bool) | hm_hit.reshape(n, L).any(1)
    got = (any_hit & 1).astype(bool)
    mism = np.nonzero(got != want)[0]
    touching = []
    for c in mism:
        sel = (k // L) == c
        if not sel.any():
            touching.append(False)
            continue
        pa, pb, qa, qb = A.hm_pairs(scene, k[sel], x[sel], y[sel], pose)
        ref, _ = S.signed_distance(pa, pb, pyoracle.pose12(qa), pyoracle.pose12(qb), threads=8)
        touching.append(bool((np.abs(ref["distance"]) <= 1e-5).any()) or abs(mind[c]) <= 1e-5)
    print(f"\n[arm+heightmap] configs={n} colliding cpu={int(want.sum())} gpu={int(got.sum())} "
          f"map-only={int((want & ~(flags & 1).astype(bool)).sum())} box pairs={npairs} map pairs gpu={n_hm} "
          f"map candidates cpu={len(k)} mismatches={len(mism)} (touching {sum(touching)}) {mism[:10].tolist()}")
    assert all(touching)
    assert int((any_hit & 2).sum()) == 0 and int((flags & 2).sum()) == 0
    assert hm_cnt.max() <= 4096
    inside = (mind < 0.5 * MARGIN) & (min_d < 0.5 * MARGIN)
    derr = np.abs(min_d[inside] - mind[inside])
    assert derr.size == 0 or derr.max() <= 1e-5
    # the shortcuts decide most colliding links before any narrowphase
    assert n_hm < len(k)


@pytest.mark.gpu
def test_heightmap_upload_rejects_bad_shapes(gpu_ctx):
    with pytest.raises(RuntimeError, match="powers of two"):
        gpu_ctx.upload_heightmap(np.zeros((6, 6), np.uint16), 0.01, 0.01)
    with pytest.raises(RuntimeError, match="powers of two"):
        gpu_ctx.upload_heightmap(np.zeros((8, 8), np.uint16), 0.0, 0.01)


def _scene_with_map(n, heights, res, pose7, seed=6):
    """arm_scene with a caller-chosen heightmap (any power-of-two halves,
    resolution and frame) appended as pixel box shapes."""
    sc = W.arm_scene(n, seed=seed)
    npix = heights.size
    sc["hm_heights"], sc["hm_res"], sc["hm_pose7"] = heights, res, np.asarray(pose7, np.float32)
    sc["hm_box_shape0"] = len(sc["types"])
    sc["types"] = np.concatenate([sc["types"], np.full(npix, W.BOX, np.int32)])
    sc["params"] = np.concatenate([sc["params"], W.heightmap_box_params(heights, res, res)])
    sc["vert_off"] = np.concatenate([sc["vert_off"], np.zeros(npix, np.uint64)])
    sc["vert_cnt"] = np.concatenate([sc["vert_cnt"], np.zeros(npix, np.uint32)])
    sc["tri_off"] = np.concatenate([sc["tri_off"], np.zeros(npix, np.int64)])
    sc["tri_cnt"] = np.concatenate([sc["tri_cnt"], np.zeros(npix, np.int64)])
    return sc


@pytest.mark.gpu
def test_arm_heightmap_non_square_tilted(gpu_ctx):
    """A 128 x 512 map of 2 cm pixels (halves 64 and 256: a 1 x 4 top layer,
    8 layers) with holes (zero-height pixels), its frame tilted 0.2 rad about
    x: collision flags equal the CPU pipeline's outside the touching band."""
    rng = np.random.default_rng(11)
    h = W.fbm_heightmap(rng, half=256)[:, :128].copy()  # rows = y (512), columns = x (128)
    h[rng.random(h.shape) < 0.05] = 0
    ang = 0.2
    pose7 = [0.2, -0.1, -0.5, np.cos(ang / 2), np.sin(ang / 2), 0.0, 0.0]
    scene = _scene_with_map(1024, h, 0.02, pose7)
    L, nb = scene["n_links"], scene["n_boxes"]
    gpu_ctx.upload_shapes(scene["types"], scene["params"], scene["vert_off"], scene["vert_cnt"], scene["verts"],
                          faces=mfcl.faces_from_triangles(scene["tris"], scene["tri_off"], scene["tri_cnt"]))
    gpu_ctx.upload_heightmap(h, 0.02, 0.02)
    mn, mx = W.box_aabbs(scene["box_pose"], scene["box_half"])
    gpu_ctx.aabbtree_build(mn, mx, np.arange(nb))
    arm = mfcl.make_arm(np.arange(L), scene["link_len"], scene["link_radius"], L, nb, broadphase_cap=CAP,
                        heightmap=dict(box_shape0=scene["hm_box_shape0"], pose7=pose7, cap=8192))
    n = len(scene["q"])
    dq, dbox = gpu_ctx.to_device(scene["q"]), gpu_ctx.to_device(scene["box_pose"])
    dany, dmin = gpu_ctx.alloc(4 * n), gpu_ctx.alloc(4 * n)
    gpu_ctx.reserve_arm(arm, n)
    gpu_ctx.arm_query_dev(arm, dq, n, MARGIN, dbox, dany, dmin)
    any_hit = np.zeros(n, np.int32)
    dany.download(any_hit)
    S = A.shape_set(scene)
    tree = pyoracle.AABBTree(mn, mx, np.arange(nb))
    pose = A.fk(scene["q"], scene["link_len"])
    flags, mind, _ = A.pipeline(scene, scene["q"], MARGIN, CAP, S, tree, threads=8, pose=pose)
    hm_hit, hm_cnt, (k, x, y) = A.hm_collide(scene, S, pose, threads=8)
    want = (flags & 1).astype(bool) | hm_hit.reshape(n, L).any(1)
    got = (any_hit & 1).astype(bool)
    mism = np.nonzero(got != want)[0]
    touching = []
    for c in mism:
        sel = (k // L) == c
        ok = abs(mind[c]) <= 1e-5
        if sel.any():
            pa, pb, qa, qb = A.hm_pairs(scene, k[sel], x[sel], y[sel], pose)
            ref, _ = S.signed_distance(pa, pb, pyoracle.pose12(qa), pyoracle.pose12(qb), threads=8)
            ok = ok or bool((np.abs(ref["distance"]) <= 1e-5).any())
        touching.append(ok)
    map_only = want & ~(flags & 1).astype(bool)
    print(f"\n[arm+tilted map] configs={n} colliding={int(want.sum())} map-only={int(map_only.sum())} "
          f"map candidates cpu={len(k)} mismatches={len(mism)} {mism[:10].tolist()}")
    assert map_only.sum() > 0
    assert all(touching)
    assert int((any_hit & 2).sum()) == 0 and hm_cnt.max() <= 8192


@pytest.mark.gpu
def test_arm_heightmap_rejects_bad_arguments():
    ctx = mfcl.Context(0)  # a context without a heightmap
    try:
        _bad_heightmap_arguments(ctx)
    finally:
        ctx.close()


def _bad_heightmap_arguments(gpu_ctx):
    scene = W.arm_scene(8, seed=6, heightmap=True)
    L, nb = scene["n_links"], scene["n_boxes"]
    gpu_ctx.upload_shapes(scene["types"], scene["params"], scene["vert_off"], scene["vert_cnt"], scene["verts"])
    mn, mx = W.box_aabbs(scene["box_pose"], scene["box_half"])
    gpu_ctx.aabbtree_build(mn, mx, np.arange(nb))
    dq, dbox = gpu_ctx.to_device(scene["q"]), gpu_ctx.to_device(scene["box_pose"])
    dany, dmin = gpu_ctx.alloc(32), gpu_ctx.alloc(32)
    hm = dict(box_shape0=scene["hm_box_shape0"], pose7=scene["hm_pose7"], cap=64)
    arm = mfcl.make_arm(np.arange(L), scene["link_len"], scene["link_radius"], L, nb, heightmap=hm)
    with pytest.raises(RuntimeError, match="no heightmap"):
        gpu_ctx.arm_query_dev(arm, dq, 8, MARGIN, dbox, dany, dmin)
    gpu_ctx.upload_heightmap(scene["hm_heights"], scene["hm_res"], scene["hm_res"])
    bad = mfcl.make_arm(np.arange(L), scene["link_len"], scene["link_radius"], L, nb, heightmap=dict(hm, cap=0))
    with pytest.raises(RuntimeError, match="hm_cap"):
        gpu_ctx.arm_query_dev(bad, dq, 8, MARGIN, dbox, dany, dmin)
    bad = mfcl.make_arm(np.arange(L), scene["link_len"], scene["link_radius"], L, nb, heightmap=dict(hm, box_shape0=0))
    with pytest.raises(RuntimeError, match="must be boxes"):
        gpu_ctx.arm_query_dev(bad, dq, 8, MARGIN, dbox, dany, dmin)
    bad = mfcl.make_arm(np.arange(L), scene["link_len"], scene["link_radius"], L, nb,
                        heightmap=dict(hm, pose7=[0, 0, 0, 0, 0, 0, 0]))
    with pytest.raises(RuntimeError, match="quaternion"):
        gpu_ctx.arm_query_dev(bad, dq, 8, MARGIN, dbox, dany, dmin)
