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
    derr = np.abs(min_d - mind)[inside]
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
