"""Robot-arm pipeline on the CPU -- TEST INFRASTRUCTURE ONLY.

The checker for mfcl_arm_query_dev: numpy forward kinematics of the serial
chain (joint i about z for even i, y for odd i; link i along its local x,
shape centred at L_i / 2; the base at the origin), the link broadphase boxes
in the same FP32 arithmetic as the device, the oracle's AABB tree
(binary_AABB_tree-inl.h:400-440) and signed distance (gjk_solver-inl.h:
820-874) per candidate pair, and the per-configuration reductions.  Only
tests and the benchmark's cpu_baseline use it.
"""
import numpy as np

import pyoracle


def fk(q: np.ndarray, lens: np.ndarray) -> np.ndarray:
    """Link poses (n, L, 7) as FP32 (t, w, x, y, z); FP64 composition."""
    q = np.asarray(q, np.float64)
    n, L = q.shape
    qw, qx, qy, qz = np.ones(n), np.zeros(n), np.zeros(n), np.zeros(n)
    p = np.zeros((n, 3))
    out = np.empty((n, L, 7), np.float32)
    for i in range(L):
        s, co = np.sin(0.5 * q[:, i]), np.cos(0.5 * q[:, i])
        if i % 2 == 0:  # right-multiply by a rotation about z
            qw, qx, qy, qz = qw * co - qz * s, qx * co + qy * s, qy * co - qx * s, qz * co + qw * s
        else:  # about y
            qw, qx, qy, qz = qw * co - qy * s, qx * co - qz * s, qy * co + qw * s, qz * co + qx * s
        a = np.stack([1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy + qz * qw), 2 * (qx * qz - qy * qw)], 1)
        h = 0.5 * float(np.float32(lens[i]))
        c = p + h * a
        out[:, i, 0:3] = c
        out[:, i, 3] = qw
        out[:, i, 4] = qx
        out[:, i, 5] = qy
        out[:, i, 6] = qz
        p = p + 2 * h * a
    return out


def link_boxes(pose: np.ndarray, radius: np.ndarray, margin: float):
    """Broadphase boxes: FP32 centre -+ (radius + margin), FP32 arithmetic."""
    rr = (radius.astype(np.float32) + np.float32(margin)).astype(np.float32)
    c = pose[..., 0:3]
    return (c - rr[None, :, None]).astype(np.float32), (c + rr[None, :, None]).astype(np.float32)


def shape_set(scene):
    L = scene["n_links"]
    hos = np.full(len(scene["types"]), -1, np.int32)
    hos[:L] = np.arange(L)
    return pyoracle.ShapeSet(scene["types"], scene["params"], scene["verts"], scene["vert_off"][:L],
                             scene["vert_cnt"][:L], scene["tris"], scene["tri_off"][:L], scene["tri_cnt"][:L],
                             hull_of_shape=hos)


def pipeline(scene, q, margin, cap, S, tree, threads=1, pose=None):
    """Per configuration: (any_hit flags, min signed distance, candidate pairs)."""
    L = scene["n_links"]
    n = len(q)
    if pose is None:
        pose = fk(q, scene["link_len"])
    mn, mx = link_boxes(pose, scene["link_radius"], margin)
    counts, ids, _, _ = tree.query(mn.reshape(-1, 3), mx.reshape(-1, 3), cap=cap, threads=threads)
    flags = np.zeros(n, np.int32)
    over = (counts > cap) | (counts < 0)
    np.bitwise_or.at(flags, np.nonzero(over)[0] // L, 2)
    m = np.clip(counts, 0, cap)
    qk = np.repeat(np.arange(n * L), m)
    slot = np.arange(len(qk)) - np.repeat(np.cumsum(m) - m, m)
    box = ids[qk, slot]
    cfg = qk // L
    pa = (qk % L).astype(np.uint32)
    pb = (L + box).astype(np.uint32)
    tfa = pyoracle.pose12(pose.reshape(-1, 7)[qk])
    tfb = pyoracle.pose12(scene["box_pose"][box])
    ref, secs = S.signed_distance(pa, pb, tfa, tfb, threads=threads)
    ok = ref["ok"] == 1
    np.bitwise_or.at(flags, cfg[~ok], 2)
    np.bitwise_or.at(flags, cfg[ok & (ref["distance"] < 0)], 1)
    mind = np.full(n, np.inf, np.float32)
    np.minimum.at(mind, cfg[ok], ref["distance"][ok].astype(np.float32))
    return flags, mind, len(qk)
