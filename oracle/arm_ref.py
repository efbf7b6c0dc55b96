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


# ---- heightmap ---------------------------------------------------------------
def hm_candidates(scene, pose, S_unused=None):
    """Heightmap candidates of every link, the flat algorithm of
    HeightMapShapeIntersect (heightmap_solver_traverse-inl.h:17-98) over the
    bottom layer: the hull's AABB in the map frame (utility-inl.h:160-173:
    R v + t per vertex, FP64), nothing when it lies below the map's base or
    above its highest pixel, else the pixels of aabbToPixelROI
    (flat_heightmap-inl.h:58-87) with non-zero height at or above the AABB's
    bottom (traverse-inl.h:78-85).  Returns (query k, pixel x, pixel y,
    the hull AABB minimum and maximum (n, 3), its lowest vertex (n, 3) and the
    interior point (n, 3) in the map frame)."""
    L = scene["n_links"]
    H = scene["hm_heights"]
    fy, fx = H.shape
    hx, hy = fx // 2, fy // 2
    res = float(np.float32(scene["hm_res"]))
    P = pose.reshape(-1, 7)
    nq = len(P)
    tl = pyoracle.pose12(P)
    th = pyoracle.pose12(scene["hm_pose7"][None])[0]
    Rh, t_h = th[:9].reshape(3, 3), th[9:]
    Rl, t_l = tl[:, :9].reshape(-1, 3, 3), tl[:, 9:]
    # relTf: M = Rh^T Rl, T = Rh^T t_l + (-(Rh^T t_h)), in the device's order
    M = np.empty_like(Rl)
    for i in range(3):
        for j in range(3):
            M[:, i, j] = Rh[0, i] * Rl[:, 0, j] + Rh[1, i] * Rl[:, 1, j] + Rh[2, i] * Rl[:, 2, j]
    ti = -np.array([Rh[0, i] * t_h[0] + Rh[1, i] * t_h[1] + Rh[2, i] * t_h[2] for i in range(3)])
    T = np.stack([(Rh[0, i] * t_l[:, 0] + Rh[1, i] * t_l[:, 1] + Rh[2, i] * t_l[:, 2]) + ti[i] for i in range(3)], 1)
    link = np.arange(nq) % L
    mn = np.full((nq, 3), np.inf)
    mx = np.full((nq, 3), -np.inf)
    low = np.zeros((nq, 3))
    interior = np.zeros((nq, 3))
    for li in range(L):
        sel = np.nonzero(link == li)[0]
        o, c = int(scene["vert_off"][li]), int(scene["vert_cnt"][li])
        V = scene["verts"][o:o + c].astype(np.float64)
        Ms, Ts = M[sel], T[sel]
        P3 = np.stack([(Ms[:, i, 0:1] * V[None, :, 0] + Ms[:, i, 1:2] * V[None, :, 1]) + Ms[:, i, 2:3] * V[None, :, 2]
                       + Ts[:, i:i + 1] for i in range(3)], 2)  # (n, c, 3)
        mn[sel] = P3.min(1)
        mx[sel] = P3.max(1)
        low[sel] = P3[np.arange(len(sel)), P3[:, :, 2].argmin(1)]
        cen = V.mean(0)  # only used by the tests' shortcut checks (vertex mean)
        interior[sel] = np.stack([(Ms[:, i, 0] * cen[0] + Ms[:, i, 1] * cen[1]) + Ms[:, i, 2] * cen[2] + Ts[:, i]
                                  for i in range(3)], 1)
    upper = float(H.max()) * 0.001
    ks, xs, ys = [], [], []

    def proj(p, half, full):
        v = np.floor(p / res) + half
        out = np.clip(v, 0, full - 1)
        return out.astype(np.int64), out != v

    x0, px0 = proj(mn[:, 0], hx, fx)
    y0, py0 = proj(mn[:, 1], hy, fy)
    x1, px1 = proj(mx[:, 0], hx, fx)
    y1, py1 = proj(mx[:, 1], hy, fy)
    live = ~(mx[:, 2] < 0) & ~(mn[:, 2] > upper)
    live &= ~((px0 | py0) & (px1 | py1) & ((x0 == x1) | (y0 == y1)))
    Hm = H.astype(np.float64) * 0.001
    for k in np.nonzero(live)[0]:
        sub = Hm[y0[k]:y1[k] + 1, x0[k]:x1[k] + 1]
        yy, xx = np.nonzero((sub > 0) & ~(mn[k, 2] > sub))
        ks.append(np.full(len(yy), k))
        xs.append(xx + x0[k])
        ys.append(yy + y0[k])
    cat = (lambda a, dt: np.concatenate(a).astype(dt) if a else np.zeros(0, dt))
    return cat(ks, np.int64), cat(xs, np.int64), cat(ys, np.int64), mn, mx, low, interior


def hm_pairs(scene, k, x, y, pose):
    """(pa, pb, pose7 a, pose7 b) of heightmap candidates: the pixel's box
    first, as boxToShapeProcessLeafPair orders them (heightmap_solver_leaf-
    inl.h:6-28; constructBox: centred on the pixel, half its height up, the
    map's rotation; world centre in FP64 rounded to FP32), then the link."""
    L = scene["n_links"]
    H = scene["hm_heights"]
    fy, fx = H.shape
    res = float(np.float32(scene["hm_res"]))
    th = pyoracle.pose12(scene["hm_pose7"][None])[0]
    Rh, t_h = th[:9].reshape(3, 3), th[9:]
    c = np.stack([(x - fx // 2 + 0.5) * res, (y - fy // 2 + 0.5) * res, 0.5 * (H[y, x].astype(np.float64) * 0.001)], 1)
    w = np.stack([((Rh[i, 0] * c[:, 0] + Rh[i, 1] * c[:, 1]) + Rh[i, 2] * c[:, 2]) + t_h[i] for i in range(3)], 1)
    qb = np.empty((len(k), 7), np.float32)
    qb[:, :3] = w.astype(np.float32)
    qb[:, 3:] = scene["hm_pose7"][3:]
    link = (k % L).astype(np.uint32)
    box = (scene["hm_box_shape0"] + y * fx + x).astype(np.uint32)
    return box, link, qb, pose.reshape(-1, 7)[k]


def hm_collide(scene, S, pose, threads=1):
    """Per query k: does the link collide with the map (some candidate
    pixel box with ShapeIntersect(Box, Convex) true, boolean request:
    MPR, then GJK when MPR fails)?  Also the candidate count per query."""
    k, x, y, *_ = hm_candidates(scene, pose)
    nq = pose.reshape(-1, 7).shape[0]
    hit = np.zeros(nq, bool)
    cnt = np.bincount(k, minlength=nq)
    if len(k):
        pa, pb, qa, qb = hm_pairs(scene, k, x, y, pose)
        ref, _ = S.collide(pa, pb, pyoracle.pose12(qa), pyoracle.pose12(qb), max_contacts=1, penetration=False,
                           threads=threads)
        np.logical_or.at(hit, k, ref["hit"] == 1)
    return hit, cnt, (k, x, y)
