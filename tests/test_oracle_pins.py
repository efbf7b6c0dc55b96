"""CPU: pin the oracle's primitive closed forms, supports and EPA depths.

Every check below compares the oracle with an independent computation that
the mathematics fixes, chosen so that a plausible restatement slip (a
dropped term, a wrong sign or index, a swapped operand, lz instead of lz/2)
fails it:

* primitive supports (gjk_shape.hpp:41-180) == the exact support function
  h(d) = max over the solid of d.x (closed forms: box sum |d_i| s_i / 2,
  sphere r|d|, capsule r|d| + |d_z| lz/2, cylinder r |d_xy| + |d_z| lz/2),
  with the returned point inside the solid and no densely sampled surface
  point beyond it; directions include the axes and the xy plane;
* sphere-cylinder and sphere-box closed forms
  (sphere_cylinder-inl.h:48-150, sphere_box-inl.h:48-136) == the exact
  point-to-solid distance: hit, depth = r - signed distance, normal = the
  direction of the nearest boundary point, and the depth is the smallest
  translation along the normal that separates (checked by exact distances
  after moving by 0.999 and 1.001 of it);
* boxBox2 (box_box-inl.h:214-803): normal = one of the 15 separating-axis
  candidates, the overlap along it between the 15-axis minimum and 1.05 of
  it (the reference's edge-axis fudge factor); on face contacts built so the
  incident face lies inside the reference face, the contacts are exactly
  the incident corners, moved half their depth along the normal, with depth
  = their penetration of the reference face, and a truncated request keeps
  the deepest of them (shape_pair_intersect-inl.h:91-113);
* EPA depth (epa.hpp:137-257) on capsule pairs == r1 + r2 - the segment
  distance, and on cylinder / box / capsule mixes within the tolerance of
  the minimum of the Minkowski difference's exact support function over
  unit directions (the distance from an interior point to the boundary of a
  convex body): an overestimate fails the upper bound, an underestimate the
  lower one.
"""
import numpy as np
import pytest

BOX, SPHERE, CAPSULE, CYLINDER = 0, 1, 3, 5
EPA_OK = 1


def _rot(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def _tf(R, t):
    out = np.zeros(12)
    out[:9] = np.asarray(R).reshape(9)
    out[9:] = t
    return out


def _small_rot(rng, angle):
    axis = rng.normal(size=3)
    axis /= np.linalg.norm(axis)
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + np.sin(angle) * K + (1 - np.cos(angle)) * K @ K


# ---------------------------------------------------------------- supports

def _exact_support_value(kind, p, d):
    """h(d) = max over the solid of d.x, from the shape's geometry alone."""
    if kind == BOX:
        return float(np.sum(np.abs(d) * np.asarray(p[:3]) / 2))
    if kind == SPHERE:
        return float(p[0] * np.linalg.norm(d))
    if kind == CAPSULE:
        return float(p[0] * np.linalg.norm(d) + abs(d[2]) * p[1] / 2)
    if kind == CYLINDER:
        return float(p[0] * np.hypot(d[0], d[1]) + abs(d[2]) * p[1] / 2)
    raise ValueError(kind)


def _inside(kind, p, x, tol):
    if kind == BOX:
        return bool(np.all(np.abs(x) <= np.asarray(p[:3]) / 2 + tol))
    if kind == SPHERE:
        return bool(np.linalg.norm(x) <= p[0] + tol)
    if kind == CAPSULE:
        z = np.clip(x[2], -p[1] / 2, p[1] / 2)
        return bool(np.linalg.norm(x - np.array([0, 0, z])) <= p[0] + tol)
    if kind == CYLINDER:
        return bool(np.hypot(x[0], x[1]) <= p[0] + tol and abs(x[2]) <= p[1] / 2 + tol)
    raise ValueError(kind)


def _surface_samples(kind, p, rng, n=40000):
    """Dense points on the boundary of the solid."""
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    if kind == BOX:
        h = np.asarray(p[:3]) / 2
        x = rng.uniform(-1, 1, (n, 3)) * h
        ax = rng.integers(0, 3, n)
        x[np.arange(n), ax] = np.sign(rng.uniform(-1, 1, n)) * h[ax]
        corners = np.array([[sx, sy, sz] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)]) * h
        return np.vstack([x, corners])
    if kind == SPHERE:
        return p[0] * u
    if kind == CAPSULE:
        z = np.where(u[:, 2] >= 0, p[1] / 2, -p[1] / 2)
        ends = p[0] * u + np.stack([0 * z, 0 * z, z], 1)
        t = rng.uniform(0, 2 * np.pi, n)
        barrel = np.stack([p[0] * np.cos(t), p[0] * np.sin(t), rng.uniform(-p[1] / 2, p[1] / 2, n)], 1)
        return np.vstack([ends, barrel])
    if kind == CYLINDER:
        t = rng.uniform(0, 2 * np.pi, n)
        rr = p[0] * np.sqrt(rng.uniform(0, 1, n))
        caps = np.stack([rr * np.cos(t), rr * np.sin(t), np.where(rng.uniform(0, 1, n) < .5, p[1] / 2, -p[1] / 2)], 1)
        rim = np.stack([p[0] * np.cos(t), p[0] * np.sin(t), np.where(t < np.pi, p[1] / 2, -p[1] / 2)], 1)
        barrel = np.stack([p[0] * np.cos(t), p[0] * np.sin(t), rng.uniform(-p[1] / 2, p[1] / 2, n)], 1)
        return np.vstack([caps, rim, barrel])
    raise ValueError(kind)


def _directions(rng):
    d = rng.normal(size=(300, 3))
    axes = np.vstack([np.eye(3), -np.eye(3)])
    near = axes + rng.normal(scale=1e-7, size=axes.shape)
    flat = rng.normal(size=(40, 3))
    flat[:, 2] = 0.0  # in the xy plane: the capsule/cylinder tie on z
    diag = np.array([[1, 1, 0], [1, 0, 1], [0, 1, -1], [1, -1, 1], [-1, -1, -1]], np.float64)
    d = np.vstack([d, axes, near, flat, diag])
    return d / np.linalg.norm(d, axis=1, keepdims=True)


@pytest.mark.parametrize("kind,params", [
    (BOX, (0.3, 0.8, 0.55)), (BOX, (1.0, 0.1, 0.4)), (SPHERE, (0.37, 0, 0)),
    (CAPSULE, (0.2, 0.9, 0)), (CAPSULE, (0.45, 0.1, 0)),
    (CYLINDER, (0.25, 0.7, 0)), (CYLINDER, (0.5, 0.12, 0)),
])
def test_primitive_support_exact(oracle, kind, params):
    rng = np.random.default_rng(100 + kind)
    S = oracle.ShapeSet(np.array([kind], np.int32), np.array([params], np.float64))
    samples = _surface_samples(kind, params, rng)
    scale = max(params)
    for d in _directions(rng):
        s = S.support(0, d)
        h = _exact_support_value(kind, params, d)
        assert abs(d @ s - h) <= 1e-12 * scale, (kind, d, s, h)
        assert _inside(kind, params, s, 1e-12 * scale), (kind, d, s)
        assert (samples @ d).max() <= d @ s + 1e-12 * scale, (kind, d)


def test_cylinder_support_axial_extent(oracle):
    """The cylinder's z extent is lz / 2 (fcl::Cylinder stores the full
    length): axial and oblique directions see +-lz/2, never +-lz."""
    for r, lz in ((0.2, 0.9), (0.6, 0.2)):
        S = oracle.ShapeSet(np.array([CYLINDER, CAPSULE], np.int32), np.array([[r, lz, 0], [r, lz, 0]], np.float64))
        for sgn in (1.0, -1.0):
            s = S.support(0, np.array([0, 0, sgn]))
            assert np.array_equal(s, [0, 0, sgn * lz / 2])
            s = S.support(0, np.array([0.6, 0, 0.8 * sgn]))
            assert np.allclose(s, [r, 0, sgn * lz / 2], atol=1e-15)
            c = S.support(1, np.array([0, 0, sgn]))
            assert np.allclose(c, [0, 0, sgn * (lz / 2 + r)], atol=1e-15)


# ------------------------------------------------------- sphere closed forms

def _cyl_signed(p, r, lz):
    """Signed distance from point p (cylinder frame) to the solid cylinder
    and the outward direction of its nearest boundary point."""
    rho = np.hypot(p[0], p[1])
    dr, dz = rho - r, abs(p[2]) - lz / 2
    if dr <= 0 and dz <= 0:
        if dr > dz:  # barrel nearer
            return dr, np.array([p[0] / rho, p[1] / rho, 0.0]), "barrel"
        return dz, np.array([0.0, 0.0, np.sign(p[2])]), "face"
    q = p.copy()
    if rho > r:
        q[:2] *= r / rho
    q[2] = np.clip(p[2], -lz / 2, lz / 2)
    d = np.linalg.norm(p - q)
    return d, (p - q) / d, "out"


def _box_signed(p, side):
    h = np.asarray(side) / 2
    q = np.clip(p, -h, h)
    if np.any(q != p):
        d = np.linalg.norm(p - q)
        return d, (p - q) / d, "out"
    gaps = h - np.abs(p)
    ax = int(np.argmin(gaps))
    n = np.zeros(3)
    n[ax] = np.sign(p[ax]) if p[ax] != 0 else 1.0
    return -gaps[ax], n, "in"


def _check_sphere_closed_form(oracle, other, params_other, signed_fn, seed):
    rng = np.random.default_rng(seed)
    n = 1500
    rs = rng.uniform(0.05, 0.5, n)
    types = np.array([SPHERE, other] * n, np.int32)
    params = np.zeros((2 * n, 3))
    params[0::2, 0] = rs
    params[1::2] = params_other(rng, n)
    S = oracle.ShapeSet(types, params)
    ta = np.stack([_tf(_rot(rng), rng.uniform(0, 0.8, 3)) for _ in range(n)])
    tb = np.stack([_tf(_rot(rng), rng.uniform(0, 0.8, 3)) for _ in range(n)])
    pa, pb = np.arange(0, 2 * n, 2), np.arange(1, 2 * n, 2)
    out, _ = S.collide(pa, pb, ta, tb, max_contacts=1, penetration=True)
    kinds = {}
    for i in range(n):
        Rb, tbx = tb[i, :9].reshape(3, 3), tb[i, 9:]
        c = Rb.T @ (ta[i, 9:] - tbx)  # sphere centre in the other shape's frame
        sd, nrm_out, kind = signed_fn(c, params[2 * i + 1])
        if abs(sd - rs[i]) < 1e-9 or (kind != "out" and abs(sd) < 1e-9):
            continue
        hit = sd <= rs[i]
        assert bool(out["hit"][i]) == hit, (i, sd, rs[i])
        if not hit:
            continue
        if kind in ("barrel", "face") and abs(_cyl_tie_gap(c, params[2 * i + 1])) < 1e-9:
            continue
        kinds[kind] = kinds.get(kind, 0) + 1
        depth = rs[i] - sd
        assert abs(out["depth"][i] - depth) <= 1e-12, (i, kind, out["depth"][i], depth)
        # normal from the sphere into the other shape: minus the outward normal
        n_world = Rb @ (-nrm_out)
        assert np.allclose(out["normal"][i], n_world, atol=1e-9), (i, kind, out["normal"][i], n_world)
        # the depth is the minimal separating translation along the normal
        for f, sep in ((0.999, False), (1.001, True)):
            moved = c + Rb.T @ (f * depth * n_world)  # sphere fixed, other shape moved by f*depth*n
            sd2, _, _ = signed_fn(c - (moved - c), params[2 * i + 1])
            assert (sd2 > rs[i]) == sep, (i, kind, f, sd2, rs[i])
    return kinds


def _cyl_tie_gap(c, p):
    rho = np.hypot(c[0], c[1])
    return (p[0] - rho) - (p[1] / 2 - abs(c[2]))


def test_sphere_cylinder_closed_form(oracle):
    kinds = _check_sphere_closed_form(
        oracle, CYLINDER,
        lambda rng, n: np.stack([rng.uniform(0.05, 0.5, n), rng.uniform(0.1, 1.0, n), np.zeros(n)], 1),
        lambda c, p: _cyl_signed(c, p[0], p[1]), 21)
    assert kinds.get("out", 0) > 100 and kinds.get("barrel", 0) > 5 and kinds.get("face", 0) > 5, kinds


def test_sphere_box_closed_form(oracle):
    kinds = _check_sphere_closed_form(oracle, BOX, lambda rng, n: rng.uniform(0.1, 1.0, (n, 3)),
                                      lambda c, p: _box_signed(c, p), 22)
    assert kinds.get("out", 0) > 100 and kinds.get("in", 0) > 10, kinds


# ----------------------------------------------------------------- box-box

def _corners(side, R, t):
    h = np.asarray(side) / 2
    s = np.array([[sx, sy, sz] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)], np.float64)
    return (s * h) @ R.T + t


def _sat_axes(R1, R2):
    axes = [R1[:, i] for i in range(3)] + [R2[:, i] for i in range(3)]
    axes += [np.cross(R1[:, i], R2[:, j]) for i in range(3) for j in range(3)]
    out = []
    for a in axes:
        n = np.linalg.norm(a)
        if n > 1e-9:
            out.append(a / n)
    return np.array(out)


def test_box_box_normal_and_depth_vs_sat(oracle):
    """Random box pairs: the contact normal is a separating-axis candidate
    pointing from box 1 to box 2, the overlap along it lies between the
    15-axis minimum and 1.05 times it, and every contact depth lies in
    [0, that overlap]; an edge-edge contact's depth is the overlap."""
    rng = np.random.default_rng(23)
    n = 2500
    sides = rng.uniform(0.1, 1.0, size=(2 * n, 3))
    S = oracle.ShapeSet(np.full(2 * n, BOX, np.int32), sides)
    ta = np.stack([_tf(_rot(rng), rng.uniform(0, 1, 3)) for _ in range(n)])
    tb = np.stack([_tf(_rot(rng), rng.uniform(0, 1, 3)) for _ in range(n)])
    out, _ = S.collide(np.arange(0, 2 * n, 2), np.arange(1, 2 * n, 2), ta, tb, max_contacts=4, penetration=True)
    n_face = n_edge = 0
    for i in np.nonzero(out["hit"])[0]:
        R1, t1, R2, t2 = ta[i, :9].reshape(3, 3), ta[i, 9:], tb[i, :9].reshape(3, 3), tb[i, 9:]
        A, B = _corners(sides[2 * i], R1, t1), _corners(sides[2 * i + 1], R2, t2)
        axes = _sat_axes(R1, R2)
        ov_all = np.minimum((A @ axes.T).max(0) - (B @ axes.T).min(0), (B @ axes.T).max(0) - (A @ axes.T).min(0))
        min_ov = ov_all.min()
        if min_ov < 1e-7:
            continue
        nrm = out["normal"][i]
        assert abs(np.linalg.norm(nrm) - 1) < 1e-12
        cos = np.abs(axes @ nrm).max()
        assert cos > 1 - 1e-12, (i, cos)
        assert nrm @ (t2 - t1) >= -1e-12
        ov = (A @ nrm).max() - (B @ nrm).min()
        assert min_ov - 1e-9 <= ov <= 1.05 * min_ov + 1e-9, (i, ov, min_ov)
        k = out["n"][i]
        assert 1 <= k <= 4
        dep = out["depth"][i]
        assert -1e-12 <= dep <= ov + 1e-9
        face_axis = max(np.abs(R1.T @ nrm).max(), np.abs(R2.T @ nrm).max()) > 1 - 1e-12
        if face_axis:
            n_face += 1
        else:
            assert k == 1
            assert abs(dep - ov) <= 1e-9, (i, dep, ov)
            n_edge += 1
    assert n_face > 200 and n_edge > 20, (n_face, n_edge)


def _face_contact_case(rng, ref_is_first):
    """A small box resting in the +z face of a large one (in the large box's
    frame), slightly tilted, penetrating by a few mm; world poses."""
    big = np.array([1.0, 1.0, 1.0]) * rng.uniform(0.8, 1.2)
    small = rng.uniform(0.1, 0.3, 3)
    Rl = _small_rot(rng, rng.uniform(0.0, 0.08))
    pen = rng.uniform(0.005, 0.04)
    # lowest point of the small box in the big box's frame sits pen below the face
    zc = big[2] / 2 + (np.abs(Rl[2, :]) * small / 2).sum() - pen
    cl = np.array([rng.uniform(-0.2, 0.2), rng.uniform(-0.2, 0.2), zc])
    Rw, tw = _rot(rng), rng.uniform(-1, 1, 3)
    R_small, t_small = Rw @ Rl, Rw @ cl + tw
    if ref_is_first:
        return (big, Rw, tw), (small, R_small, t_small)
    return (small, R_small, t_small), (big, Rw, tw)


@pytest.mark.parametrize("ref_is_first", [True, False])
def test_box_box_face_contacts_closed_form(oracle, ref_is_first):
    """Incident face inside the reference face: boxBox2 returns the incident
    face's four corners, each moved half its depth along the normal, with
    depth = its penetration below the reference face; CollisionRequest(k)
    keeps the k deepest."""
    rng = np.random.default_rng(24 + ref_is_first)
    n_full = 0
    for _ in range(120):
        b1, b2 = _face_contact_case(rng, ref_is_first)
        S = oracle.ShapeSet(np.array([BOX, BOX], np.int32), np.array([b1[0], b2[0]]))
        tf1, tf2 = _tf(b1[1], b1[2]), _tf(b2[1], b2[2])
        big, small = (b1, b2) if ref_is_first else (b2, b1)
        up = big[1][:, 2]  # the big box's +z face normal (world)
        n_expect = up if ref_is_first else -up  # normal points from box 1 to box 2
        face_off = up @ big[2] + big[0][2] / 2
        C = _corners(small[0], small[1], small[2])
        depth_c = face_off - C @ up
        inc = np.argsort(depth_c)[-4:]  # the incident face: the four lowest corners
        inc = inc[depth_c[inc] >= 0]  # boxBox2 keeps the penetrating ones
        exp_pos = C[inc] + np.outer(depth_c[inc] / 2, up)
        exp_dep = depth_c[inc]
        out, _ = S.collide([0], [1], tf1[None], tf2[None], max_contacts=4, penetration=True)
        assert out["hit"][0] == 1 and out["n"][0] == len(inc)
        assert np.allclose(out["normal"][0], n_expect, atol=1e-12)
        # all four contacts through the full-contact oracle entry
        full, _ = S.collide([0], [1], tf1[None], tf2[None], max_contacts=4, penetration=True)
        got = _all_contacts(oracle, S, tf1, tf2, 4)
        assert len(got) == len(inc)
        n_full += len(inc) == 4
        for pos, dep in zip(exp_pos, exp_dep):
            j = int(np.argmin([np.linalg.norm(g[0] - pos) for g in got]))
            assert np.linalg.norm(got[j][0] - pos) < 1e-12, (got[j][0], pos)
            assert abs(got[j][1] - dep) < 1e-12
        for k in (1, 2, 3):
            kept = sorted(g[1] for g in _all_contacts(oracle, S, tf1, tf2, k))
            assert np.allclose(kept, sorted(exp_dep)[-min(k, len(inc)):], atol=1e-12), (k, kept, exp_dep)
        # nothing truncated: contact 0 is the first of the clipping order
        assert full["depth"][0] == got[0][1]
    assert n_full > 40, n_full


def _all_contacts(oracle, S, tf1, tf2, k):
    """(pos, depth) of every contact of CollisionRequest(k) on one pair."""
    rec = oracle.collide_contacts(S, tf1, tf2, k)
    return [(rec["pos"][j], rec["depth"][j]) for j in range(len(rec))]


# --------------------------------------------------------------- EPA depth

def _seg_seg_distance(p1, q1, p2, q2):
    """Closest distance between segments p1q1 and p2q2 (Ericson, Real-Time
    Collision Detection, 5.1.9)."""
    d1, d2, r = q1 - p1, q2 - p2, p1 - p2
    a, e, f = d1 @ d1, d2 @ d2, d2 @ r
    c, b = d1 @ r, d1 @ d2
    den = a * e - b * b
    s = np.clip((b * f - c * e) / den, 0, 1) if den > 1e-15 else 0.0
    t = (b * s + f) / e
    if t < 0:
        t, s = 0.0, np.clip(-c / a, 0, 1)
    elif t > 1:
        t, s = 1.0, np.clip((b - c) / a, 0, 1)
    return float(np.linalg.norm((p1 + d1 * s) - (p2 + d2 * t)))


def test_capsule_capsule_epa_depth(oracle):
    """Two capsules penetrate by r1 + r2 - (segment distance); EPA (status
    OK) is within its tolerance of it."""
    rng = np.random.default_rng(25)
    n = 600
    r = rng.uniform(0.05, 0.4, 2 * n)
    lz = rng.uniform(0.1, 0.8, 2 * n)
    S = oracle.ShapeSet(np.full(2 * n, CAPSULE, np.int32), np.stack([r, lz, 0 * r], 1))
    ta = np.stack([_tf(_rot(rng), rng.uniform(0, 0.5, 3)) for _ in range(n)])
    tb = np.stack([_tf(_rot(rng), rng.uniform(0, 0.5, 3)) for _ in range(n)])
    out, _ = S.collide(np.arange(0, 2 * n, 2), np.arange(1, 2 * n, 2), ta, tb, max_contacts=1, penetration=True)
    checked = 0
    for i in range(n):
        segs = []
        for tf, L in ((ta[i], lz[2 * i]), (tb[i], lz[2 * i + 1])):
            R, t = tf[:9].reshape(3, 3), tf[9:]
            segs += [t + R[:, 2] * L / 2, t - R[:, 2] * L / 2]
        dseg = _seg_seg_distance(*segs)
        depth = r[2 * i] + r[2 * i + 1] - dseg
        if abs(depth) < 1e-7:
            continue
        assert bool(out["hit"][i]) == (depth > 0), (i, depth)
        if depth <= 0 or out["epa_status"][i] != EPA_OK:
            continue
        assert abs(out["depth"][i] - depth) <= 1.5e-6, (i, out["depth"][i], depth)
        checked += 1
    assert checked > 150, checked


def _fib_sphere(n):
    i = np.arange(n) + 0.5
    phi = np.arccos(1 - 2 * i / n)
    th = np.pi * (1 + 5 ** 0.5) * i
    return np.stack([np.cos(th) * np.sin(phi), np.sin(th) * np.sin(phi), np.cos(phi)], 1)


def _support_value(kind, p, R, t, D):
    """Exact support function of a posed primitive for many directions."""
    L = D @ R  # directions in the shape frame (rows)
    if kind == BOX:
        h = (np.abs(L) * np.asarray(p[:3]) / 2).sum(1)
    elif kind == CAPSULE:
        h = p[0] * np.linalg.norm(L, axis=1) + np.abs(L[:, 2]) * p[1] / 2
    elif kind == CYLINDER:
        h = p[0] * np.hypot(L[:, 0], L[:, 1]) + np.abs(L[:, 2]) * p[1] / 2
    else:
        raise ValueError(kind)
    return h + D @ t


def _kink_axes(kind, R):
    """Unit vectors a such that the shape's support function is not smooth
    where n.a = 0 (box: its three axes; capsule, cylinder: the axis)."""
    return [R[:, i] for i in range(3)] if kind == BOX else [R[:, 2]]


def _min_support_difference(k1, p1, tf1, k2, p2, tf2, rng):
    """min over unit n of h_{A-B}(n) = h_A(n) + h_B(-n): the distance from
    the origin to the boundary of A - B when it is inside.  The function is
    smooth except on the great circles orthogonal to the kink axes, so the
    minimum is searched over the sphere, over each such circle (1-D) and at
    their crossings, each refined to ~1e-12; the result is an upper bound
    that equals the minimum up to that refinement."""
    R1, t1, R2, t2 = tf1[:9].reshape(3, 3), tf1[9:], tf2[:9].reshape(3, 3), tf2[9:]

    def f(D):
        return _support_value(k1, p1, R1, t1, D) + _support_value(k2, p2, R2, t2, -D)

    axes = _kink_axes(k1, R1) + _kink_axes(k2, R2)
    vals = []
    # crossings of two kink circles and the axes themselves
    pts = [a for a in axes]
    for i in range(len(axes)):
        for j in range(i + 1, len(axes)):
            c = np.cross(axes[i], axes[j])
            if np.linalg.norm(c) > 1e-9:
                pts.append(c / np.linalg.norm(c))
    P = np.array(pts)
    vals.append(f(np.vstack([P, -P])).min())
    # each kink circle: dense angles, then golden-section refinement
    for a in axes:
        u = np.cross(a, [1.0, 0, 0] if abs(a[0]) < 0.9 else [0, 1.0, 0])
        u /= np.linalg.norm(u)
        w = np.cross(a, u)
        g = lambda th: f(np.outer(np.cos(th), u) + np.outer(np.sin(th), w))
        th = np.linspace(0, 2 * np.pi, 20001)
        v = g(th)
        k = int(np.argmin(v))
        lo, hi = th[k] - 2 * np.pi / 20000, th[k] + 2 * np.pi / 20000
        gr = (5 ** 0.5 - 1) / 2
        for _ in range(80):
            m1, m2 = hi - gr * (hi - lo), lo + gr * (hi - lo)
            if g(np.array([m1]))[0] < g(np.array([m2]))[0]:
                hi = m2
            else:
                lo = m1
        vals.append(min(v[k], g(np.array([(lo + hi) / 2]))[0]))
    # the smooth part: Fibonacci sampling, then shrinking random search
    D = _fib_sphere(20000)
    v = f(D)
    for b in D[np.argsort(v)[:6]]:
        cur, cv = b, f(b[None])[0]
        step = 0.02
        while step > 1e-12:
            cand = cur + rng.normal(scale=step, size=(64, 3))
            cand /= np.linalg.norm(cand, axis=1, keepdims=True)
            cvals = f(cand)
            j = int(np.argmin(cvals))
            if cvals[j] < cv:
                cur, cv = cand[j], cvals[j]
            else:
                step *= 0.6
        vals.append(cv)
    return float(min(vals))


@pytest.mark.parametrize("k1,k2", [(CYLINDER, CYLINDER), (CAPSULE, CYLINDER), (BOX, CYLINDER), (CAPSULE, BOX)])
def test_epa_depth_vs_support_function_minimum(oracle, k1, k2):
    rng = np.random.default_rng(26 + k1 + 7 * k2)

    def params(k):
        if k == BOX:
            return rng.uniform(0.2, 0.8, 3)
        return np.array([rng.uniform(0.1, 0.4), rng.uniform(0.2, 0.8), 0.0])

    checked = 0
    for _ in range(60):
        p1, p2 = params(k1), params(k2)
        S = oracle.ShapeSet(np.array([k1, k2], np.int32), np.array([p1, p2]))
        tf1 = _tf(_rot(rng), rng.uniform(0, 0.3, 3))
        tf2 = _tf(_rot(rng), rng.uniform(0, 0.3, 3))
        out, _ = S.collide([0], [1], tf1[None], tf2[None], max_contacts=1, penetration=True)
        if not out["hit"][0] or out["epa_status"][0] != EPA_OK or out["method"][0] != 1:
            continue
        ref = _min_support_difference(k1, p1, tf1, k2, p2, tf2, rng)
        if ref < 1e-4:
            continue
        got = out["depth"][0]
        # the EPA polytope lies inside A - B: its depth cannot exceed the true one
        assert got <= ref + 1e-9, (got, ref)
        # and on termination it is within the tolerance (1e-6) of it
        assert got >= ref - 2e-6, (got, ref)
        checked += 1
    assert checked >= 20, checked


# --------------------------------------------------- triangle-pair contacts

def test_triangle_contacts_properties(oracle):
    """intersect_Triangle with contacts (intersect-inl.h:796-842) on random
    intersecting triangle pairs: the normal is the unit normal of one
    triangle's plane (+ the first's, or - the second's); each contact point
    is a vertex of the other triangle; the depth is that vertex's depth
    below the plane (t - n.v) and no vertex of the other triangle lies
    deeper; at most two points; the pair order reverses the normal."""
    rng = np.random.default_rng(27)
    I3 = np.eye(3).reshape(9)
    n_hit = 0
    for _ in range(3000):
        P = rng.uniform(-1, 1, (3, 3))
        Q = rng.uniform(-1, 1, (3, 3))
        hit, pts, depth, nrm = oracle.tri_tri_contacts(P.reshape(9), Q.reshape(9), I3, np.zeros(3))
        if not hit:
            continue
        n_hit += 1
        nP = np.cross(P[1] - P[0], P[2] - P[0])
        nP /= np.linalg.norm(nP)
        nQ = np.cross(Q[1] - Q[0], Q[2] - Q[0])
        nQ /= np.linalg.norm(nQ)
        assert abs(np.linalg.norm(nrm) - 1) < 1e-12
        use_p = np.allclose(nrm, nP, atol=1e-12)
        assert use_p or np.allclose(nrm, -nQ, atol=1e-12)
        other = Q if use_p else P
        plane_n, plane_pt = (nP, P[0]) if use_p else (nQ, Q[0])
        d_other = (plane_pt @ plane_n) - other @ plane_n  # depth below the plane
        assert len(pts) <= 2
        assert abs(depth - d_other.max()) < 1e-12
        for p in pts:
            j = int(np.argmin(np.linalg.norm(other - p, axis=1)))
            assert np.linalg.norm(other[j] - p) < 1e-15
            assert abs(d_other[j] - depth) <= 1e-6 + 1e-12
    assert n_hit > 300


def test_triangle_contacts_known_answer(oracle):
    """A small triangle piercing a large one in the z = 0 plane with one
    vertex 0.1 below it: the contact is that vertex, depth 0.1, normal +z
    (the large triangle's side: its own vertices lie far on both sides of
    the small triangle's plane, so that depth is the larger)."""
    P = np.array([[-2.0, -2.0, 0.0], [3.0, -1.0, 0.0], [-1.0, 3.0, 0.0]])
    Q = np.array([[0.1, 0.2, -0.1], [0.4, 0.1, 0.6], [0.0, 0.5, 0.5]])
    hit, pts, depth, nrm = oracle.tri_tri_contacts(P.reshape(9), Q.reshape(9), np.eye(3).reshape(9), np.zeros(3))
    assert hit
    assert np.allclose(nrm, [0, 0, 1], atol=1e-15)
    assert abs(depth - 0.1) < 1e-15
    assert len(pts) == 1 and np.allclose(pts[0], Q[0], atol=1e-15)
    # the same pair with Q given in a rotated frame (R, T) gives the same answer
    rng = np.random.default_rng(28)
    R = _rot(rng)
    T = rng.uniform(-1, 1, 3)
    Ql = (Q - T) @ R  # R^T (q - T)
    hit2, pts2, depth2, nrm2 = oracle.tri_tri_contacts(P.reshape(9), Ql.reshape(9), R.reshape(9), T)
    assert hit2 and abs(depth2 - 0.1) < 1e-12 and np.allclose(nrm2, [0, 0, 1], atol=1e-12)
    assert np.allclose(pts2[0], Q[0], atol=1e-12)
